// mech_frequency.cpp -- natural frequencies of a shell model read from a BDF
// with the GPU classes: the cylinder-buckling driver's set-up (one isotropic
// shell per component, aluminium) followed by TACSFrequencyAnalysis with the
// Lanczos constructor (src/TACSBuckling.cpp:614-702), as a TACS user would
// write it.  The reference ships no frequency example; this one exercises
// TACSFrequencyAnalysis::solve / extractEigenvalue / evalEigenDVSens /
// evalEigenXptSens.
//
// Usage: mech_frequency [file.bdf] [thickness] [sigma] [num_eigvals] [max_lanczos]
// Prints lambda[i] = omega_i^2, the frequency in Hz and the Ritz residual,
// the sum over elements of d(lambda_0)/dt and X . d(lambda_0)/dX.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "TACSShellGPU.h"

using namespace shellgpu;

int main(int argc, char *argv[]) {
  const char *bdf = argc > 1 ? argv[1] : "mech-cylinder.bdf";
  const TacsScalar thick = argc > 2 ? atof(argv[2]) : 0.010;
  const TacsScalar sigma = argc > 3 ? atof(argv[3]) : 0.0;
  const int num_eigvals = argc > 4 ? atoi(argv[4]) : 10;
  const int max_lanczos = argc > 5 ? atoi(argv[5]) : 100;

  TACSMeshLoader *mesh = new TACSMeshLoader();
  mesh->incref();
  if (mesh->scanBDFFile(bdf)) return 1;
  const int num_components = mesh->getNumComponents();
  TACSShellTransform *transform = new TACSShellNaturalTransform();
  transform->incref();
  TACSMaterialProperties *mat = new TACSMaterialProperties(2718.0, 0.0, 72.0e9, 0.33, 1e11, 10e-6, 0.0);
  mat->incref();
  for (int icomp = 0; icomp < num_components; icomp++) {
    TACSIsoShellConstitutive *con = new TACSIsoShellConstitutive(mat, thick);
    TACSElement *shell = TacsCreateShellByName(mesh->getElementDescript(icomp), transform, con);
    if (!shell) {
      fprintf(stderr, "no shell element for %s\n", mesh->getElementDescript(icomp));
      return 1;
    }
    mesh->setElement(icomp, shell);
  }
  TACSAssembler *assembler = mesh->createTACS(6);
  if (!assembler) return 1;
  assembler->incref();
  mesh->decref();
  printf("nodes %d elements %d\n", assembler->getNumNodes(), assembler->getNumElements());

  TACSSchurMat *kmat = assembler->createSchurMat();
  TACSSchurMat *mmat = assembler->createSchurMat();
  TACSSchurPc *pc = new TACSSchurPc(kmat, 1000000, 10.0, 1);
  pc->incref();
  GMRES *solver = new GMRES(kmat, pc, 10, 15, 0);
  solver->incref();
  solver->setTolerances(1e-12, 1e-12);

  TACSFrequencyAnalysis *freq =
      new TACSFrequencyAnalysis(assembler, sigma, mmat, kmat, solver, max_lanczos, num_eigvals, 1e-12);
  freq->incref();
  KSMPrint *ksm_print = new KSMPrintStdout("FrequencyAnalysis", 0, 10);
  ksm_print->incref();
  freq->solve(ksm_print);

  for (int i = 0; i < num_eigvals; i++) {
    TacsScalar error;
    TacsScalar eig = freq->extractEigenvalue(i, &error);
    printf("lambda[%d] = %.16e freq %.6f Hz error %.3e\n", i, eig, std::sqrt(std::fabs(eig)) / (2.0 * M_PI), error);
  }
  freq->checkEigenvector(0);

  TACSBVec *dfdx = assembler->createDesignVec();
  dfdx->incref();
  freq->evalEigenDVSens(0, dfdx);
  std::vector<TacsScalar> g;
  dfdx->getValues(g);
  double sum = 0.0;
  for (double v : g) sum += v;
  printf("sum d(lambda_0)/dt = %.16e\n", sum);

  // shape sensitivity: X . d(lambda_0)/dX is the derivative under a uniform
  // scaling of the mesh about the origin
  TACSBVec *X = assembler->createNodeVec(), *dfdX = assembler->createNodeVec();
  X->incref();
  dfdX->incref();
  assembler->getNodes(X);
  freq->evalEigenXptSens(0, dfdX);
  printf("X . d(lambda_0)/dX = %.16e\n", X->dot(dfdX));
  X->decref();
  dfdX->decref();

  dfdx->decref();
  freq->decref();
  ksm_print->decref();
  solver->decref();
  pc->decref();
  mat->decref();
  transform->decref();
  assembler->decref();
  return 0;
}
