// mech_buckling.cpp -- the reference's cylinder buckling driver
// (examples/cylinder-buckling/mechBuckling.cpp) on the GPU classes: scan a
// BDF, one isotropic shell per component, static path, shift-invert
// buckling, eigenvalues and their residual check.
//
// Usage: mech_buckling [file.bdf] [thickness] [sigma] [num_eigvals] [max_lanczos] [f5 prefix]
// Defaults are the reference driver's: mech-cylinder.bdf, t = 0.010,
// sigma = 10, 50 eigenvalues, 100 Lanczos vectors, tolerance 1e-12.  With
// an f5 prefix every mode is written, as the reference driver does, to
// <prefix><mode>.f5 with TACSToFH5 (connectivity, nodes, displacements,
// strains, stresses, extras).
// Difference from the reference driver: the solve is called with u0 =
// NULL, so the static path is solved (the reference passes a zero vector,
// which skips that solve and builds G from the BC values alone).
#include <cstdio>
#include <cstdlib>
#include <string>

#include "TACSShellGPU.h"

using namespace shellgpu;

int main(int argc, char *argv[]) {
  const char *bdf = argc > 1 ? argv[1] : "mech-cylinder.bdf";
  const TacsScalar thick = argc > 2 ? atof(argv[2]) : 0.010;
  const TacsScalar sigma = argc > 3 ? atof(argv[3]) : 10.0;
  const int num_eigvals = argc > 4 ? atoi(argv[4]) : 50;
  const int max_lanczos_vecs = argc > 5 ? atoi(argv[5]) : 100;
  const char *f5prefix = argc > 6 ? argv[6] : NULL;

  printf("Scanning BDF file\n");
  TACSMeshLoader *mesh = new TACSMeshLoader();
  mesh->incref();
  if (mesh->scanBDFFile(bdf)) return 1;
  int num_components = mesh->getNumComponents();

  TACSShellTransform *transform = new TACSShellNaturalTransform();
  transform->incref();

  // aluminium, as in the reference driver
  TacsScalar rho = 2718;
  TacsScalar specific_heat = 0.0;
  TacsScalar E = 72.0e9;
  TacsScalar nu = 0.33;
  TacsScalar ys = 1e11;
  TacsScalar cte = 10e-6;
  TacsScalar kappa = 0.0;
  TACSMaterialProperties *mat = new TACSMaterialProperties(rho, specific_heat, E, nu, ys, cte, kappa);
  mat->incref();

  for (int icomp = 0; icomp < num_components; icomp++) {
    const char *descriptor = mesh->getElementDescript(icomp);
    TACSIsoShellConstitutive *con = new TACSIsoShellConstitutive(mat, thick);
    TACSElement *shell = TacsCreateShellByName(descriptor, transform, con);
    if (!shell) {
      fprintf(stderr, "no shell element for %s\n", descriptor);
      return 1;
    }
    mesh->setElement(icomp, shell);
  }

  int vars_per_node = 6;
  TACSAssembler *assembler = mesh->createTACS(vars_per_node);
  if (!assembler) return 1;
  assembler->incref();
  mesh->decref();
  printf("nodes %d elements %d\n", assembler->getNumNodes(), assembler->getNumElements());

  // mechanical buckling: zero temperature on every element
  TacsScalar temperature = 0.0;
  for (int ielem = 0; ielem < assembler->getNumElements(); ielem++)
    assembler->getElement(ielem)->setTemperature(temperature);

  TACSBVec *f = assembler->createVec();
  f->incref();

  TACSSchurMat *kmat = assembler->createSchurMat();
  TACSSchurMat *gmat = assembler->createSchurMat();
  TACSSchurMat *aux_mat = assembler->createSchurMat();

  int lev = 1000000;
  double fill = 10.0;
  int reorder_schur = 1;
  TACSSchurPc *pc = new TACSSchurPc(aux_mat, lev, fill, reorder_schur);
  pc->incref();

  int subspaceSize = 10, nrestarts = 15, isFlexible = 0;
  GMRES *solver = new GMRES(aux_mat, pc, subspaceSize, nrestarts, isFlexible);
  solver->incref();
  solver->setTolerances(1e-12, 1e-12);

  double eig_tol = 1e-12;
  TACSLinearBuckling *buckling = new TACSLinearBuckling(assembler, sigma, gmat, kmat, aux_mat, solver,
                                                        max_lanczos_vecs, num_eigvals, eig_tol);
  buckling->incref();

  KSMPrint *ksm_print = new KSMPrintStdout("BucklingAnalysis", 0, 10);
  ksm_print->incref();

  buckling->solve(f, NULL, ksm_print);

  TacsScalar error;
  for (int imode = 0; imode < num_eigvals; imode++) {
    TacsScalar eig = buckling->extractEigenvalue(imode, &error);
    printf("lambda[%d] = %.16e error %.3e\n", imode, eig, error);
  }
  buckling->checkEigenvector(0);

  if (f5prefix) {
    // write each of the buckling modes to a file
    int write_flag = (TACS_OUTPUT_CONNECTIVITY | TACS_OUTPUT_NODES | TACS_OUTPUT_DISPLACEMENTS |
                      TACS_OUTPUT_STRAINS | TACS_OUTPUT_STRESSES | TACS_OUTPUT_EXTRAS);
    TACSToFH5 *f5 = new TACSToFH5(assembler, TACS_BEAM_OR_SHELL_ELEMENT, write_flag);
    f5->incref();
    TACSBVec *phi = assembler->createVec();
    phi->incref();
    for (int imode = 0; imode < num_eigvals; imode++) {
      buckling->extractEigenvector(imode, phi, &error);
      assembler->setVariables(phi);
      std::string filename = std::string(f5prefix) + std::to_string(imode) + ".f5";
      f5->writeToFile(filename.c_str());
    }
    phi->decref();
    f5->decref();
  }

  buckling->decref();
  ksm_print->decref();
  solver->decref();
  pc->decref();
  f->decref();
  mat->decref();
  transform->decref();
  assembler->decref();
  return 0;
}
