// Golden-vector driver for the reference shell element (test infrastructure).
//
// This program is compiled only in the development container, against the
// reference's own element sources (see oracle/Makefile, target `ref`).  It
// links nothing from this repository's product and is never shipped.  It reads
// fixed-size input records and writes, for each record, what the reference's
// TACSQuad4Shell / TACSQuad4NonlinearShell compute:
//
//   record in  (66 doubles): Xpts[12] vars[24] psi[24] t E nu rho cte temperature
//   record out (2516 doubles):
//     res_lin[24]   TACSShellElement::addResidual, linear model      (SE.h:303-449)
//     res_nl[24]    same, nonlinear model
//     K_lin[576]    getMatType(TACS_STIFFNESS_MATRIX), linear model   (SE.h:675-771)
//     K_nl[576]     addJacobian(alpha=1, beta=0, gamma=0), nonlinear (SE.h:455-672)
//     G[576]        getMatType(TACS_GEOMETRIC_STIFFNESS_MATRIX)       (SE.h:705-766)
//     adj_lin       addAdjResProduct(scale=1, psi), linear model      (SE.h:774-868)
//     dRdt_lin      psi^T dR_lin/dt from a 5-point stencil in t
//     dRdt_nl       psi^T dR_nl/dt from a 5-point stencil in t
//     energy_lin    strain energy from computeEnergies, linear model  (SE.h:202-297)
//     M[576]        getMatType(TACS_MASS_MATRIX), linear model        (SE.h:675-704)
//     out_lin[160]  getOutputData(TACS_BEAM_OR_SHELL_ELEMENT, nodes | displacements |
//                   strains | stresses | extras), linear model: 4 nodes x 40 (SE.h:1402-1500)
//
// The residual is a polynomial of degree 3 in the thickness (A ~ t, D ~ t^3,
// drilling ~ t), so the 5-point central stencil with step 0.1 t is exact up to
// round-off; it pins the thickness adjoint independently of addAdjResProduct.
//
// Usage: ref_driver <in.bin> <out.bin>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "TACSIsoShellConstitutive.h"
#include "TACSMaterialProperties.h"
#include "TACSShellElementDefs.h"
#include "TACSShellElementTransform.h"

static const int NIN = 66;
static const int NOUT = 24 + 24 + 3 * 576 + 4 + 576 + 160;

static double psi_dot_res(TACSElement *elem, const double *X, const double *q,
                          const double *psi, const double *zero) {
  double res[24];
  memset(res, 0, sizeof(res));
  elem->addResidual(0, 0.0, X, q, zero, zero, res);
  double s = 0.0;
  for (int i = 0; i < 24; i++) s += psi[i] * res[i];
  return s;
}

// 5-point derivative of psi^T R(t) with respect to t at t0
static double dres_dt(TACSElement *elem, double t0, const double *X,
                      const double *q, const double *psi, const double *zero) {
  const double h = 0.1 * t0;
  double f[4];
  const double off[4] = {h, -h, 2.0 * h, -2.0 * h};
  for (int k = 0; k < 4; k++) {
    double tt = t0 + off[k];
    elem->setDesignVars(0, 1, &tt);
    f[k] = psi_dot_res(elem, X, q, psi, zero);
  }
  elem->setDesignVars(0, 1, &t0);
  return (8.0 * (f[0] - f[1]) - (f[2] - f[3])) / (12.0 * h);
}

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
    return 2;
  }
  MPI_Init(&argc, &argv);
  FILE *fin = fopen(argv[1], "rb");
  if (!fin) {
    fprintf(stderr, "cannot open %s\n", argv[1]);
    return 1;
  }
  std::vector<double> in;
  double buf[NIN];
  while (fread(buf, sizeof(double), NIN, fin) == (size_t)NIN) {
    in.insert(in.end(), buf, buf + NIN);
  }
  fclose(fin);
  const int nrec = (int)(in.size() / NIN);
  std::vector<double> out((size_t)nrec * NOUT, 0.0);

  double zero[24];
  memset(zero, 0, sizeof(zero));

  for (int r = 0; r < nrec; r++) {
    const double *X = &in[(size_t)r * NIN];
    const double *q = X + 12;
    const double *psi = X + 36;
    const double t = X[60], E = X[61], nu = X[62], rho = X[63], cte = X[64];
    const double temp = X[65];
    double *o = &out[(size_t)r * NOUT];

    TACSMaterialProperties *mat =
        new TACSMaterialProperties(rho, 0.0, E, nu, 1e11, cte, 0.0);
    mat->incref();
    TACSShellTransform *tr = new TACSShellNaturalTransform();
    tr->incref();
    TACSIsoShellConstitutive *con = new TACSIsoShellConstitutive(mat, t, 0);
    con->incref();

    TACSQuad4Shell *lin = new TACSQuad4Shell(tr, con);
    lin->incref();
    lin->setTemperature(temp);
    TACSQuad4NonlinearShell *nl = new TACSQuad4NonlinearShell(tr, con);
    nl->incref();
    nl->setTemperature(temp);

    double *res_lin = o;
    double *res_nl = o + 24;
    double *K_lin = o + 48;
    double *K_nl = o + 48 + 576;
    double *G = o + 48 + 2 * 576;
    double *tail = o + 48 + 3 * 576;

    lin->addResidual(0, 0.0, X, q, zero, zero, res_lin);
    nl->addResidual(0, 0.0, X, q, zero, zero, res_nl);
    lin->getMatType(TACS_STIFFNESS_MATRIX, 0, 0.0, X, q, K_lin);
    double rdummy[24];
    memset(rdummy, 0, sizeof(rdummy));
    nl->addJacobian(0, 0.0, 1.0, 0.0, 0.0, X, q, zero, zero, rdummy, K_nl);
    lin->getMatType(TACS_GEOMETRIC_STIFFNESS_MATRIX, 0, 0.0, X, q, G);

    double dfdx = 0.0;
    lin->addAdjResProduct(0, 0.0, 1.0, psi, X, q, zero, zero, 1, &dfdx);
    tail[0] = dfdx;
    tail[1] = dres_dt(lin, t, X, q, psi, zero);
    tail[2] = dres_dt(nl, t, X, q, psi, zero);
    double Te = 0.0, Ue = 0.0;
    lin->computeEnergies(0, 0.0, X, q, zero, &Te, &Ue);
    tail[3] = Ue;
    lin->getMatType(TACS_MASS_MATRIX, 0, 0.0, X, q, tail + 4);
    const int flag = TACS_OUTPUT_NODES | TACS_OUTPUT_DISPLACEMENTS | TACS_OUTPUT_STRAINS |
                     TACS_OUTPUT_STRESSES | TACS_OUTPUT_EXTRAS;
    lin->getOutputData(0, TACS_BEAM_OR_SHELL_ELEMENT, flag, X, q, zero, zero, 40, tail + 4 + 576);

    nl->decref();
    lin->decref();
    con->decref();
    tr->decref();
    mat->decref();
  }

  FILE *fout = fopen(argv[2], "wb");
  if (!fout) {
    fprintf(stderr, "cannot open %s\n", argv[2]);
    return 1;
  }
  fwrite(out.data(), sizeof(double), out.size(), fout);
  fclose(fout);
  MPI_Finalize();
  return 0;
}
