// CPU timing of the reference's own shell element (benchmark infrastructure).
//
// Compiled only in the development container, against the reference's
// element sources (oracle/Makefile, target `ref`); it links nothing from
// this repository's product and is never shipped.  It times
// TACSQuad4Shell::getMatType(TACS_STIFFNESS_MATRIX) (src/elements/shell/
// TACSShellElement.h:675-771, linear model, natural transform, isotropic
// material with the element's thickness as its design variable) over a set
// of elements, repeated for about the requested time on one core.
//
//   record in (13 doubles): Xpts[12] t
//   usage: ref_timer <in.bin> <seconds>
//   prints one line: <element Jacobians> <seconds> <checksum>
//
// The checksum (sum of all K entries of the last pass) keeps the work alive
// and lets the caller compare it with the product's K.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "TACSIsoShellConstitutive.h"
#include "TACSMaterialProperties.h"
#include "TACSShellElementDefs.h"
#include "TACSShellElementTransform.h"

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s in.bin seconds\n", argv[0]);
    return 2;
  }
  MPI_Init(&argc, &argv);
  const double seconds = atof(argv[2]);
  FILE *fin = fopen(argv[1], "rb");
  if (!fin) {
    fprintf(stderr, "cannot open %s\n", argv[1]);
    return 1;
  }
  std::vector<double> in;
  double buf[13];
  while (fread(buf, sizeof(double), 13, fin) == 13) in.insert(in.end(), buf, buf + 13);
  fclose(fin);
  const int nrec = (int)(in.size() / 13);
  if (nrec == 0) {
    fprintf(stderr, "no element records\n");
    return 1;
  }

  TACSMaterialProperties *mat = new TACSMaterialProperties(2700.0, 0.0, 70e9, 0.3, 1e11, 0.0, 0.0);
  mat->incref();
  TACSShellTransform *tr = new TACSShellNaturalTransform();
  tr->incref();
  TACSIsoShellConstitutive *con = new TACSIsoShellConstitutive(mat, in[12], 0);
  con->incref();
  TACSQuad4Shell *elem = new TACSQuad4Shell(tr, con);
  elem->incref();

  double vars[24];
  memset(vars, 0, sizeof(vars));
  std::vector<double> K(576);
  double checksum = 0.0;
  auto pass = [&]() {
    checksum = 0.0;
    for (int r = 0; r < nrec; r++) {
      const double *X = &in[(size_t)r * 13];
      double t = X[12];
      elem->setDesignVars(0, 1, &t);
      elem->getMatType(TACS_STIFFNESS_MATRIX, 0, 0.0, X, vars, K.data());
      for (int k = 0; k < 576; k++) checksum += K[k];
    }
  };
  pass();  // warm-up
  long done = 0;
  const auto t0 = std::chrono::steady_clock::now();
  double used = 0.0;
  while (used < seconds) {
    pass();
    done += nrec;
    used = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  printf("%ld %.6f %.17g\n", done, used, checksum);

  elem->decref();
  con->decref();
  tr->decref();
  mat->decref();
  MPI_Finalize();
  return 0;
}
