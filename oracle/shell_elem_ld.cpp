// shell_elem_ld.cpp -- the oracle's element restatement (shell_elem.c)
// evaluated in x87 extended precision (long double, 64-bit significand).
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// The geometric stiffness of the reference is a central difference of the
// nonlinear tangent along the state (TACSShellElement.h:705-766).  The
// tangent is exactly quadratic in the state, so the difference has no
// truncation error, but in double precision its round-off grows on fine
// meshes (the quadratic-strain term scales with the inverse square of the
// element size): on the 200 x 400 cylinder the double value moves by ~1e-10
// (row-floored relative) between steps 1 and 0.1.  The same code compiled
// with `double` replaced by `long double` removes that round-off (about
// 2,000 times smaller unit round-off) and is the reference for the
// geometric stiffness of the full-size tests.  The arithmetic is otherwise
// identical to shell_elem.c, which the golden vectors pin.
#include <cmath>
#include <cstring>

#include "oracle.h"

namespace ld {
using std::exp;
using std::log;
using std::sqrt;
#define OR_NO_EXTERN_C
#define double long double
#include "shell_elem.c"
#undef double
#undef OR_NO_EXTERN_C
}  // namespace ld

extern "C" void or_elem_geometric_step_ld(const or_material *m, double t, double temp, const double X[12],
                                          const double u[24], double dh, double mat[576]) {
  long double Xl[12], ul[24], Ml[576];
  for (int i = 0; i < 12; i++) Xl[i] = X[i];
  for (int i = 0; i < 24; i++) ul[i] = u[i];
  ld::or_elem_geometric_step(m, t, temp, Xl, ul, dh, Ml);
  for (int i = 0; i < 576; i++) mat[i] = (double)Ml[i];
}
