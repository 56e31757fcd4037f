// flopcount.cpp -- algorithmic flop count of one element Jacobian of the
// reference algorithm (test infrastructure).
//
// The oracle's restatement of the reference element (shell_elem.c, which
// follows TACSShellElement.h operation by operation) is compiled as C++ with
// `double` replaced by an operation-counting scalar: +, -, *, / and sqrt each
// count one flop; sign changes, comparisons and copies count none.  The
// counts are data independent (no data-dependent branches on the path).
//
// Usage: flopcount  -> prints a JSON object of flops per element
#include <cmath>
#include <cstdio>
#include <cstring>

static unsigned long long g_flops = 0;

struct CD {
  double v;
  CD() = default;
  CD(double x) : v(x) {}
  CD &operator+=(CD o) { v += o.v; g_flops++; return *this; }
  CD &operator-=(CD o) { v -= o.v; g_flops++; return *this; }
  CD &operator*=(CD o) { v *= o.v; g_flops++; return *this; }
  CD &operator/=(CD o) { v /= o.v; g_flops++; return *this; }
  CD operator-() const { return CD(-v); }
};
static inline CD operator+(CD a, CD b) { g_flops++; return CD(a.v + b.v); }
static inline CD operator-(CD a, CD b) { g_flops++; return CD(a.v - b.v); }
static inline CD operator*(CD a, CD b) { g_flops++; return CD(a.v * b.v); }
static inline CD operator/(CD a, CD b) { g_flops++; return CD(a.v / b.v); }
static inline bool operator==(CD a, CD b) { return a.v == b.v; }
static inline bool operator!=(CD a, CD b) { return a.v != b.v; }
static inline CD sqrt(CD a) { g_flops++; return CD(std::sqrt(a.v)); }

#define OR_NO_EXTERN_C
#define double CD
#include "shell_elem.c"
#undef double

int main() {
  or_material m;
  m.E = 70e9; m.nu = 0.3; m.rho = 2700.0; m.cte = 0.0; m.k_drill = 10.0;
  CD X[12], q[24], res[24], mat[576], psi[24];
  const double Xv[12] = {0, 1, 0, 0.1, 1, 0, 0, 0.98, 0.15, 0.1, 0.98, 0.15};
  for (int i = 0; i < 12; i++) X[i] = Xv[i];
  for (int i = 0; i < 24; i++) { q[i] = 1e-6 * (i + 1); psi[i] = 1.0; res[i] = 0.0; }
  unsigned long long c[6];
  g_flops = 0; or_elem_stiffness(OR_LINEAR, &m, CD(0.002), CD(0.0), X, q, mat); c[0] = g_flops;
  g_flops = 0; or_elem_stiffness(OR_NONLINEAR, &m, CD(0.002), CD(0.0), X, q, mat); c[1] = g_flops;
  g_flops = 0; or_elem_geometric_fd(&m, CD(0.002), CD(0.0), X, q, mat); c[2] = g_flops;
  g_flops = 0; or_elem_residual(OR_LINEAR, &m, CD(0.002), CD(0.0), X, q, res); c[3] = g_flops;
  g_flops = 0; or_elem_residual(OR_NONLINEAR, &m, CD(0.002), CD(0.0), X, q, res); c[4] = g_flops;
  CD dfdt = 0.0;
  g_flops = 0; or_elem_adj_res_dv(OR_LINEAR, &m, CD(0.002), CD(0.0), CD(1.0), psi, X, q, &dfdt); c[5] = g_flops;
  printf("{\"stiffness_linear\": %llu, \"tangent_nonlinear\": %llu, \"geometric_fd\": %llu, "
         "\"residual_linear\": %llu, \"residual_nonlinear\": %llu, \"adjoint_thickness\": %llu}\n",
         c[0], c[1], c[2], c[3], c[4], c[5]);
  return 0;
}
