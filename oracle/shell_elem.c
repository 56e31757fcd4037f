/*
 * shell_elem.c -- oracle restatement of the reference's 4-node director shell
 * (TACSQuad4Shell / TACSQuad4NonlinearShell with TACSLinearizedRotation,
 * TACSShellNaturalTransform, TACSIsoShellConstitutive).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The structure follows the reference routine by routine: node normals,
 * nodal drilling strain, director, tying strain, per-quadrature-point
 * displacement gradient / strain / stress, then the reverse (residual) and
 * second-order (Jacobian) passes mapped back through the same chain.  The
 * long per-entry Hessian tables of the strain models (SM.h:528-620,
 * SM.h:1213-1720) are restated in closed form from the strain definitions
 * (their values are the same; the summation order inside an entry differs).
 *
 * Pinned by: tests/golden/element_golden.npz (reference element compiled from
 * source, records of residual, K, nonlinear K, G and thickness adjoint on
 * random and cylinder elements) -- see tests/test_oracle_golden.py.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

#define NN 4
#define VPN 6
#define NV 24
#define NTY 9

/* TACSGaussQuadrature.h:26-27 (truncated literal, kept as is) */
static const double GAUSS2 = 0.577350269189626;

/* ------------------------------------------------------------------ */
/* 3x3 row-major algebra (EA.h:39, 108, 825-970, 1094-1290, 1331-1400, 1980) */
/* ------------------------------------------------------------------ */
static void m_mul(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
static void m_mul_add(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      C[3 * i + j] += A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
/* C = A B^T */
static void m_mul_bt(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      C[3 * i + j] = A[3 * i] * B[3 * j] + A[3 * i + 1] * B[3 * j + 1] + A[3 * i + 2] * B[3 * j + 2];
}
static void m_mul_bt_add(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      C[3 * i + j] += A[3 * i] * B[3 * j] + A[3 * i + 1] * B[3 * j + 1] + A[3 * i + 2] * B[3 * j + 2];
}
/* C = A^T B */
static void m_mul_at(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      C[3 * i + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
}
/* inverse by cofactors, determinant expanded on the third row (EA.h:1980) */
static double m_inv(const double *A, double *Ai) {
  double det = (A[8] * (A[0] * A[4] - A[3] * A[1]) - A[7] * (A[0] * A[5] - A[3] * A[2]) +
                A[6] * (A[1] * A[5] - A[2] * A[4]));
  double r = 1.0 / det;
  Ai[0] = (A[4] * A[8] - A[5] * A[7]) * r;
  Ai[1] = -(A[1] * A[8] - A[2] * A[7]) * r;
  Ai[2] = (A[1] * A[5] - A[2] * A[4]) * r;
  Ai[3] = -(A[3] * A[8] - A[5] * A[6]) * r;
  Ai[4] = (A[0] * A[8] - A[2] * A[6]) * r;
  Ai[5] = -(A[0] * A[5] - A[2] * A[3]) * r;
  Ai[6] = (A[3] * A[7] - A[4] * A[6]) * r;
  Ai[7] = -(A[0] * A[7] - A[1] * A[6]) * r;
  Ai[8] = (A[0] * A[4] - A[1] * A[3]) * r;
  return det;
}
static void cross3(const double *x, const double *y, double *o) {
  o[0] = x[1] * y[2] - x[2] * y[1];
  o[1] = x[2] * y[0] - x[0] * y[2];
  o[2] = x[0] * y[1] - x[1] * y[0];
}
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* symmetric 6-vector S = [s11 s12 s13 s22 s23 s33]; A = T^T S T (EA.h:1094) */
static void sym_tt(const double *T, const double *S, double *A) {
  double W[9];
  W[0] = S[0] * T[0] + S[1] * T[3] + S[2] * T[6];
  W[1] = S[0] * T[1] + S[1] * T[4] + S[2] * T[7];
  W[2] = S[0] * T[2] + S[1] * T[5] + S[2] * T[8];
  W[3] = S[1] * T[0] + S[3] * T[3] + S[4] * T[6];
  W[4] = S[1] * T[1] + S[3] * T[4] + S[4] * T[7];
  W[5] = S[1] * T[2] + S[3] * T[5] + S[4] * T[8];
  W[6] = S[2] * T[0] + S[4] * T[3] + S[5] * T[6];
  W[7] = S[2] * T[1] + S[4] * T[4] + S[5] * T[7];
  W[8] = S[2] * T[2] + S[4] * T[5] + S[5] * T[8];
  A[0] = T[0] * W[0] + T[3] * W[3] + T[6] * W[6];
  A[1] = T[0] * W[1] + T[3] * W[4] + T[6] * W[7];
  A[2] = T[0] * W[2] + T[3] * W[5] + T[6] * W[8];
  A[3] = T[1] * W[1] + T[4] * W[4] + T[7] * W[7];
  A[4] = T[1] * W[2] + T[4] * W[5] + T[7] * W[8];
  A[5] = T[2] * W[2] + T[5] * W[5] + T[8] * W[8];
}
/* transpose of the linear map S -> A above, acting on dA (EA.h:1179) */
static void sym_tt_sens(const double *T, const double *dA, double *dS) {
  double dW[9];
  dW[0] = T[0] * dA[0];
  dW[1] = T[0] * dA[1] + T[1] * dA[3];
  dW[2] = T[0] * dA[2] + T[1] * dA[4] + T[2] * dA[5];
  dW[3] = T[3] * dA[0];
  dW[4] = T[3] * dA[1] + T[4] * dA[3];
  dW[5] = T[3] * dA[2] + T[4] * dA[4] + T[5] * dA[5];
  dW[6] = T[6] * dA[0];
  dW[7] = T[6] * dA[1] + T[7] * dA[3];
  dW[8] = T[6] * dA[2] + T[7] * dA[4] + T[8] * dA[5];
  dS[0] = T[0] * dW[0] + T[1] * dW[1] + T[2] * dW[2];
  dS[1] = T[3] * dW[0] + T[4] * dW[1] + T[5] * dW[2] + T[0] * dW[3] + T[1] * dW[4] + T[2] * dW[5];
  dS[2] = T[6] * dW[0] + T[7] * dW[1] + T[8] * dW[2] + T[0] * dW[6] + T[1] * dW[7] + T[2] * dW[8];
  dS[3] = T[3] * dW[3] + T[4] * dW[4] + T[5] * dW[5];
  dS[4] = T[6] * dW[3] + T[7] * dW[4] + T[8] * dW[5] + T[3] * dW[6] + T[4] * dW[7] + T[5] * dW[8];
  dS[5] = T[6] * dW[6] + T[7] * dW[7] + T[8] * dW[8];
}
/* second-order version: d2S = L^T d2A L for the map L = sym_tt (EA.h:1220) */
static void sym_tt_hess(const double *T, const double *d2A, double *d2S) {
  double tmp[36], col[6], out[6];
  for (int i = 0; i < 6; i++) sym_tt_sens(T, &d2A[6 * i], &tmp[6 * i]);
  for (int j = 0; j < 6; j++) {
    for (int i = 0; i < 6; i++) col[i] = tmp[6 * i + j];
    sym_tt_sens(T, col, out);
    for (int i = 0; i < 6; i++) d2S[6 * i + j] = out[i];
  }
}
/* D = [a]x B [c]x (EA.h:1362) and D = [a]x B (EA.h:1331) */
static void skew_mat_skew(const double *a, const double *B, const double *c, double *D) {
  double t[9];
  t[0] = a[1] * B[6] - a[2] * B[3];
  t[1] = a[1] * B[7] - a[2] * B[4];
  t[2] = a[1] * B[8] - a[2] * B[5];
  t[3] = a[2] * B[0] - a[0] * B[6];
  t[4] = a[2] * B[1] - a[0] * B[7];
  t[5] = a[2] * B[2] - a[0] * B[8];
  t[6] = a[0] * B[3] - a[1] * B[0];
  t[7] = a[0] * B[4] - a[1] * B[1];
  t[8] = a[0] * B[5] - a[1] * B[2];
  D[0] = c[2] * t[1] - c[1] * t[2];
  D[3] = c[2] * t[4] - c[1] * t[5];
  D[6] = c[2] * t[7] - c[1] * t[8];
  D[1] = c[0] * t[2] - c[2] * t[0];
  D[4] = c[0] * t[5] - c[2] * t[3];
  D[7] = c[0] * t[8] - c[2] * t[6];
  D[2] = c[1] * t[0] - c[0] * t[1];
  D[5] = c[1] * t[3] - c[0] * t[4];
  D[8] = c[1] * t[6] - c[0] * t[7];
}
static void skew_mat(const double *a, const double *B, double *D) {
  D[0] = a[1] * B[6] - a[2] * B[3];
  D[1] = a[1] * B[7] - a[2] * B[4];
  D[2] = a[1] * B[8] - a[2] * B[5];
  D[3] = a[2] * B[0] - a[0] * B[6];
  D[4] = a[2] * B[1] - a[0] * B[7];
  D[5] = a[2] * B[2] - a[0] * B[8];
  D[6] = a[0] * B[3] - a[1] * B[0];
  D[7] = a[0] * B[4] - a[1] * B[1];
  D[8] = a[0] * B[5] - a[1] * B[2];
}

/* Second derivative through C = T1^T A B1 on the column index and
 * C = T2^T A B2 on the row index of an 81-entry Hessian (SU.h:167-299).
 * mode 0: set, 1: add, 2: add symmetric */
static void hess_tmm(const double *T1, const double *T2, const double *B1, const double *B2,
                     const double *d2C, double *d2A, int mode) {
  double tmp2[81], t1[9], t2[9];
  for (int i = 0; i < 9; i++) {
    m_mul_bt(&d2C[9 * i], B1, t1);
    m_mul(T1, t1, &tmp2[9 * i]);
  }
  for (int i = 0; i < 9; i++) {
    for (int j = 0; j < 9; j++) t1[j] = tmp2[i + 9 * j];
    m_mul_bt(t1, B2, t2);
    m_mul(T2, t2, t1);
    for (int j = 0; j < 9; j++) {
      if (mode == 0) {
        d2A[i + 9 * j] = t1[j];
      } else if (mode == 1) {
        d2A[i + 9 * j] += t1[j];
      } else {
        d2A[i + 9 * j] += t1[j];
        d2A[j + 9 * i] += t1[j];
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* Bilinear basis on [-1,1]^2, node n = i + 2j at (-1+2i, -1+2j)        */
/* (SB.h:119-300: Lagrange-Lobatto order 2)                             */
/* ------------------------------------------------------------------ */
static void node_point(int n, double pt[2]) {
  pt[0] = -1.0 + 2.0 * (n % 2);
  pt[1] = -1.0 + 2.0 * (n / 2);
}
static void quad_point(int n, double pt[2]) {
  pt[0] = (n % 2) ? GAUSS2 : -GAUSS2;
  pt[1] = (n / 2) ? GAUSS2 : -GAUSS2;
}
static void basis_eval(const double pt[2], double N[4], double Na[4], double Nb[4]) {
  double na[2] = {0.5 * (1.0 - pt[0]), 0.5 * (1.0 + pt[0])};
  double nb[2] = {0.5 * (1.0 - pt[1]), 0.5 * (1.0 + pt[1])};
  const double dn[2] = {-0.5, 0.5};
  for (int j = 0; j < 2; j++)
    for (int i = 0; i < 2; i++) {
      N[i + 2 * j] = na[i] * nb[j];
      Na[i + 2 * j] = dn[i] * nb[j];
      Nb[i + 2 * j] = na[i] * dn[j];
    }
}
/* interpFields<stride,3> / interpFieldsGrad<stride,3> (SB.h:180-230) */
static void interp3(const double pt[2], int stride, const double *v, double out[3]) {
  double N[4], Na[4], Nb[4];
  basis_eval(pt, N, Na, Nb);
  out[0] = out[1] = out[2] = 0.0;
  for (int n = 0; n < NN; n++)
    for (int k = 0; k < 3; k++) out[k] += N[n] * v[stride * n + k];
}
static double interp1(const double pt[2], const double *v) {
  double N[4], Na[4], Nb[4];
  basis_eval(pt, N, Na, Nb);
  double s = 0.0;
  for (int n = 0; n < NN; n++) s += N[n] * v[n];
  return s;
}
static void grad3(const double pt[2], int stride, const double *v, double g[6]) {
  double N[4], Na[4], Nb[4];
  basis_eval(pt, N, Na, Nb);
  for (int k = 0; k < 6; k++) g[k] = 0.0;
  for (int n = 0; n < NN; n++)
    for (int k = 0; k < 3; k++) {
      g[2 * k] += Na[n] * v[stride * n + k];
      g[2 * k + 1] += Nb[n] * v[stride * n + k];
    }
}
/* transposes (SB.h:200-260) */
static void interp3_t(const double pt[2], int stride, const double f[3], double *v) {
  double N[4], Na[4], Nb[4];
  basis_eval(pt, N, Na, Nb);
  for (int n = 0; n < NN; n++)
    for (int k = 0; k < 3; k++) v[stride * n + k] += N[n] * f[k];
}
static void grad3_t(const double pt[2], int stride, const double g[6], double *v) {
  double N[4], Na[4], Nb[4];
  basis_eval(pt, N, Na, Nb);
  for (int n = 0; n < NN; n++)
    for (int k = 0; k < 3; k++) v[stride * n + k] += Na[n] * g[2 * k] + Nb[n] * g[2 * k + 1];
}
/* addInterpFieldsOuterProduct<nb,nb,3,3> (SB.h:262-301):
 * mat[(nb*j+r)*ld + nb*i + c] += Ni Nj jac[3r+c] */
static void outer_ff(const double pt[2], int nb, const double jac[9], double *mat) {
  double N[4], Na[4], Nb[4];
  basis_eval(pt, N, Na, Nb);
  const int ld = NN * nb;
  for (int j = 0; j < NN; j++)
    for (int r = 0; r < 3; r++)
      for (int i = 0; i < NN; i++)
        for (int c = 0; c < 3; c++) mat[(nb * j + r) * ld + nb * i + c] += N[i] * N[j] * jac[3 * r + c];
}
/* addInterpGradOuterProduct<nb,nb,3,3> (SB.h:318-355); jac is 6x6 over
 * gradient indices 2k+a */
static void outer_gg(const double pt[2], int nb, const double jac[36], double *mat) {
  double N[4], Na[4], Nb[4];
  basis_eval(pt, N, Na, Nb);
  const int ld = NN * nb;
  for (int j = 0; j < NN; j++)
    for (int r = 0; r < 3; r++)
      for (int i = 0; i < NN; i++) {
        double Naa = Na[j] * Na[i], Nab = Na[j] * Nb[i], Nba = Nb[j] * Na[i], Nbb = Nb[j] * Nb[i];
        const double *j1 = &jac[6 * (2 * r)], *j2 = &jac[6 * (2 * r + 1)];
        for (int c = 0; c < 3; c++)
          mat[(nb * j + r) * ld + nb * i + c] +=
              Naa * j1[2 * c] + Nab * j1[2 * c + 1] + Nba * j2[2 * c] + Nbb * j2[2 * c + 1];
      }
}
/* addInterpGradMixedOuterProduct<3,3,3,3> (SB.h:370-460): jac is 3x6 (row
 * field, column gradient), jacT is the transposed coupling (row gradient,
 * column field) stored 3x6 with the row index of the field. */
static void outer_fg(const double pt[2], const double *jac, const double *jacT, double *mat) {
  double N[4], Na[4], Nb[4];
  basis_eval(pt, N, Na, Nb);
  const int ld = NN * 3;
  for (int j = 0; j < NN; j++)
    for (int r = 0; r < 3; r++)
      for (int i = 0; i < NN; i++)
        for (int c = 0; c < 3; c++) {
          double v = 0.0;
          if (jac) v += N[j] * Na[i] * jac[6 * r + 2 * c] + N[j] * Nb[i] * jac[6 * r + 2 * c + 1];
          if (jacT) v += N[i] * Na[j] * jacT[6 * c + 2 * r] + N[i] * Nb[j] * jacT[6 * c + 2 * r + 1];
          mat[(3 * j + r) * ld + 3 * i + c] += v;
        }
}

/* ------------------------------------------------------------------ */
/* MITC tying points (SB.h:463-565), order 2: 9 points                   */
/* 0,1: g11 at (0,-1),(0,1); 2,3: g22 at (-1,0),(1,0); 4: g12 at (0,0);  */
/* 5,6: g23 at (-1,0),(1,0); 7,8: g13 at (0,-1),(0,1)                     */
/* ------------------------------------------------------------------ */
enum { F11 = 0, F22 = 1, F12 = 2, F23 = 3, F13 = 4 };
static const int TY_FIELD[NTY] = {F11, F11, F22, F22, F12, F23, F23, F13, F13};
static const double TY_PT[NTY][2] = {{0.0, -1.0}, {0.0, 1.0}, {-1.0, 0.0}, {1.0, 0.0}, {0.0, 0.0},
                                     {-1.0, 0.0}, {1.0, 0.0}, {0.0, -1.0}, {0.0, 1.0}};
/* gty slot of each field: g11->0, g22->3, g12->1, g23->4, g13->2 */
static const int FIELD_SLOT[5] = {0, 3, 1, 4, 2};

/* evalTyingInterp (SB.h:568-607): Lagrange on knots {-1,1}, reduced {0} */
static void tying_interp(const double pt[2], double N[NTY]) {
  double na[2] = {(pt[0] - 1.0) * (-0.5), (pt[0] + 1.0) * 0.5};
  double nb[2] = {(pt[1] - 1.0) * (-0.5), (pt[1] + 1.0) * 0.5};
  N[0] = nb[0];
  N[1] = nb[1];
  N[2] = na[0];
  N[3] = na[1];
  N[4] = 1.0;
  N[5] = na[0];
  N[6] = na[1];
  N[7] = nb[0];
  N[8] = nb[1];
}
static void interp_tying(const double pt[2], const double ety[NTY], double gty[6]) {
  double N[NTY];
  tying_interp(pt, N);
  for (int k = 0; k < 6; k++) gty[k] = 0.0;
  for (int t = 0; t < NTY; t++) gty[FIELD_SLOT[TY_FIELD[t]]] += N[t] * ety[t];
}
static void interp_tying_t(const double pt[2], const double dgty[6], double dety[NTY]) {
  double N[NTY];
  tying_interp(pt, N);
  for (int t = 0; t < NTY; t++) dety[t] += dgty[FIELD_SLOT[TY_FIELD[t]]] * N[t];
}
static void interp_tying_hess(const double pt[2], const double d2gty[36], double d2ety[81]) {
  double N[NTY];
  tying_interp(pt, N);
  for (int a = 0; a < NTY; a++)
    for (int b = 0; b < NTY; b++)
      d2ety[NTY * a + b] += N[a] * N[b] * d2gty[6 * FIELD_SLOT[TY_FIELD[a]] + FIELD_SLOT[TY_FIELD[b]]];
}

/* ------------------------------------------------------------------ */
/* TACSShellNaturalTransform::computeTransform (ST.h:21-93).  The      */
/* in-plane direction is NOT orthogonalised against the normal: the    */
/* reference subtracts d*n[0] from t1[0] three times (ST.h:42-44).     */
/* ------------------------------------------------------------------ */
static void natural_transform(const double Xxi[6], const double n0[3], double T[9]) {
  double n[3] = {n0[0], n0[1], n0[2]};
  double inv = 1.0 / sqrt(dot3(n, n));
  n[0] *= inv;
  n[1] *= inv;
  n[2] *= inv;
  double t1[3] = {Xxi[0], Xxi[2], Xxi[4]};
  double d = dot3(n, t1);
  t1[0] = t1[0] - d * n[0];
  t1[0] = t1[0] - d * n[0];
  t1[0] = t1[0] - d * n[0];
  inv = 1.0 / sqrt(dot3(t1, t1));
  t1[0] *= inv;
  t1[1] *= inv;
  t1[2] *= inv;
  double t2[3];
  cross3(n, t1, t2);
  T[0] = t1[0];
  T[3] = t1[1];
  T[6] = t1[2];
  T[1] = t2[0];
  T[4] = t2[1];
  T[7] = t2[2];
  T[2] = n[0];
  T[5] = n[1];
  T[8] = n[2];
}

/* frame [a(2 cols interleaved) ; c] (SU.h:8-35) */
static void frame(const double xi[6], const double *c, double Xd[9]) {
  Xd[0] = xi[0];
  Xd[1] = xi[1];
  Xd[2] = c ? c[0] : 0.0;
  Xd[3] = xi[2];
  Xd[4] = xi[3];
  Xd[5] = c ? c[1] : 0.0;
  Xd[6] = xi[4];
  Xd[7] = xi[5];
  Xd[8] = c ? c[2] : 0.0;
}
/* TacsShellExtractFrame (SU.h:69-95) */
static void extract_frame(const double Xd[9], double xi[6], double *c) {
  xi[0] = Xd[0];
  xi[1] = Xd[1];
  xi[2] = Xd[3];
  xi[3] = Xd[4];
  xi[4] = Xd[6];
  xi[5] = Xd[7];
  if (c) {
    c[0] = Xd[2];
    c[1] = Xd[5];
    c[2] = Xd[8];
  }
}

/* ------------------------------------------------------------------ */
/* constitutive (ISC:153-226, TACSMaterialProperties.cpp:295-313)       */
/* ------------------------------------------------------------------ */
typedef struct {
  double A[6], B[6], D[6], As[3], drill;
} abd_t;
static void tangent_2d(const or_material *m, double C[6]) {
  double G = 0.5 * m->E / (1.0 + m->nu);
  double Dm = m->E / (1.0 - m->nu * m->nu);
  C[0] = Dm;
  C[1] = m->nu * Dm;
  C[2] = 0.0;
  C[3] = Dm;
  C[4] = 0.0;
  C[5] = G;
}
static void eval_abd(const or_material *m, double t, abd_t *c) {
  tangent_2d(m, c->A);
  double I = t * t * t / 12.0;
  for (int i = 0; i < 6; i++) {
    c->D[i] = I * c->A[i];
    c->A[i] *= t;
    c->B[i] = 0.0; /* tOffset = 0 */
  }
  c->As[0] = c->As[2] = (5.0 / 6.0) * c->A[5];
  c->As[1] = 0.0;
  c->drill = 0.5 * m->k_drill * (c->As[0] + c->As[2]);
}
/* TACSShellConstitutive::computeStress (TACSShellConstitutive.h:125-147) */
static void compute_stress(const abd_t *c, const double e[9], double s[9]) {
  const double *A = c->A, *B = c->B, *D = c->D, *As = c->As;
  s[0] = A[0] * e[0] + A[1] * e[1] + A[2] * e[2] + B[0] * e[3] + B[1] * e[4] + B[2] * e[5];
  s[1] = A[1] * e[0] + A[3] * e[1] + A[4] * e[2] + B[1] * e[3] + B[3] * e[4] + B[4] * e[5];
  s[2] = A[2] * e[0] + A[4] * e[1] + A[5] * e[2] + B[2] * e[3] + B[4] * e[4] + B[5] * e[5];
  s[3] = B[0] * e[0] + B[1] * e[1] + B[2] * e[2] + D[0] * e[3] + D[1] * e[4] + D[2] * e[5];
  s[4] = B[1] * e[0] + B[3] * e[1] + B[4] * e[2] + D[1] * e[3] + D[3] * e[4] + D[4] * e[5];
  s[5] = B[2] * e[0] + B[4] * e[1] + B[5] * e[2] + D[2] * e[3] + D[4] * e[4] + D[5] * e[5];
  s[6] = As[0] * e[6] + As[1] * e[7];
  s[7] = As[1] * e[6] + As[2] * e[7];
  s[8] = c->drill * e[8];
}
/* symmetric 3x3 stored [0 1 2; 1 3 4; 2 4 5] */
static double sym_inner(const double *S, const double *x, const double *y) {
  return y[0] * (S[0] * x[0] + S[1] * x[1] + S[2] * x[2]) + y[1] * (S[1] * x[0] + S[3] * x[1] + S[4] * x[2]) +
         y[2] * (S[2] * x[0] + S[4] * x[1] + S[5] * x[2]);
}

/* ------------------------------------------------------------------ */
/* strain models (SM.h:440-620 linear, SM.h:1115-1720 nonlinear)       */
/* e[0..2] membrane, e[3..5] bending, e[6..7] shear, e[8] drill        */
/* ------------------------------------------------------------------ */
/* derivative of the three bending strains w.r.t. u0x[a] (g) and u1x[b] (h) */
static void bend_grads(int model, const double u0x[9], const double u1x[9], double g[9][3], double h[9][3]) {
  memset(g, 0, 27 * sizeof(double));
  memset(h, 0, 27 * sizeof(double));
  if (model == OR_NONLINEAR) {
    g[0][0] = u1x[0]; g[0][2] = u1x[1];
    g[1][1] = u1x[1]; g[1][2] = u1x[0];
    g[3][0] = u1x[3]; g[3][2] = u1x[4];
    g[4][1] = u1x[4]; g[4][2] = u1x[3];
    g[6][0] = u1x[6]; g[6][2] = u1x[7];
    g[7][1] = u1x[7]; g[7][2] = u1x[6];
    h[0][0] = 1.0 + u0x[0]; h[0][2] = u0x[1];
    h[1][1] = u0x[1];       h[1][2] = 1.0 + u0x[0];
    h[3][0] = u0x[3];       h[3][2] = 1.0 + u0x[4];
    h[4][1] = 1.0 + u0x[4]; h[4][2] = u0x[3];
    h[6][0] = u0x[6];       h[6][2] = u0x[7];
    h[7][1] = u0x[7];       h[7][2] = u0x[6];
  } else {
    h[0][0] = 1.0;
    h[1][2] = 1.0;
    h[3][2] = 1.0;
    h[4][1] = 1.0;
  }
}
static void eval_strain(int model, const double u0x[9], const double u1x[9], const double e0ty[6], double e[9]) {
  e[0] = e0ty[0];
  e[1] = e0ty[3];
  e[2] = 2.0 * e0ty[1];
  if (model == OR_NONLINEAR) {
    e[3] = u1x[0] + (u0x[0] * u1x[0] + u0x[3] * u1x[3] + u0x[6] * u1x[6]);
    e[4] = u1x[4] + (u0x[1] * u1x[1] + u0x[4] * u1x[4] + u0x[7] * u1x[7]);
    e[5] = u1x[1] + u1x[3] +
           (u0x[0] * u1x[1] + u0x[3] * u1x[4] + u0x[6] * u1x[7] + u1x[0] * u0x[1] + u1x[3] * u0x[4] +
            u1x[6] * u0x[7]);
  } else {
    e[3] = u1x[0];
    e[4] = u1x[4];
    e[5] = u1x[1] + u1x[3];
  }
  e[6] = 2.0 * e0ty[4];
  e[7] = 2.0 * e0ty[2];
}
/* evalStrainSens (SM.h:460-493 / 1136-1170) */
static void strain_sens(int model, double scale, const double s[9], const double u0x[9], const double u1x[9],
                        double du0x[9], double du1x[9], double de0ty[6]) {
  double g[9][3], h[9][3];
  bend_grads(model, u0x, u1x, g, h);
  de0ty[0] = scale * s[0];
  de0ty[1] = 2.0 * scale * s[2];
  de0ty[2] = 2.0 * scale * s[7];
  de0ty[3] = scale * s[1];
  de0ty[4] = 2.0 * scale * s[6];
  de0ty[5] = 0.0;
  for (int a = 0; a < 9; a++) {
    du0x[a] = scale * (s[3] * g[a][0] + s[4] * g[a][1] + s[5] * g[a][2]);
    du1x[a] = scale * (s[3] * h[a][0] + s[4] * h[a][1] + s[5] * h[a][2]);
  }
}
/* evalStrainDeriv (SM.h:495-526 / 1172-1211): forward derivative ed */
static void strain_deriv(int model, const double u0x[9], const double u1x[9], const double u0xd[9],
                         const double u1xd[9], const double e0tyd[6], double ed[9]) {
  double g[9][3], h[9][3];
  bend_grads(model, u0x, u1x, g, h);
  ed[0] = e0tyd[0];
  ed[1] = e0tyd[3];
  ed[2] = 2.0 * e0tyd[1];
  for (int i = 0; i < 3; i++) {
    double v = 0.0;
    for (int a = 0; a < 9; a++) v += g[a][i] * u0xd[a] + h[a][i] * u1xd[a];
    ed[3 + i] = v;
  }
  ed[6] = 2.0 * e0tyd[4];
  ed[7] = 2.0 * e0tyd[2];
}
/* evalStrainHessian, restated in closed form (SM.h:528-620, SM.h:1213-1720):
 * the Hessian of scale * 0.5 e^T C e w.r.t. (u0x, u1x, e0ty) plus, for the
 * nonlinear model, scale * s . d2e/du0x du1x. */
static void strain_hessian(int model, double scale, const double s[9], const abd_t *c, const double u0x[9],
                           const double u1x[9], double d2u0x[81], double d2u1x[81], double d2u0xu1x[81],
                           double d2e0ty[36], double d2e0tyu0x[54], double d2e0tyu1x[54]) {
  double g[9][3], h[9][3];
  bend_grads(model, u0x, u1x, g, h);
  const double *A = c->A, *B = c->B, *D = c->D, *As = c->As;
  /* symmetric 3x3 -> full */
  double Df[9] = {D[0], D[1], D[2], D[1], D[3], D[4], D[2], D[4], D[5]};
  double Bf[9] = {B[0], B[1], B[2], B[1], B[3], B[4], B[2], B[4], B[5]};
  double Dg[9][3], Dh[9][3];
  for (int a = 0; a < 9; a++)
    for (int i = 0; i < 3; i++) {
      Dg[a][i] = Df[3 * i] * g[a][0] + Df[3 * i + 1] * g[a][1] + Df[3 * i + 2] * g[a][2];
      Dh[a][i] = Df[3 * i] * h[a][0] + Df[3 * i + 1] * h[a][1] + Df[3 * i + 2] * h[a][2];
    }
  for (int a = 0; a < 9; a++)
    for (int b = 0; b < 9; b++) {
      d2u0x[9 * a + b] = scale * (Dg[a][0] * g[b][0] + Dg[a][1] * g[b][1] + Dg[a][2] * g[b][2]);
      d2u1x[9 * a + b] = scale * (Dh[a][0] * h[b][0] + Dh[a][1] * h[b][1] + Dh[a][2] * h[b][2]);
      d2u0xu1x[9 * a + b] = Dg[a][0] * h[b][0] + Dg[a][1] * h[b][1] + Dg[a][2] * h[b][2];
    }
  if (model == OR_NONLINEAR) {
    /* second derivatives of the bending strains: d2e3/du0x_a du1x_a etc. */
    const int idx3[3][2] = {{0, 0}, {3, 3}, {6, 6}};
    const int idx4[3][2] = {{1, 1}, {4, 4}, {7, 7}};
    const int idx5[6][2] = {{0, 1}, {3, 4}, {6, 7}, {1, 0}, {4, 3}, {7, 6}};
    for (int k = 0; k < 3; k++) d2u0xu1x[9 * idx3[k][0] + idx3[k][1]] += s[3];
    for (int k = 0; k < 3; k++) d2u0xu1x[9 * idx4[k][0] + idx4[k][1]] += s[4];
    for (int k = 0; k < 6; k++) d2u0xu1x[9 * idx5[k][0] + idx5[k][1]] += s[5];
  }
  for (int k = 0; k < 81; k++) d2u0xu1x[k] *= scale;

  /* membrane / shear strains from e0ty: em = (e0ty0, e0ty3, 2 e0ty1),
   * es = (2 e0ty4, 2 e0ty2) */
  memset(d2e0ty, 0, 36 * sizeof(double));
  const int mslot[3] = {0, 3, 1};
  const double mfac[3] = {1.0, 1.0, 2.0};
  double Af[9] = {A[0], A[1], A[2], A[1], A[3], A[4], A[2], A[4], A[5]};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) d2e0ty[6 * mslot[i] + mslot[j]] = scale * mfac[i] * mfac[j] * Af[3 * i + j];
  const int sslot[2] = {4, 2};
  double Asf[4] = {As[0], As[1], As[1], As[2]};
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) d2e0ty[6 * sslot[i] + sslot[j]] = 4.0 * scale * Asf[2 * i + j];

  /* membrane-bending coupling through B */
  memset(d2e0tyu0x, 0, 54 * sizeof(double));
  memset(d2e0tyu1x, 0, 54 * sizeof(double));
  for (int i = 0; i < 3; i++)
    for (int a = 0; a < 9; a++) {
      double vg = Bf[3 * i] * g[a][0] + Bf[3 * i + 1] * g[a][1] + Bf[3 * i + 2] * g[a][2];
      double vh = Bf[3 * i] * h[a][0] + Bf[3 * i + 1] * h[a][1] + Bf[3 * i + 2] * h[a][2];
      d2e0tyu0x[9 * mslot[i] + a] = scale * mfac[i] * vg;
      d2e0tyu1x[9 * mslot[i] + a] = scale * mfac[i] * vh;
    }
}

/* ------------------------------------------------------------------ */
/* tying strain at the 9 tying points (SM.h:32-77 linear, 643-697 nl)  */
/* ------------------------------------------------------------------ */
typedef struct {
  double Xxi[NTY][6], n0[NTY][3], Uxi[NTY][6], d0[NTY][3];
} tying_data;

static void tying_setup(const double X[12], const double fn[12], const double q[24], const double d[12],
                        tying_data *td) {
  for (int t = 0; t < NTY; t++) {
    grad3(TY_PT[t], 3, X, td->Xxi[t]);
    interp3(TY_PT[t], 3, fn, td->n0[t]);
    grad3(TY_PT[t], VPN, q, td->Uxi[t]);
    interp3(TY_PT[t], 3, d, td->d0[t]);
  }
}
/* value of the tying strain and its first derivative coefficients:
 * ety = f(Uxi, d0); dUxi[6] = d ety / d Uxi, dd0[3] = d ety / d d0 */
static double tying_eval(int model, int t, const tying_data *td, double dUxi[6], double dd0[3]) {
  const double *Xxi = td->Xxi[t], *U = td->Uxi[t], *n0 = td->n0[t], *d0 = td->d0[t];
  const double nl = (model == OR_NONLINEAR) ? 1.0 : 0.0;
  double a[6], nd[3];
  for (int k = 0; k < 6; k++) a[k] = Xxi[k] + nl * U[k];
  for (int k = 0; k < 3; k++) nd[k] = n0[k] + nl * d0[k];
  for (int k = 0; k < 6; k++) dUxi[k] = 0.0;
  dd0[0] = dd0[1] = dd0[2] = 0.0;
  double v = 0.0;
  switch (TY_FIELD[t]) {
    case F11:
      v = U[0] * Xxi[0] + U[2] * Xxi[2] + U[4] * Xxi[4];
      if (nl != 0.0) v += 0.5 * (U[0] * U[0] + U[2] * U[2] + U[4] * U[4]);
      dUxi[0] = a[0];
      dUxi[2] = a[2];
      dUxi[4] = a[4];
      break;
    case F22:
      v = U[1] * Xxi[1] + U[3] * Xxi[3] + U[5] * Xxi[5];
      if (nl != 0.0) v += 0.5 * (U[1] * U[1] + U[3] * U[3] + U[5] * U[5]);
      dUxi[1] = a[1];
      dUxi[3] = a[3];
      dUxi[5] = a[5];
      break;
    case F12:
      v = U[0] * Xxi[1] + U[2] * Xxi[3] + U[4] * Xxi[5] + U[1] * Xxi[0] + U[3] * Xxi[2] + U[5] * Xxi[4];
      if (nl != 0.0) v += U[0] * U[1] + U[2] * U[3] + U[4] * U[5];
      v *= 0.5;
      dUxi[0] = 0.5 * a[1];
      dUxi[1] = 0.5 * a[0];
      dUxi[2] = 0.5 * a[3];
      dUxi[3] = 0.5 * a[2];
      dUxi[4] = 0.5 * a[5];
      dUxi[5] = 0.5 * a[4];
      break;
    case F23:
      v = 0.5 * (Xxi[1] * d0[0] + Xxi[3] * d0[1] + Xxi[5] * d0[2] + nd[0] * U[1] + nd[1] * U[3] + nd[2] * U[5]);
      dUxi[1] = 0.5 * nd[0];
      dUxi[3] = 0.5 * nd[1];
      dUxi[5] = 0.5 * nd[2];
      dd0[0] = 0.5 * a[1];
      dd0[1] = 0.5 * a[3];
      dd0[2] = 0.5 * a[5];
      break;
    default: /* F13 */
      v = 0.5 * (Xxi[0] * d0[0] + Xxi[2] * d0[1] + Xxi[4] * d0[2] + nd[0] * U[0] + nd[1] * U[2] + nd[2] * U[4]);
      dUxi[0] = 0.5 * nd[0];
      dUxi[2] = 0.5 * nd[1];
      dUxi[4] = 0.5 * nd[2];
      dd0[0] = 0.5 * a[0];
      dd0[1] = 0.5 * a[2];
      dd0[2] = 0.5 * a[4];
      break;
  }
  return v;
}

/* addComputeTyingStrainTranspose (SM.h:79-161 / 699-780) */
static void tying_transpose(int model, const tying_data *td, const double dety[NTY], double res[24],
                            double dd[12]) {
  for (int t = 0; t < NTY; t++) {
    double dU[6], dd0[3];
    tying_eval(model, t, td, dU, dd0);
    for (int k = 0; k < 6; k++) dU[k] *= dety[t];
    for (int k = 0; k < 3; k++) dd0[k] *= dety[t];
    if (TY_FIELD[t] == F23 || TY_FIELD[t] == F13) interp3_t(TY_PT[t], 3, dd0, dd);
    grad3_t(TY_PT[t], VPN, dU, res);
  }
}

/* addComputeTyingStrainHessian (SM.h:163-373 linear / 782-1031 nonlinear).
 * d2ety[9x9], d2etyu[9x12], d2etyd[9x12]; mat 24x24, d2d and d2du 12x12. */
static void tying_hessian(int model, double alpha, const tying_data *td, const double dety[NTY],
                          const double d2ety[81], const double d2etyu[108], const double d2etyd[108],
                          double mat[576], double d2d[144], double d2du[144]) {
  /* per tying point derivative vectors w.r.t. the 12 displacement and 12
   * director coefficients */
  double Ju[NTY][12], Jd[NTY][12];
  for (int t = 0; t < NTY; t++) {
    double dU[6], dd0[3];
    tying_eval(model, t, td, dU, dd0);
    memset(Ju[t], 0, sizeof(Ju[t]));
    memset(Jd[t], 0, sizeof(Jd[t]));
    grad3_t(TY_PT[t], 3, dU, Ju[t]);
    if (TY_FIELD[t] == F23 || TY_FIELD[t] == F13) interp3_t(TY_PT[t], 3, dd0, Jd[t]);
  }
  for (int t1 = 0; t1 < NTY; t1++) {
    double du2[12], dd2[12];
    memset(du2, 0, sizeof(du2));
    memset(dd2, 0, sizeof(dd2));
    for (int t2 = 0; t2 < NTY; t2++) {
      double v = d2ety[NTY * t1 + t2];
      for (int k = 0; k < 12; k++) {
        du2[k] += v * Ju[t2][k];
        dd2[k] += v * Jd[t2][k];
      }
    }
    const double *du1 = Ju[t1], *dd1 = Jd[t1];
    const double *etu = &d2etyu[12 * t1], *etd = &d2etyd[12 * t1];
    const int shear = (TY_FIELD[t1] == F23 || TY_FIELD[t1] == F13);
    if (shear) {
      for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++) {
          d2d[12 * i + j] += dd1[i] * dd2[j] + dd1[i] * etd[j] + etd[i] * dd1[j];
          d2du[12 * i + j] += dd1[i] * du2[j] + dd1[i] * etu[j];
        }
    }
    for (int i = 0; i < 12; i++)
      for (int j = 0; j < 12; j++) d2du[12 * i + j] += etd[i] * du1[j];
    for (int i = 0; i < 12; i++) {
      int ii = VPN * (i / 3) + i % 3;
      for (int j = 0; j < 12; j++) {
        int jj = VPN * (j / 3) + j % 3;
        mat[NV * ii + jj] += du1[i] * du2[j] + du1[i] * etu[j] + etu[i] * du1[j];
      }
    }
    if (model == OR_NONLINEAR) {
      /* second derivatives of the quadratic tying strain terms */
      double c = alpha * dety[t1];
      double d2U[36], d2dU[18];
      memset(d2U, 0, sizeof(d2U));
      memset(d2dU, 0, sizeof(d2dU));
      switch (TY_FIELD[t1]) {
        case F11:
          d2U[0] = d2U[14] = d2U[28] = c;
          break;
        case F22:
          d2U[7] = d2U[21] = d2U[35] = c;
          break;
        case F12:
          d2U[1] = d2U[6] = d2U[15] = d2U[20] = d2U[29] = d2U[34] = 0.5 * c;
          break;
        case F23:
          d2dU[1] = d2dU[9] = d2dU[17] = 0.5 * c;
          break;
        default:
          d2dU[0] = d2dU[8] = d2dU[16] = 0.5 * c;
          break;
      }
      if (shear) {
        outer_fg(TY_PT[t1], d2dU, NULL, d2du);
      } else {
        outer_gg(TY_PT[t1], VPN, d2U, mat);
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* per-element geometric / kinematic set-up                             */
/* ------------------------------------------------------------------ */
typedef struct {
  double fn[12], Xdn[36];
  double Tn[36], XdinvTn[36], u0xn[36], Ctn[36], etn[4];
  double d[12];
  tying_data td;
  double ety[NTY];
} elem_setup;

/* TacsShellComputeNodeNormals (SU.h:301-343) */
static void node_normals(const double X[12], double fn[12], double Xdn[36]) {
  for (int n = 0; n < NN; n++) {
    double pt[2], Xxi[6];
    node_point(n, pt);
    grad3(pt, 3, X, Xxi);
    double a[3] = {Xxi[0], Xxi[2], Xxi[4]};
    double b[3] = {Xxi[1], Xxi[3], Xxi[5]};
    cross3(a, b, &fn[3 * n]);
    double nrm = sqrt(dot3(&fn[3 * n], &fn[3 * n]));
    if (nrm != 0.0) {
      double inv = 1.0 / nrm;
      fn[3 * n] *= inv;
      fn[3 * n + 1] *= inv;
      fn[3 * n + 2] *= inv;
    }
    frame(Xxi, &fn[3 * n], &Xdn[9 * n]);
  }
}

/* TacsShellComputeDrillStrain (SU.h:649-707) with TACSLinearizedRotation
 * (SD.h:21-37, 560-564): C = I - q^x, et = 0.5 (Ct[3] + u0x[3] - Ct[1] - u0x[1]) */
static void drill_setup(const double q[24], elem_setup *es) {
  for (int n = 0; n < NN; n++) {
    double pt[2], Xxi[6], u0xi[6], Xdinv[9], C[9], tmp[9];
    node_point(n, pt);
    extract_frame(&es->Xdn[9 * n], Xxi, NULL);
    natural_transform(Xxi, &es->fn[3 * n], &es->Tn[9 * n]);
    grad3(pt, VPN, q, u0xi);
    m_inv(&es->Xdn[9 * n], Xdinv);
    m_mul(Xdinv, &es->Tn[9 * n], &es->XdinvTn[9 * n]);
    frame(u0xi, NULL, &es->u0xn[9 * n]);
    const double *qr = &q[VPN * n + 3];
    C[0] = 1.0; C[1] = qr[2];  C[2] = -qr[1];
    C[3] = -qr[2]; C[4] = 1.0; C[5] = qr[0];
    C[6] = qr[1];  C[7] = -qr[0]; C[8] = 1.0;
    m_mul_at(&es->Tn[9 * n], C, tmp);
    m_mul(tmp, &es->Tn[9 * n], &es->Ctn[9 * n]);
    m_mul(&es->u0xn[9 * n], &es->XdinvTn[9 * n], tmp);
    m_mul_at(&es->Tn[9 * n], tmp, &es->u0xn[9 * n]);
    const double *Ct = &es->Ctn[9 * n], *u0x = &es->u0xn[9 * n];
    es->etn[n] = 0.5 * (Ct[3] + u0x[3] - Ct[1] - u0x[1]);
  }
}

static void elem_setup_all(int model, const double X[12], const double q[24], elem_setup *es) {
  node_normals(X, es->fn, es->Xdn);
  drill_setup(q, es);
  /* director d = q_rot x t (SD.h:215-242) */
  for (int n = 0; n < NN; n++) cross3(&q[VPN * n + 3], &es->fn[3 * n], &es->d[3 * n]);
  tying_setup(X, es->fn, q, es->d, &es->td);
  for (int t = 0; t < NTY; t++) {
    double dU[6], dd0[3];
    es->ety[t] = tying_eval(model, t, &es->td, dU, dd0);
  }
}

/* quadrature-point kinematics (SE.h:341-362 + TacsShellComputeDispGrad,
 * SU.h:361-422) */
typedef struct {
  double X[3], Xxi[6], n0[3], T[9], et;
  double XdinvT[9], XdinvzT[9], u0x[9], u1x[9], detXd;
} qp_data;

static void qp_kinematics(const double pt[2], double weight, const double Xpts[12], const double q[24],
                          const elem_setup *es, qp_data *qp) {
  interp3(pt, 3, Xpts, qp->X);
  grad3(pt, 3, Xpts, qp->Xxi);
  interp3(pt, 3, es->fn, qp->n0);
  qp->et = interp1(pt, es->etn);
  natural_transform(qp->Xxi, qp->n0, qp->T);

  double nxi[6], Xd[9], Xdz[9], Xdinv[9], nXdz[9];
  grad3(pt, 3, es->fn, nxi);
  frame(qp->Xxi, qp->n0, Xd);
  frame(nxi, NULL, Xdz);
  qp->detXd = m_inv(Xd, Xdinv);
  m_mul(Xdinv, Xdz, nXdz);
  for (int i = 0; i < 9; i++) nXdz[i] *= -1.0;
  m_mul(Xdinv, qp->T, qp->XdinvT);
  m_mul(nXdz, qp->XdinvT, qp->XdinvzT);

  double d0[3], d0xi[6], u0xi[6], tmp[9], u1d[9];
  interp3(pt, 3, es->d, d0);
  grad3(pt, 3, es->d, d0xi);
  grad3(pt, VPN, q, u0xi);
  frame(u0xi, d0, qp->u0x); /* u0d */
  frame(d0xi, NULL, u1d);
  m_mul(u1d, qp->XdinvT, tmp);
  m_mul_add(qp->u0x, qp->XdinvzT, tmp);
  m_mul_at(qp->T, tmp, qp->u1x);
  m_mul(qp->u0x, qp->XdinvT, tmp);
  m_mul_at(qp->T, tmp, qp->u0x);
  qp->detXd *= weight;
}

/* TacsShellAddDispGradSens (SU.h:530-580) */
static void add_dispgrad_sens(const double pt[2], const qp_data *qp, const double du0x[9], const double du1x[9],
                              double res[24], double dd[12]) {
  double du0d[9], du1d[9], tmp[9];
  m_mul_bt(du1x, qp->XdinvzT, tmp);
  m_mul_bt_add(du0x, qp->XdinvT, tmp);
  m_mul(qp->T, tmp, du0d);
  m_mul_bt(du1x, qp->XdinvT, tmp);
  m_mul(qp->T, tmp, du1d);
  double du0xi[6], dd0[3], dd0xi[6];
  extract_frame(du0d, du0xi, dd0);
  extract_frame(du1d, dd0xi, NULL);
  interp3_t(pt, 3, dd0, dd);
  grad3_t(pt, 3, dd0xi, dd);
  if (res) grad3_t(pt, VPN, du0xi, res);
}

/* TacsShellAddDispGradHessian (SU.h:582-647) */
static void add_dispgrad_hess(const double pt[2], const qp_data *qp, const double d2u0x[81],
                              const double d2u1x[81], const double d2u0xu1x[81], double mat[576],
                              double d2d[144], double d2du[144]) {
  const double *T = qp->T, *XT = qp->XdinvT, *XzT = qp->XdinvzT;
  double d2u0d[81], d2u1d[81], d2u0du1d[81];
  hess_tmm(T, T, XT, XT, d2u0x, d2u0d, 0);
  hess_tmm(T, T, XzT, XzT, d2u1x, d2u0d, 1);
  hess_tmm(T, T, XzT, XT, d2u0xu1x, d2u0d, 2);
  hess_tmm(T, T, XT, XT, d2u1x, d2u1d, 0);
  hess_tmm(T, T, XT, XT, d2u0xu1x, d2u0du1d, 0);
  hess_tmm(T, T, XT, XzT, d2u1x, d2u0du1d, 1);

  /* TacsShellExtractFrameSens / ExtractFrameMixedSens (SU.h:97-165):
   * u0d index 3k+c <-> u0xi index 2k+c (c < 2), 3k+2 <-> d0[k] */
  double d2u0xi[36], d2d0[9], d2d0u0xi[18], d2d0xi[36], d2d0xiu0xi[36], d2d0d0xi[18];
  for (int j = 0; j < 6; j++) {
    int jj = 3 * (j / 2) + (j % 2);
    for (int i = 0; i < 6; i++) {
      int ii = 3 * (i / 2) + (i % 2);
      d2u0xi[6 * j + i] = d2u0d[9 * jj + ii];
      d2d0xi[6 * j + i] = d2u1d[9 * jj + ii];
      d2d0xiu0xi[6 * i + j] = d2u0du1d[9 * jj + ii];
    }
  }
  for (int j = 0; j < 3; j++) {
    int jj = 3 * j + 2;
    for (int i = 0; i < 3; i++) d2d0[3 * j + i] = d2u0d[9 * jj + 3 * i + 2];
    for (int i = 0; i < 6; i++) {
      int ii = 3 * (i / 2) + (i % 2);
      d2d0u0xi[6 * j + i] = d2u0d[9 * jj + ii];
      d2d0d0xi[6 * j + i] = d2u0du1d[9 * jj + ii];
    }
  }
  outer_ff(pt, 3, d2d0, d2d);
  outer_gg(pt, 3, d2d0xi, d2d);
  outer_fg(pt, d2d0d0xi, d2d0d0xi, d2d);
  outer_fg(pt, d2d0u0xi, NULL, d2du);
  outer_gg(pt, 3, d2d0xiu0xi, d2du);
  if (mat) outer_gg(pt, VPN, d2u0xi, mat);
}

/* TacsShellAddTyingDispCoupling (SU.h:961-1063) */
static void add_tying_disp_coupling(const double pt[2], const qp_data *qp, const double d2e0tyu0x[54],
                                    const double d2e0tyu1x[54], double d2etyu[108], double d2etyd[108]) {
  const double *T = qp->T, *XT = qp->XdinvT, *XzT = qp->XdinvzT;
  double tmp0d[54], tmp1d[54], d2gtyu0d[54], d2gtyu1d[54];
  for (int k = 0; k < 6; k++) {
    double tmp[9];
    m_mul_bt(&d2e0tyu1x[9 * k], XzT, tmp);
    m_mul_bt_add(&d2e0tyu0x[9 * k], XT, tmp);
    m_mul(T, tmp, &tmp0d[9 * k]);
    m_mul_bt(&d2e0tyu1x[9 * k], XT, tmp);
    m_mul(T, tmp, &tmp1d[9 * k]);
  }
  for (int k = 0; k < 9; k++) {
    double t0[6], out[6], t1[6], out1[6];
    for (int kk = 0; kk < 6; kk++) {
      t0[kk] = tmp0d[9 * kk + k];
      t1[kk] = tmp1d[9 * kk + k];
    }
    sym_tt_sens(XT, t0, out);
    sym_tt_sens(XT, t1, out1);
    for (int kk = 0; kk < 6; kk++) {
      d2gtyu0d[9 * kk + k] = out[kk];
      d2gtyu1d[9 * kk + k] = out1[kk];
    }
  }
  double d2gtyu[72], d2gtyd[72];
  memset(d2gtyu, 0, sizeof(d2gtyu));
  memset(d2gtyd, 0, sizeof(d2gtyd));
  for (int k = 0; k < 6; k++) {
    double du0xi[6], dd0[3], dd0xi[6];
    extract_frame(&d2gtyu0d[9 * k], du0xi, dd0);
    extract_frame(&d2gtyu1d[9 * k], dd0xi, NULL);
    interp3_t(pt, 3, dd0, &d2gtyd[12 * k]);
    grad3_t(pt, 3, dd0xi, &d2gtyd[12 * k]);
    grad3_t(pt, 3, du0xi, &d2gtyu[12 * k]);
  }
  for (int k = 0; k < 12; k++) {
    double t1[6], t2[NTY], t3[6], t4[NTY];
    memset(t2, 0, sizeof(t2));
    memset(t4, 0, sizeof(t4));
    for (int kk = 0; kk < 6; kk++) {
      t1[kk] = d2gtyu[12 * kk + k];
      t3[kk] = d2gtyd[12 * kk + k];
    }
    interp_tying_t(pt, t1, t2);
    interp_tying_t(pt, t3, t4);
    for (int t = 0; t < NTY; t++) {
      d2etyu[12 * t + k] += t2[t];
      d2etyd[12 * t + k] += t4[t];
    }
  }
}

/* derivative of the nodal drilling strain of node n w.r.t. the 24 element
 * dofs (the "drot" rows of TacsShellAddDrillStrainHessian, SU.h:837-890;
 * the same map as TacsShellAddDrillStrainSens, SU.h:787-835) */
static void drill_row(int n, const elem_setup *es, double scale, double out[24]) {
  double pt[2];
  node_point(n, pt);
  const double *T = &es->Tn[9 * n], *XT = &es->XdinvTn[9 * n];
  double du0x[9] = {0, -0.5 * scale, 0, 0.5 * scale, 0, 0, 0, 0, 0};
  double dCt[9] = {0, -0.5 * scale, 0, 0.5 * scale, 0, 0, 0, 0, 0};
  double tmp[9], dC[9], du0d[9], du0xi[6];
  m_mul(T, dCt, tmp);
  m_mul_bt(tmp, T, dC);
  /* addRotationMatResidual (SD.h:74-88) */
  double *r = &out[VPN * n + 3];
  r[0] += -(dC[7] - dC[5]);
  r[1] += -(dC[2] - dC[6]);
  r[2] += -(dC[3] - dC[1]);
  m_mul_bt(du0x, XT, tmp);
  m_mul(T, tmp, du0d);
  extract_frame(du0d, du0xi, NULL);
  grad3_t(pt, VPN, du0xi, out);
}

/* addDirectorJacobian (SD.h:368-501) with beta = gamma = 0, and the
 * residual part of addDirectorResidual (SD.h:354-366) */
static void add_director(const elem_setup *es, const double dd[12], const double *d2d, const double *d2du,
                         double res[24], double *mat) {
  const double *t = es->fn;
  if (mat) {
    for (int i = 0; i < NN; i++)
      for (int j = 0; j < NN; j++) {
        double blk[9], tmp[9];
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) blk[3 * a + b] = d2d[12 * (3 * i + a) + 3 * j + b];
        skew_mat_skew(&t[3 * i], blk, &t[3 * j], tmp);
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) mat[NV * (VPN * i + 3 + a) + VPN * j + 3 + b] -= tmp[3 * a + b];
      }
    for (int i = 0; i < NN; i++)
      for (int j = 0; j < NN; j++) {
        double blk[9], tmp[9];
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) blk[3 * a + b] = d2du[12 * (3 * i + a) + 3 * j + b];
        skew_mat(&t[3 * i], blk, tmp);
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            mat[NV * (VPN * i + a + 3) + VPN * j + b] += tmp[3 * a + b];
            mat[NV * (VPN * j + b) + VPN * i + a + 3] += tmp[3 * a + b];
          }
      }
  }
  if (res) {
    for (int n = 0; n < NN; n++) {
      double c[3];
      cross3(&t[3 * n], &dd[3 * n], c);
      res[VPN * n + 3] += c[0];
      res[VPN * n + 4] += c[1];
      res[VPN * n + 5] += c[2];
    }
  }
}

/* ------------------------------------------------------------------ */
/* addResidual / addJacobian core (SE.h:303-449, 455-672)              */
/* ------------------------------------------------------------------ */
static void elem_core(int model, const or_material *m, double t, double temp, double alpha,
                      const double Xpts[12], const double q[24], double res[24], double *mat) {
  elem_setup es;
  elem_setup_all(model, Xpts, q, &es);
  abd_t c;
  eval_abd(m, t, &c);

  double dd[12], dety[NTY], detn[4];
  memset(dd, 0, sizeof(dd));
  memset(dety, 0, sizeof(dety));
  memset(detn, 0, sizeof(detn));
  double d2d[144], d2du[144], d2ety[81], d2etyu[108], d2etyd[108], d2etn[16];
  if (mat) {
    memset(d2d, 0, sizeof(d2d));
    memset(d2du, 0, sizeof(d2du));
    memset(d2ety, 0, sizeof(d2ety));
    memset(d2etyu, 0, sizeof(d2etyu));
    memset(d2etyd, 0, sizeof(d2etyd));
    memset(d2etn, 0, sizeof(d2etn));
  }

  for (int iq = 0; iq < 4; iq++) {
    double pt[2];
    quad_point(iq, pt);
    qp_data qp;
    qp_kinematics(pt, 1.0, Xpts, q, &es, &qp);

    double gty[6], e0ty[6], e[9], em[9], s[9];
    interp_tying(pt, es.ety, gty);
    sym_tt(qp.XdinvT, gty, e0ty);
    eval_strain(model, qp.u0x, qp.u1x, e0ty, e);
    e[8] = qp.et;
    /* thermal strain (ISC:438-456): alpha*T on the two normal membrane terms */
    for (int i = 0; i < 9; i++) em[i] = e[i];
    em[0] -= m->cte * temp;
    em[1] -= m->cte * temp;
    compute_stress(&c, em, s);

    double du0x[9], du1x[9], de0ty[6];
    strain_sens(model, qp.detXd, s, qp.u0x, qp.u1x, du0x, du1x, de0ty);

    double N[4], Na[4], Nb[4];
    basis_eval(pt, N, Na, Nb);
    for (int n = 0; n < NN; n++) detn[n] += N[n] * (qp.detXd * s[8]);

    add_dispgrad_sens(pt, &qp, du0x, du1x, res, dd);

    double dgty[6];
    sym_tt_sens(qp.XdinvT, de0ty, dgty);
    interp_tying_t(pt, dgty, dety);

    if (mat) {
      double d2u0x[81], d2u1x[81], d2u0xu1x[81], d2e0ty[36], d2e0tyu0x[54], d2e0tyu1x[54];
      strain_hessian(model, alpha * qp.detXd, s, &c, qp.u0x, qp.u1x, d2u0x, d2u1x, d2u0xu1x, d2e0ty,
                     d2e0tyu0x, d2e0tyu1x);
      double d2et = qp.detXd * alpha * c.drill;
      for (int i = 0; i < NN; i++)
        for (int j = 0; j < NN; j++) d2etn[4 * j + i] += N[i] * N[j] * d2et;
      add_dispgrad_hess(pt, &qp, d2u0x, d2u1x, d2u0xu1x, mat, d2d, d2du);
      double d2gty[36];
      sym_tt_hess(qp.XdinvT, d2e0ty, d2gty);
      interp_tying_hess(pt, d2gty, d2ety);
      add_tying_disp_coupling(pt, &qp, d2e0tyu0x, d2e0tyu1x, d2etyu, d2etyd);
    }
  }

  /* drilling strain: residual + Hessian (SU.h:787-959) */
  double drot[4][24];
  for (int n = 0; n < NN; n++) {
    memset(drot[n], 0, sizeof(drot[n]));
    drill_row(n, &es, 1.0, drot[n]);
  }
  for (int n = 0; n < NN; n++)
    for (int j = 0; j < NV; j++) res[j] += detn[n] * drot[n][j];
  if (mat) {
    for (int i = 0; i < NN; i++) {
      double tv[24];
      memset(tv, 0, sizeof(tv));
      for (int j = 0; j < NN; j++)
        for (int k = 0; k < NV; k++) tv[k] += d2etn[i * 4 + j] * drot[j][k];
      for (int a = 0; a < NV; a++)
        for (int b = 0; b < NV; b++) mat[NV * a + b] += tv[a] * drot[i][b];
    }
  }

  tying_transpose(model, &es.td, dety, res, dd);
  if (mat) tying_hessian(model, alpha, &es.td, dety, d2ety, d2etyu, d2etyd, mat, d2d, d2du);
  add_director(&es, dd, mat ? d2d : NULL, mat ? d2du : NULL, res, mat);
}

void or_elem_residual(int model, const or_material *m, double t, double temp, const double X[12],
                      const double q[24], double res[24]) {
  elem_core(model, m, t, temp, 1.0, X, q, res, NULL);
}

void or_elem_jacobian(int model, const or_material *m, double t, double temp, double alpha, const double X[12],
                      const double q[24], double res[24], double mat[576]) {
  elem_core(model, m, t, temp, alpha, X, q, res, mat);
}

void or_elem_stiffness(int model, const or_material *m, double t, double temp, const double X[12],
                       const double q[24], double mat[576]) {
  double res[24];
  memset(res, 0, sizeof(res));
  memset(mat, 0, 576 * sizeof(double));
  elem_core(model, m, t, temp, 1.0, X, q, res, mat);
}

/* getOutputData for TACS_BEAM_OR_SHELL_ELEMENT (SE.h:1402-1500) with the
 * strain, stress and extra groups (TACSToFH5's element zone): at each of
 * the 4 nodes, the strain e[9] (u0x, u1x and the tying strain interpolated at
 * the node, e[8] = drilling strain there), the stress s = C e of
 * TACSIsoShellConstitutive::evalStress (ISC:153-189, no thermal strain),
 * failure0 = evalFailure (ISC:254-297: KS, weight 100, of the von Mises
 * values of the top and bottom surface stresses, TACSMaterialProperties.cpp:
 * 453-457), failure1..5 = 0, dv1 = t, dv2..7 = 0 (ISC:503-512).
 * out[31 * node + k].
 * pinned by: golden vectors of the compiled reference getOutputData */
#ifndef OR_NO_EXTERN_C /* not part of the operation-counting build (flopcount.cpp) */
void or_elem_output(int model, const or_material *m, double t, double ys, const double Xpts[12],
                    const double q[24], double out[124]) {
  elem_setup es;
  elem_setup_all(model, Xpts, q, &es);
  abd_t c;
  eval_abd(m, t, &c);
  double Q[6];
  tangent_2d(m, Q);
  const double ksWeight = 100.0, tOffset = 0.0;
  for (int n = 0; n < NN; n++) {
    double pt[2];
    node_point(n, pt);
    qp_data qp;
    qp_kinematics(pt, 1.0, Xpts, q, &es, &qp);
    double gty[6], e0ty[6], e[9], sg[9];
    interp_tying(pt, es.ety, gty);
    sym_tt(qp.XdinvT, gty, e0ty);
    eval_strain(model, qp.u0x, qp.u1x, e0ty, e);
    e[8] = qp.et;
    compute_stress(&c, e, sg);
    /* evalFailure: top and bottom surface strains and stresses */
    const double ht = (0.5 - tOffset) * t, hb = -(0.5 + tOffset) * t;
    double et[3], eb[3], st[3], sb[3];
    for (int i = 0; i < 3; i++) {
      et[i] = e[i] + ht * e[3 + i];
      eb[i] = e[i] + hb * e[3 + i];
    }
    st[0] = Q[0] * et[0] + Q[1] * et[1] + Q[2] * et[2];
    st[1] = Q[1] * et[0] + Q[3] * et[1] + Q[4] * et[2];
    st[2] = Q[2] * et[0] + Q[4] * et[1] + Q[5] * et[2];
    sb[0] = Q[0] * eb[0] + Q[1] * eb[1] + Q[2] * eb[2];
    sb[1] = Q[1] * eb[0] + Q[3] * eb[1] + Q[4] * eb[2];
    sb[2] = Q[2] * eb[0] + Q[4] * eb[1] + Q[5] * eb[2];
    const double top = sqrt(st[0] * st[0] + st[1] * st[1] - st[0] * st[1] + 3.0 * st[2] * st[2]) / ys;
    const double bottom = sqrt(sb[0] * sb[0] + sb[1] * sb[1] - sb[0] * sb[1] + 3.0 * sb[2] * sb[2]) / ys;
    const double ksMax = top > bottom ? top : bottom;
    const double ksSum = exp(ksWeight * (top - ksMax)) + exp(ksWeight * (bottom - ksMax));
    double *o = &out[31 * n];
    for (int i = 0; i < 9; i++) {
      o[i] = e[i];
      o[9 + i] = sg[i];
    }
    for (int i = 0; i < 13; i++) o[18 + i] = 0.0;
    o[18] = ksMax + log(ksSum) / ksWeight;
    o[24] = t;
  }
}
#endif

/* getMatType(TACS_MASS_MATRIX) (SE.h:675-704 with gamma = 1, alpha = beta =
 * 0): the inertial terms of addJacobian (SE.h:613-648) -- the translational
 * block gamma detXd m0 I through addInterpFieldsOuterProduct and the
 * director block detXd m2 I accumulated in d2Tdotd -- and the gamma d2Tdotd
 * part of addDirectorJacobian for TACSLinearizedRotation (SD.h:368-430).
 * With alpha = 0 every stiffness Hessian (d2d, d2du, tying and drilling
 * terms) is scaled to zero.  Mass moments of TACSIsoShellConstitutive
 * (ISC:120-129) with tOffset = 0: m0 = rho t, m1 = -rho t^2 tOffset = 0 (so
 * d2Tdotu vanishes), m2 = rho t^3 (tOffset^2 + 1/12).  mat is overwritten.
 * pinned by: golden vectors of the compiled reference getMatType(MASS);
 * rigid translation / in-plane rotation energies (closed forms) */
void or_elem_mass(const or_material *m, double t, const double Xpts[12], double mat[576]) {
  memset(mat, 0, 576 * sizeof(double));
  double fn[12], Xdn[36];
  node_normals(Xpts, fn, Xdn);
  const double gamma = 1.0, tOffset = 0.0;
  const double m0 = m->rho * t;
  const double m2 = m->rho * t * t * t * (tOffset * tOffset + 1.0 / 12.0);
  double d2Tdotd[144];
  memset(d2Tdotd, 0, sizeof(d2Tdotd));
  for (int iq = 0; iq < 4; iq++) {
    double pt[2], Xxi[6], n0[3], Xd[9], Xdinv[9];
    quad_point(iq, pt);
    grad3(pt, 3, Xpts, Xxi);
    interp3(pt, 3, fn, n0);
    frame(Xxi, n0, Xd);
    const double detXd = m_inv(Xd, Xdinv) * 1.0; /* Gauss weight 1 */
    double d2u0dot[9], d2Td[9];
    memset(d2u0dot, 0, sizeof(d2u0dot));
    d2u0dot[0] = d2u0dot[4] = d2u0dot[8] = gamma * detXd * m0;
    outer_ff(pt, VPN, d2u0dot, mat);
    memset(d2Td, 0, sizeof(d2Td));
    d2Td[0] = d2Td[4] = d2Td[8] = detXd * m2;
    outer_ff(pt, 3, d2Td, d2Tdotd);
  }
  /* d = q x t: mat[q_i, q_j] -= [t_i]x (gamma d2Tdotd_ij) [t_j]x */
  for (int i = 0; i < NN; i++)
    for (int j = 0; j < NN; j++) {
      double blk[9], tmp[9];
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) blk[3 * a + b] = gamma * d2Tdotd[12 * (3 * i + a) + 3 * j + b];
      skew_mat_skew(&fn[3 * i], blk, &fn[3 * j], tmp);
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) mat[NV * (VPN * i + 3 + a) + VPN * j + 3 + b] -= tmp[3 * a + b];
    }
}

/* getMatType(TACS_GEOMETRIC_STIFFNESS_MATRIX) (SE.h:705-766): central
 * difference of the nonlinear tangent along the state with the temperature
 * perturbed along with it.  The reference's step is dh = 1e-4.  Because the
 * nonlinear tangent is an exact quadratic polynomial of the state (strains
 * at most quadratic in q), the same formula with any step returns the same
 * matrix up to round-off; a unit step avoids the cancellation of the 1e-4
 * step and is used as the accurate reference value (or_elem_geometric_step). */
void or_elem_geometric_step(const or_material *m, double t, double temp, const double X[12], const double u[24],
                            double dh, double mat[576]) {
  double res[24], path[24];
  memset(mat, 0, 576 * sizeof(double));
  memset(res, 0, sizeof(res));
  double norm = 0.0;
  for (int i = 0; i < NV; i++) norm += u[i] * u[i];
  norm += temp * temp;
  if (norm == 0.0) {
    norm = 1.0;
  } else {
    norm = sqrt(norm);
  }
  double alpha = 0.5 * norm / dh;
  for (int i = 0; i < NV; i++) path[i] = dh * u[i] / norm;
  elem_core(OR_NONLINEAR, m, t, temp + dh * temp / norm, alpha, X, path, res, mat);
  for (int i = 0; i < NV; i++) path[i] = -dh * u[i] / norm;
  elem_core(OR_NONLINEAR, m, t, temp - dh * temp / norm, -alpha, X, path, res, mat);
}

void or_elem_geometric_fd(const or_material *m, double t, double temp, const double X[12], const double u[24],
                          double mat[576]) {
  or_elem_geometric_step(m, t, temp, X, u, 1e-4, mat);
}

/* addAdjResProduct (SE.h:774-868) + addStressDVSens (ISC:229-251) for the
 * static case.  Two deliberate departures, both pinned by the 5-point
 * thickness derivative of the reference's own residual:
 *  - the drilling-strain derivative etnd is computed; the reference's
 *    TACSLinearizedRotation::evalDrillStrainDeriv (SD.h:584-590) never
 *    writes it, leaving an uninitialised value;
 *  - the mechanical strain e - e_th is used, as in the residual; the
 *    reference passes e (SE.h:848-849), which differs only when T != 0. */
void or_elem_adj_res_dv(int model, const or_material *m, double t, double temp, double scale,
                        const double psi[24], const double Xpts[12], const double q[24], double *dfdt) {
  elem_setup es;
  elem_setup_all(model, Xpts, q, &es);
  /* forward derivatives along psi */
  double etnd[4], dd[12], etyd[NTY];
  for (int n = 0; n < NN; n++) {
    double row[24];
    memset(row, 0, sizeof(row));
    drill_row(n, &es, 1.0, row);
    double v = 0.0;
    for (int k = 0; k < NV; k++) v += row[k] * psi[k];
    etnd[n] = v;
  }
  for (int n = 0; n < NN; n++) cross3(&psi[VPN * n + 3], &es.fn[3 * n], &dd[3 * n]);
  {
    tying_data tdd = es.td;
    for (int tp = 0; tp < NTY; tp++) {
      grad3(TY_PT[tp], VPN, psi, tdd.Uxi[tp]);
      interp3(TY_PT[tp], 3, dd, tdd.d0[tp]);
    }
    for (int tp = 0; tp < NTY; tp++) {
      double dU[6], dd0[3];
      tying_eval(model, tp, &es.td, dU, dd0);
      double v = 0.0;
      for (int k = 0; k < 6; k++) v += dU[k] * tdd.Uxi[tp][k];
      if (TY_FIELD[tp] == F23 || TY_FIELD[tp] == F13)
        for (int k = 0; k < 3; k++) v += dd0[k] * tdd.d0[tp][k];
      etyd[tp] = v;
    }
  }
  double A2[6];
  tangent_2d(m, A2);
  double dI = 0.25 * t * t;
  for (int iq = 0; iq < 4; iq++) {
    double pt[2];
    quad_point(iq, pt);
    qp_data qp;
    qp_kinematics(pt, 1.0, Xpts, q, &es, &qp);
    /* directional derivatives of u0x, u1x (TacsShellComputeDispGradDeriv, SU.h:444-528) */
    double d0d[3], d0xid[6], u0xid[6], u0xd[9], u1xd[9], tmp[9], u1dd[9];
    interp3(pt, 3, dd, d0d);
    grad3(pt, 3, dd, d0xid);
    grad3(pt, VPN, psi, u0xid);
    frame(u0xid, d0d, u0xd);
    frame(d0xid, NULL, u1dd);
    m_mul(u1dd, qp.XdinvT, tmp);
    m_mul_add(u0xd, qp.XdinvzT, tmp);
    m_mul_at(qp.T, tmp, u1xd);
    m_mul(u0xd, qp.XdinvT, tmp);
    m_mul_at(qp.T, tmp, u0xd);

    double gty[6], gtyd[6], e0ty[6], e0tyd[6], e[9], ed[9];
    interp_tying(pt, es.ety, gty);
    interp_tying(pt, etyd, gtyd);
    sym_tt(qp.XdinvT, gty, e0ty);
    sym_tt(qp.XdinvT, gtyd, e0tyd);
    eval_strain(model, qp.u0x, qp.u1x, e0ty, e);
    strain_deriv(model, qp.u0x, qp.u1x, u0xd, u1xd, e0tyd, ed);
    e[8] = qp.et;
    ed[8] = interp1(pt, etnd);
    e[0] -= m->cte * temp;
    e[1] -= m->cte * temp;
    double sc = scale * qp.detXd;
    *dfdt += sc * (sym_inner(A2, &ed[0], &e[0]) + dI * sym_inner(A2, &ed[3], &e[3]) +
                   (5.0 / 6.0) * A2[5] * (ed[6] * e[6] + ed[7] * e[7] + m->k_drill * ed[8] * e[8]));
  }
}

/* computeEnergies (SE.h:202-297), potential part */
double or_elem_energy(int model, const or_material *m, double t, double temp, const double Xpts[12],
                      const double q[24]) {
  elem_setup es;
  elem_setup_all(model, Xpts, q, &es);
  abd_t c;
  eval_abd(m, t, &c);
  double U = 0.0;
  for (int iq = 0; iq < 4; iq++) {
    double pt[2];
    quad_point(iq, pt);
    qp_data qp;
    qp_kinematics(pt, 1.0, Xpts, q, &es, &qp);
    double gty[6], e0ty[6], e[9], s[9];
    interp_tying(pt, es.ety, gty);
    sym_tt(qp.XdinvT, gty, e0ty);
    eval_strain(model, qp.u0x, qp.u1x, e0ty, e);
    e[8] = qp.et;
    compute_stress(&c, e, s);
    double se = 0.0;
    for (int i = 0; i < 9; i++) se += s[i] * e[i];
    U += 0.5 * qp.detXd * se;
  }
  return U;
}
