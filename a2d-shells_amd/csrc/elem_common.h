// elem_common.h -- device helpers for the 4-node MITC director shell
// (TACSQuad4Shell / TACSQuad4NonlinearShell with TACSLinearizedRotation,
// TACSShellNaturalTransform and TACSIsoShellConstitutive, tOffset = 0).
//
// Reference semantics are cited per helper (paths relative to the reference
// repository, src/elements/shell/...).
#pragma once

#include <hip/hip_runtime.h>

namespace shellgpu {

// TACSGaussQuadrature.h:26 -- the reference's truncated Gauss literal
constexpr double kGauss2 = 0.577350269189626;

// Bilinear Lagrange basis, node n = i + 2j at (-1 + 2i, -1 + 2j)
// (TACSShellElementQuadBasis.h:119-230, order 2).
__device__ __forceinline__ void shape4(double u, double v, double N[4],
                                       double Na[4], double Nb[4]) {
  const double na0 = 0.5 * (1.0 - u), na1 = 0.5 * (1.0 + u);
  const double nb0 = 0.5 * (1.0 - v), nb1 = 0.5 * (1.0 + v);
  N[0] = na0 * nb0;
  N[1] = na1 * nb0;
  N[2] = na0 * nb1;
  N[3] = na1 * nb1;
  Na[0] = -0.5 * nb0;
  Na[1] = 0.5 * nb0;
  Na[2] = -0.5 * nb1;
  Na[3] = 0.5 * nb1;
  Nb[0] = na0 * -0.5;
  Nb[1] = na1 * -0.5;
  Nb[2] = na0 * 0.5;
  Nb[3] = na1 * 0.5;
}

__device__ __forceinline__ void node_point(int n, double &u, double &v) {
  u = -1.0 + 2.0 * (n & 1);
  v = -1.0 + 2.0 * (n >> 1);
}

__device__ __forceinline__ void quad_point(int n, double &u, double &v) {
  u = (n & 1) ? kGauss2 : -kGauss2;
  v = (n >> 1) ? kGauss2 : -kGauss2;
}

// MITC tying points (TACSShellElementQuadBasis.h:463-565): 0,1 g11 at
// (0,-1),(0,1); 2,3 g22 at (-1,0),(1,0); 4 g12 at (0,0); 5,6 g23 at
// (-1,0),(1,0); 7,8 g13 at (0,-1),(0,1).
enum TyingField { kG11 = 0, kG22 = 1, kG12 = 2, kG23 = 3, kG13 = 4 };

__device__ __forceinline__ int tying_field(int t) {
  return t < 2 ? kG11 : (t < 4 ? kG22 : (t == 4 ? kG12 : (t < 7 ? kG23 : kG13)));
}

__device__ __forceinline__ void tying_point(int t, double &u, double &v) {
  const int f = tying_field(t);
  const double s = (t == 0 || t == 2 || t == 5 || t == 7) ? -1.0 : 1.0;
  if (f == kG11 || f == kG13) {
    u = 0.0;
    v = s;
  } else if (f == kG12) {
    u = 0.0;
    v = 0.0;
  } else {
    u = s;
    v = 0.0;
  }
}

// Interpolation weights of the 9 tying values at (u, v)
// (TACSShellElementQuadBasis.h:568-607).
__device__ __forceinline__ void tying_interp(double u, double v, double N[9]) {
  const double na0 = (u - 1.0) * (-0.5), na1 = (u + 1.0) * 0.5;
  const double nb0 = (v - 1.0) * (-0.5), nb1 = (v + 1.0) * 0.5;
  N[0] = nb0;
  N[1] = nb1;
  N[2] = na0;
  N[3] = na1;
  N[4] = 1.0;
  N[5] = na0;
  N[6] = na1;
  N[7] = nb0;
  N[8] = nb1;
}

// slot of each tying field in the symmetric 3x3 gty = [g11 g12 g13 g22 g23 g33]
__device__ __forceinline__ int field_slot(int f) {
  return f == kG11 ? 0 : (f == kG22 ? 3 : (f == kG12 ? 1 : (f == kG23 ? 4 : 2)));
}

// ---- 3x3 row-major algebra, generic in the scalar type S (double, or the
// forward-mode Dual of dual.h for coordinate derivatives) ----
template <typename S>
__device__ __forceinline__ void cross3(const S *a, const S *b, S *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
template <typename S>
__device__ __forceinline__ S dot3(const S *a, const S *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
template <typename S>
__device__ __forceinline__ void mm3(const S *A, const S *B, S *C) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      C[3 * i + j] =
          A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
// C = A^T B
template <typename S>
__device__ __forceinline__ void mtm3(const S *A, const S *B, S *C) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      C[3 * i + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
}
// inverse by cofactors; determinant expanded on the third row
// (TACSElementAlgebra.h:1980-2009)
template <typename S>
__device__ __forceinline__ S inv3(const S *A, S *Ai) {
  const S det =
      (A[8] * (A[0] * A[4] - A[3] * A[1]) - A[7] * (A[0] * A[5] - A[3] * A[2]) +
       A[6] * (A[1] * A[5] - A[2] * A[4]));
  const S r = 1.0 / det;
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

// TACSShellNaturalTransform::computeTransform
// (TACSShellElementTransform.h:21-93).  The in-plane axis is deliberately
// left un-orthogonalised: the reference subtracts d*n[0] from t1[0] three
// times and never touches t1[1], t1[2] (lines 42-44).  T = [t1 t2 n] as
// columns.
template <typename S>
__device__ __forceinline__ void natural_transform(const S Xxi[6], const S n0[3],
                                                  S T[9]) {
  S n[3] = {n0[0], n0[1], n0[2]};
  S inv = 1.0 / sqrt(dot3(n, n));
  n[0] = n[0] * inv;
  n[1] = n[1] * inv;
  n[2] = n[2] * inv;
  S t1[3] = {Xxi[0], Xxi[2], Xxi[4]};
  const S d = dot3(n, t1);
  t1[0] = t1[0] - d * n[0];
  t1[0] = t1[0] - d * n[0];
  t1[0] = t1[0] - d * n[0];
  inv = 1.0 / sqrt(dot3(t1, t1));
  t1[0] = t1[0] * inv;
  t1[1] = t1[1] * inv;
  t1[2] = t1[2] * inv;
  S t2[3];
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

// 3x3 frame [xi-derivatives (interleaved) | c] (TACSShellUtilities.h:8-35)
template <typename S>
__device__ __forceinline__ void frame3(const S xi[6], const S c[3], S Xd[9]) {
  Xd[0] = xi[0];
  Xd[1] = xi[1];
  Xd[2] = c[0];
  Xd[3] = xi[2];
  Xd[4] = xi[3];
  Xd[5] = c[1];
  Xd[6] = xi[4];
  Xd[7] = xi[5];
  Xd[8] = c[2];
}
// the same frame with a zero third column
template <typename S>
__device__ __forceinline__ void frame3z(const S xi[6], S Xd[9]) {
  Xd[0] = xi[0];
  Xd[1] = xi[1];
  Xd[2] = 0.0;
  Xd[3] = xi[2];
  Xd[4] = xi[3];
  Xd[5] = 0.0;
  Xd[6] = xi[4];
  Xd[7] = xi[5];
  Xd[8] = 0.0;
}

}  // namespace shellgpu
