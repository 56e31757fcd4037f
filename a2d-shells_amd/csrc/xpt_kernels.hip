// xpt_kernels.hip -- nodal-coordinate derivatives of the element scalars of
// the buckling sensitivities:
//
//   s_K = phi^T K_e phi,   s_G = phi^T G_e(u) phi,   s_R = psi^T R_e(u)
//
// The reference differentiates these products with respect to Xpts by
// central differences (TACSElement::addMatXptSensInnerProduct and
// addAdjResXptProduct, src/elements/TACSElement.cpp:270-332, 434-507,
// dh = 1e-7).  Here they are exact: one thread per (element, coordinate)
// evaluates the scalars in forward-mode dual numbers seeded on that
// coordinate.  The scalars are written through strains only,
//
//   e(q) = e1(q) + Q(q, q) / 2           (Q symmetric bilinear)
//   s_K = sum_qp det e1(phi)^T C e1(phi)
//   s_G = sum_qp det [2 Q(u, phi)^T C e1(phi) + (e1(u) - e_th)^T C Q(phi, phi)]
//   s_R = sum_qp det e1(psi)^T C (e1(u) - e_th)                  (linear model)
//       = sum_qp det (e1(psi) + Q(u, psi))^T C (e(u) - e_th)     (nonlinear)
//   s_M = phi^T M_e phi = sum_qp det (m0 |u0(phi)|^2 + m2 |d0(phi)|^2)
//
// The last is the consistent mass of the linearized-rotation shell
// (TACSShellElement.h:613-648 with the moments of
// TACSIsoShellConstitutive::evalMassMoments, m0 = rho t, m2 = rho t^3/12):
// its rotational block -[t_i]x S_ij [t_j]x contracts with phi to
// S_ij (q_i x t_i).(q_j x t_j), and the interpolated director is
// d0 = sum N_m (q_m x t_m).
//
// with the strain evaluation of the element kernels: node normals and
// drilling strains (TACSShellUtilities.h:301-343, 649-707), MITC tying
// strains (TACSShellElementModel.h:32-77, 643-697) interpolated and
// transformed at the quadrature points, bending strains from the
// displacement gradients (TACSShellUtilities.h:361-422) and the natural
// transform with its un-orthogonalised in-plane axis.  Elements run colour
// by colour, so the 12 threads of an element add into distinct coordinates
// and no two elements of a launch share a node.
#include "dual.h"
#include "elem_common.h"
#include "shellgpu_internal.h"

namespace shellgpu {
namespace {

constexpr int kThreads = 192;  // 16 elements x 12 coordinates

template <typename S>
__device__ __forceinline__ void apply_cs(const Section &c, const S &w, const S b[9], S y[9]) {
  y[0] = w * (c.A[0] * b[0] + c.A[1] * b[1] + c.A[2] * b[2]);
  y[1] = w * (c.A[1] * b[0] + c.A[3] * b[1] + c.A[4] * b[2]);
  y[2] = w * (c.A[2] * b[0] + c.A[4] * b[1] + c.A[5] * b[2]);
  y[3] = w * (c.D[0] * b[3] + c.D[1] * b[4] + c.D[2] * b[5]);
  y[4] = w * (c.D[1] * b[3] + c.D[3] * b[4] + c.D[4] * b[5]);
  y[5] = w * (c.D[2] * b[3] + c.D[4] * b[4] + c.D[5] * b[5]);
  y[6] = w * (c.As[0] * b[6] + c.As[1] * b[7]);
  y[7] = w * (c.As[1] * b[6] + c.As[2] * b[7]);
  y[8] = w * (c.drill * b[8]);
}

template <typename S>
__device__ __forceinline__ S dot9(const S a[9], const S b[9]) {
  S s = a[0] * b[0];
#pragma unroll
  for (int i = 1; i < 9; i++) s = s + a[i] * b[i];
  return s;
}

// linear tying strain of a field: U = interleaved xi/eta displacement
// gradient, d0 = director, X = interleaved xi/eta coordinate gradient,
// n0 = interpolated normal (TACSShellElementModel.h:32-77)
template <typename S>
__device__ __forceinline__ S tying_lin(int f, const S U[6], const S d0[3], const S X[6], const S n0[3]) {
  if (f == kG11) return U[0] * X[0] + U[2] * X[2] + U[4] * X[4];
  if (f == kG22) return U[1] * X[1] + U[3] * X[3] + U[5] * X[5];
  if (f == kG12)
    return 0.5 * (U[0] * X[1] + U[2] * X[3] + U[4] * X[5] + U[1] * X[0] + U[3] * X[2] + U[5] * X[4]);
  if (f == kG23) return 0.5 * (X[1] * d0[0] + X[3] * d0[1] + X[5] * d0[2] + n0[0] * U[1] + n0[1] * U[3] + n0[2] * U[5]);
  return 0.5 * (X[0] * d0[0] + X[2] * d0[1] + X[4] * d0[2] + n0[0] * U[0] + n0[1] * U[2] + n0[2] * U[4]);
}

// symmetric bilinear form of the quadratic tying terms
// (TACSShellElementModel.h:643-697): the nonlinear tying strain is
// lin(q) + Q(q, q) / 2
template <typename S>
__device__ __forceinline__ S tying_quad(int f, const S Ua[6], const S da[3], const S Ub[6], const S db[3]) {
  if (f == kG11) return Ua[0] * Ub[0] + Ua[2] * Ub[2] + Ua[4] * Ub[4];
  if (f == kG22) return Ua[1] * Ub[1] + Ua[3] * Ub[3] + Ua[5] * Ub[5];
  if (f == kG12)
    return 0.5 * (Ua[0] * Ub[1] + Ua[2] * Ub[3] + Ua[4] * Ub[5] + Ub[0] * Ua[1] + Ub[2] * Ua[3] + Ub[4] * Ua[5]);
  if (f == kG23)
    return 0.5 * (da[0] * Ub[1] + da[1] * Ub[3] + da[2] * Ub[5] + db[0] * Ua[1] + db[1] * Ua[3] + db[2] * Ua[5]);
  return 0.5 * (da[0] * Ub[0] + da[1] * Ub[2] + da[2] * Ub[4] + db[0] * Ua[0] + db[1] * Ua[2] + db[2] * Ua[4]);
}

// vectors: 0 phi, 1 u, 2 psi; bilinear pairs (u,phi), (phi,phi), (u,psi), (u,u)
constexpr int kPairA[4] = {1, 0, 1, 1};
constexpr int kPairB[4] = {0, 0, 2, 1};

// tying strains of the needed vectors and pairs at tying point t
// (TACSShellElementModel.h:32-77 linear, :643-697 quadratic)
template <typename S, int WM, int PM>
__device__ __forceinline__ void tying_at(int t, const S X[12], const S fn[12], const double q[3][24], S tl[3],
                                         S tq[4]) {
  double u, v, N[4], Na[4], Nb[4];
  tying_point(t, u, v);
  shape4(u, v, N, Na, Nb);
  S Xxi[6], n0[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < 6; i++) Xxi[i] = 0.0;
  for (int m = 0; m < 4; m++)
    for (int d = 0; d < 3; d++) {
      Xxi[2 * d] = Xxi[2 * d] + Na[m] * X[3 * m + d];
      Xxi[2 * d + 1] = Xxi[2 * d + 1] + Nb[m] * X[3 * m + d];
      n0[d] = n0[d] + N[m] * fn[3 * m + d];
    }
  S U[3][6], d0[3][3];
  for (int w = 0; w < 3; w++) {
    if (!((WM >> w) & 1)) continue;
    for (int i = 0; i < 6; i++) U[w][i] = 0.0;
    for (int i = 0; i < 3; i++) d0[w][i] = 0.0;
    for (int m = 0; m < 4; m++) {
      const S qr[3] = {q[w][6 * m + 3], q[w][6 * m + 4], q[w][6 * m + 5]};
      S dm[3];
      cross3(qr, &fn[3 * m], dm);
      for (int d = 0; d < 3; d++) {
        d0[w][d] = d0[w][d] + N[m] * dm[d];
        U[w][2 * d] = U[w][2 * d] + Na[m] * q[w][6 * m + d];
        U[w][2 * d + 1] = U[w][2 * d + 1] + Nb[m] * q[w][6 * m + d];
      }
    }
  }
  const int f = tying_field(t);
  for (int w = 0; w < 3; w++)
    if ((WM >> w) & 1) tl[w] = tying_lin(f, U[w], d0[w], Xxi, n0);
  for (int p = 0; p < 4; p++)
    if ((PM >> p) & 1) tq[p] = tying_quad(f, U[kPairA[p]], d0[kPairA[p]], U[kPairB[p]], d0[kPairB[p]]);
}

// WM: the vectors a variant evaluates (bit w), PM: the pairs (bit p).  Each
// scalar needs: s_K and s_M phi; s_G phi, u and the pairs (u,phi),
// (phi,phi); s_R u, psi and, for the nonlinear model, (u,psi), (u,u).
// Vectors and pairs outside the masks are never formed, which keeps the
// dual-number strains of a variant in registers.
template <typename S, int WM, int PM>
__device__ void element_scalars(const S X[12], const double q[3][24], const Section &sec, double eth, bool nonlin,
                                const double c[4], double m0, double m2, S &out) {
  constexpr bool kPhi = WM & 1, kU = (WM >> 1) & 1, kPsi = (WM >> 2) & 1;
  constexpr bool kGeo = kPhi && kU && (PM & 3) == 3;
  constexpr bool kRes = kU && kPsi;
  constexpr bool kResNl = kRes && (PM & 12) == 12;
  // ---- node normals ----
  S fn[12];
  for (int n = 0; n < 4; n++) {
    double u, v, N[4], Na[4], Nb[4];
    node_point(n, u, v);
    shape4(u, v, N, Na, Nb);
    S a3[3] = {0.0, 0.0, 0.0}, b3[3] = {0.0, 0.0, 0.0};
    for (int m = 0; m < 4; m++)
      for (int d = 0; d < 3; d++) {
        a3[d] = a3[d] + Na[m] * X[3 * m + d];
        b3[d] = b3[d] + Nb[m] * X[3 * m + d];
      }
    S f[3];
    cross3(a3, b3, f);
    const S nrm = sqrt(dot3(f, f));
    if (value(nrm) != 0.0) {
      const S inv = 1.0 / nrm;
      f[0] = f[0] * inv;
      f[1] = f[1] * inv;
      f[2] = f[2] * inv;
    }
    fn[3 * n] = f[0];
    fn[3 * n + 1] = f[1];
    fn[3 * n + 2] = f[2];
  }
  // ---- drilling strain at the nodes (TacsShellComputeDrillStrain with
  // TACSLinearizedRotation: et = 0.5 (Ct[3] + u0x[3] - Ct[1] - u0x[1])) ----
  S etn[3][4];
  for (int n = 0; n < 4; n++) {
    double u, v, N[4], Na[4], Nb[4];
    node_point(n, u, v);
    shape4(u, v, N, Na, Nb);
    S Xxi[6];
    for (int i = 0; i < 6; i++) Xxi[i] = 0.0;
    for (int m = 0; m < 4; m++)
      for (int d = 0; d < 3; d++) {
        Xxi[2 * d] = Xxi[2 * d] + Na[m] * X[3 * m + d];
        Xxi[2 * d + 1] = Xxi[2 * d + 1] + Nb[m] * X[3 * m + d];
      }
    S Tn[9], Xd[9], Xdi[9], XTn[9];
    natural_transform(Xxi, &fn[3 * n], Tn);
    frame3(Xxi, &fn[3 * n], Xd);
    inv3(Xd, Xdi);
    mm3(Xdi, Tn, XTn);
    // dC = Tn dA Tn^T, du0d = Tn dA XTn^T with dA = d et / d(Ct or u0x)
    const double dA[9] = {0, -0.5, 0, 0.5, 0, 0, 0, 0, 0};
    S tmp[9], dC[9], du0d[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        tmp[3 * i + j] = dA[3 * i] * Tn[3 * j] + dA[3 * i + 1] * Tn[3 * j + 1] + dA[3 * i + 2] * Tn[3 * j + 2];
    mm3(Tn, tmp, dC);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        tmp[3 * i + j] = dA[3 * i] * XTn[3 * j] + dA[3 * i + 1] * XTn[3 * j + 1] + dA[3 * i + 2] * XTn[3 * j + 2];
    mm3(Tn, tmp, du0d);
    const S rot[3] = {dC[5] - dC[7], dC[6] - dC[2], dC[1] - dC[3]};
    const S g[6] = {du0d[0], du0d[1], du0d[3], du0d[4], du0d[6], du0d[7]};
    for (int w = 0; w < 3; w++) {
      if (!((WM >> w) & 1)) continue;
      S et = rot[0] * q[w][6 * n + 3] + rot[1] * q[w][6 * n + 4] + rot[2] * q[w][6 * n + 5];
      for (int m = 0; m < 4; m++)
        for (int d = 0; d < 3; d++) et = et + (Na[m] * g[2 * d] + Nb[m] * g[2 * d + 1]) * q[w][6 * m + d];
      etn[w][n] = et;
    }
  }
  // ---- tying strains: linear parts of the vectors and the bilinear
  // quadratic terms of the pairs.  Variants with pairs evaluate them again at
  // every quadrature point instead of holding 9 points' worth of dual
  // numbers through the quadrature loop (which spilled to scratch memory) ----
  constexpr bool kLate = PM != 0;
  S tyl[3][kLate ? 1 : 9];
  if constexpr (!kLate) {
    for (int t = 0; t < 9; t++) {
      S tl[3], tq[4];
      tying_at<S, WM, PM>(t, X, fn, q, tl, tq);
      for (int w = 0; w < 3; w++)
        if ((WM >> w) & 1) tyl[w][t] = tl[w];
    }
  }
  // ---- quadrature points ----
  S acc = 0.0;
  // rolled: an unrolled loop keeps four points' dual numbers live at once
#pragma unroll 1
  for (int iq = 0; iq < 4; iq++) {
    double u, v, N[4], Na[4], Nb[4];
    quad_point(iq, u, v);
    shape4(u, v, N, Na, Nb);
    S Xxi[6], nxi[6], n0[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 6; i++) Xxi[i] = nxi[i] = 0.0;
    for (int m = 0; m < 4; m++)
      for (int d = 0; d < 3; d++) {
        Xxi[2 * d] = Xxi[2 * d] + Na[m] * X[3 * m + d];
        Xxi[2 * d + 1] = Xxi[2 * d + 1] + Nb[m] * X[3 * m + d];
        nxi[2 * d] = nxi[2 * d] + Na[m] * fn[3 * m + d];
        nxi[2 * d + 1] = nxi[2 * d + 1] + Nb[m] * fn[3 * m + d];
        n0[d] = n0[d] + N[m] * fn[3 * m + d];
      }
    S T[9], Xd[9], Xdz[9], Xdi[9], XT[9], nXdz[9], XzT[9];
    natural_transform(Xxi, n0, T);
    frame3(Xxi, n0, Xd);
    frame3z(nxi, Xdz);
    const S det = inv3(Xd, Xdi);
    mm3(Xdi, Xdz, nXdz);
    for (int i = 0; i < 9; i++) nXdz[i] = -nXdz[i];
    mm3(Xdi, T, XT);
    mm3(nXdz, XT, XzT);
    // displacement gradients u0x, u1x of the three vectors
    S U0[3][9], U1[3][9];
    for (int w = 0; w < 3; w++) {
      if (!((WM >> w) & 1)) continue;
      S d0[3] = {0.0, 0.0, 0.0}, d0xi[6], u0xi[6];
      for (int i = 0; i < 6; i++) d0xi[i] = u0xi[i] = 0.0;
      for (int m = 0; m < 4; m++) {
        const S qr[3] = {q[w][6 * m + 3], q[w][6 * m + 4], q[w][6 * m + 5]};
        S dm[3];
        cross3(qr, &fn[3 * m], dm);
        for (int d = 0; d < 3; d++) {
          d0[d] = d0[d] + N[m] * dm[d];
          d0xi[2 * d] = d0xi[2 * d] + Na[m] * dm[d];
          d0xi[2 * d + 1] = d0xi[2 * d + 1] + Nb[m] * dm[d];
          u0xi[2 * d] = u0xi[2 * d] + Na[m] * q[w][6 * m + d];
          u0xi[2 * d + 1] = u0xi[2 * d + 1] + Nb[m] * q[w][6 * m + d];
        }
      }
      S u0d[9], u1d[9], tmp[9], tmp2[9];
      frame3(u0xi, d0, u0d);
      frame3z(d0xi, u1d);
      mm3(u1d, XT, tmp);
      mm3(u0d, XzT, tmp2);
      for (int i = 0; i < 9; i++) tmp[i] = tmp[i] + tmp2[i];
      mtm3(T, tmp, U1[w]);
      mm3(u0d, XT, tmp);
      mtm3(T, tmp, U0[w]);
    }
    // strains: e1 of the vectors, Q of the pairs
    S e1[3][9], eq[4][9];
    double Nt[9];
    tying_interp(u, v, Nt);
    const int kk[5] = {0, 3, 1, 4, 2};
    const double ff[5] = {1.0, 1.0, 2.0, 2.0, 2.0};
    const int pa[5] = {0, 0, 0, 1, 1}, pb[5] = {0, 1, 2, 1, 2};
    S al[5][3], aq[5][4];
    for (int r = 0; r < 5; r++) {
      for (int w = 0; w < 3; w++) al[r][w] = 0.0;
      for (int p = 0; p < 4; p++) aq[r][p] = 0.0;
    }
    for (int t = 0; t < 9; t++) {
      S tl[3], tq[4];
      if constexpr (kLate) {
        tying_at<S, WM, PM>(t, X, fn, q, tl, tq);
      } else {
        for (int w = 0; w < 3; w++)
          if ((WM >> w) & 1) tl[w] = tyl[w][t];
      }
      const int sl = field_slot(tying_field(t));
      const int cc = pa[sl], dd = pb[sl];
      for (int r = 0; r < 5; r++) {
        const int ka = pa[kk[r]], kb = pb[kk[r]];
        const S M = (cc == dd) ? XT[3 * cc + ka] * XT[3 * cc + kb]
                               : XT[3 * cc + ka] * XT[3 * dd + kb] + XT[3 * dd + ka] * XT[3 * cc + kb];
        const S W = (ff[r] * Nt[t]) * M;
        for (int w = 0; w < 3; w++)
          if ((WM >> w) & 1) al[r][w] = al[r][w] + W * tl[w];
        for (int p = 0; p < 4; p++)
          if ((PM >> p) & 1) aq[r][p] = aq[r][p] + W * tq[p];
      }
    }
    for (int r = 0; r < 5; r++) {
      const int row = r < 3 ? r : (r == 3 ? 6 : 7);
      for (int w = 0; w < 3; w++)
        if ((WM >> w) & 1) e1[w][row] = al[r][w];
      for (int p = 0; p < 4; p++)
        if ((PM >> p) & 1) eq[p][row] = aq[r][p];
    }
    for (int w = 0; w < 3; w++) {
      if (!((WM >> w) & 1)) continue;
      e1[w][3] = U1[w][0];
      e1[w][4] = U1[w][4];
      e1[w][5] = U1[w][1] + U1[w][3];
      e1[w][8] = N[0] * etn[w][0] + N[1] * etn[w][1] + N[2] * etn[w][2] + N[3] * etn[w][3];
    }
    for (int p = 0; p < 4; p++) {
      if (!((PM >> p) & 1)) continue;
      const S *A0 = U0[kPairA[p]], *A1 = U1[kPairA[p]], *B0 = U0[kPairB[p]], *B1 = U1[kPairB[p]];
      eq[p][3] = A0[0] * B1[0] + A0[3] * B1[3] + A0[6] * B1[6] + B0[0] * A1[0] + B0[3] * A1[3] + B0[6] * A1[6];
      eq[p][4] = A0[1] * B1[1] + A0[4] * B1[4] + A0[7] * B1[7] + B0[1] * A1[1] + B0[4] * A1[4] + B0[7] * A1[7];
      eq[p][5] = A0[0] * B1[1] + A0[3] * B1[4] + A0[6] * B1[7] + A1[0] * B0[1] + A1[3] * B0[4] + A1[6] * B0[7] +
                 B0[0] * A1[1] + B0[3] * A1[4] + B0[6] * A1[7] + B1[0] * A0[1] + B1[3] * A0[4] + B1[6] * A0[7];
      eq[p][8] = 0.0;
    }
    // mechanical strain of the path (u = 0 where u is not evaluated)
    S eu[9];
    for (int i = 0; i < 9; i++) {
      if constexpr (kU)
        eu[i] = e1[1][i];
      else
        eu[i] = 0.0;
    }
    eu[0] = eu[0] - eth;
    eu[1] = eu[1] - eth;
    S y[9];
    if (kPhi && (c[0] != 0.0 || c[1] != 0.0)) {
      apply_cs(sec, det, e1[0], y);  // det C e1(phi)
      if (c[0] != 0.0) acc = acc + c[0] * dot9(e1[0], y);
      if constexpr (kGeo) {
        if (c[1] != 0.0) {
          S z[9];
          apply_cs(sec, det, eq[1], z);  // det C Q(phi, phi)
          acc = acc + c[1] * (2.0 * dot9(eq[0], y) + dot9(eu, z));
        }
      }
    }
    if (kPhi && c[3] != 0.0) {
      S u0[3] = {0.0, 0.0, 0.0}, d0[3] = {0.0, 0.0, 0.0};
      for (int m = 0; m < 4; m++) {
        const S qr[3] = {q[0][6 * m + 3], q[0][6 * m + 4], q[0][6 * m + 5]};
        S dm[3];
        cross3(qr, &fn[3 * m], dm);
        for (int d = 0; d < 3; d++) {
          u0[d] = u0[d] + N[m] * q[0][6 * m + d];
          d0[d] = d0[d] + N[m] * dm[d];
        }
      }
      acc = acc + c[3] * det * (m0 * dot3(u0, u0) + m2 * dot3(d0, d0));
    }
    if constexpr (kRes) {
      if (c[2] != 0.0) {
        S ep[9];
        for (int i = 0; i < 9; i++) ep[i] = e1[2][i];
        if constexpr (kResNl) {
          if (nonlin) {
            for (int i = 0; i < 9; i++) {
              ep[i] = ep[i] + eq[2][i];
              eu[i] = eu[i] + 0.5 * eq[3][i];
            }
          }
        }
        apply_cs(sec, det, eu, y);
        acc = acc + c[2] * dot9(ep, y);
      }
    }
  }
  out = acc;
}

template <int WM, int PM>
__global__ __launch_bounds__(kThreads) void xpt_kernel(XptArgs a, int nel) {
  // the element's three vectors, shared by its 12 threads (read-only)
  __shared__ double qs[kThreads / 12][3][24];
  const int tid = blockIdx.x * kThreads + threadIdx.x;
  const int le = tid / 12, dir = tid % 12, ll = threadIdx.x / 12;
  const bool active = le < nel;
  const int e = a.elems[active ? le : nel - 1];
  const long design = blockIdx.y;
  const int *cn = a.conn + 4 * e;
  const double *src[3] = {a.phi, a.u, a.psi};
  for (int k = dir; k < 72; k += 12) {
    const int w = k / 24, i = k % 24;
    qs[ll][w][i] = ((WM >> w) & 1) && src[w] ? src[w][design * a.stride + 6 * cn[i / 6] + i % 6] : 0.0;
  }
  __syncthreads();
  if (!active) return;
  const double(*q)[24] = qs[ll];
  Dual X[12];
  for (int i = 0; i < 12; i++) X[i] = Dual(a.Xpts[3 * cn[i / 3] + i % 3], i == dir ? 1.0 : 0.0);
  const double thick = a.thick[design * a.nelems + e];
  const double temp = a.temp ? a.temp[design * a.nelems + e] : 0.0;
  const Section sec = make_section(a.mat, thick);
  const double c[4] = {a.coef[4 * design], a.coef[4 * design + 1], a.coef[4 * design + 2], a.coef[4 * design + 3]};
  // mass moments of TACSIsoShellConstitutive with tOffset = 0
  const double m0 = a.mat.rho * thick, m2 = a.mat.rho * thick * thick * thick / 12.0;
  Dual s;
  element_scalars<Dual, WM, PM>(X, q, sec, a.mat.cte * temp, a.nonlin != 0, c, m0, m2, s);
  a.dfdX[design * a.dX_stride + 3 * cn[dir / 3] + dir % 3] += s.d;
}

}  // namespace

hipError_t launch_xpt_kernel(const XptArgs &args, int nel, int nbatch, hipStream_t stream) {
  if (nel <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((12L * nel + kThreads - 1) / kThreads);
  const dim3 grid(blocks, nbatch);
  const int t = args.terms;
  if ((t & kXptR) && (t & (kXptK | kXptG | kXptM)))  // every vector and pair
    hipLaunchKernelGGL((xpt_kernel<7, 15>), grid, dim3(kThreads), 0, stream, args, nel);
  else if (t & kXptR)  // u, psi; (u, psi) and (u, u) for the nonlinear model
    if (args.nonlin)
      hipLaunchKernelGGL((xpt_kernel<6, 12>), grid, dim3(kThreads), 0, stream, args, nel);
    else
      hipLaunchKernelGGL((xpt_kernel<6, 0>), grid, dim3(kThreads), 0, stream, args, nel);
  else if (t & kXptG)  // phi, u; (u, phi) and (phi, phi)
    hipLaunchKernelGGL((xpt_kernel<3, 3>), grid, dim3(kThreads), 0, stream, args, nel);
  else  // phi alone: s_K, s_M
    hipLaunchKernelGGL((xpt_kernel<1, 0>), grid, dim3(kThreads), 0, stream, args, nel);
  return hipGetLastError();
}

}  // namespace shellgpu
