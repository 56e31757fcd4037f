// elem_kernels.hip -- per-element residual / stiffness / geometric-stiffness
// kernels for the 4-node director shell, two elements per wavefront.
//
// Formulation.  The reference (TACSShellElement.h:303-672) builds R_e and
// K_e by chaining first/second derivatives of the strain energy through
// (u0x, u1x, tying strains, drilling strain) back to the element dofs.  Here
// the same energy is differentiated in strain-Jacobian form, which maps onto
// a wavefront as a small GEMM:
//
//   e_qp(q)  = 9 strains at quadrature point qp (membrane 3, bending 3,
//              transverse shear 2 from MITC tying, drilling 1)
//   B_qp(q)  = de_qp/dq                                       (9 x 24)
//   sig_qp   = w detXd C (e_qp - e_thermal)                   (9)
//   R_e      = sum_qp B_qp^T sig_qp
//   K_e      = sum_qp B_qp^T (w detXd C) B_qp + sum_qp sum_i sig_i d2e_i/dq2
//
// The strains of both reference models are at most quadratic in q, so
// B_qp(q) = B0 + (terms linear in q) and the reference's geometric stiffness
// -- a central difference of the nonlinear tangent along the state
// (TACSShellElement.h:705-766) -- equals, up to round-off, the exact
// directional derivative computed here:
//
//   G_e = sum_qp [B0^T (wC) B1 + B1^T (wC) B0] + sum_qp sum_i sig_i(u) d2e_i/dq2
//
// with B1 = dB/d(eps) along u and sig(u) = w detXd C (B0 u - e_thermal).
//
// Work split: each half-wavefront (32 lanes) owns one element and its LDS.
// Strain-Jacobian columns run on lanes 0-23 (B1 after B) and the 24 x 24
// product as one 3 x 6 register tile per lane.  The geometry per node /
// quadrature point / tying point (4 to 9 tasks per element) is spread over
// the workgroup by (element, point): with two elements per workgroup (one
// wavefront) it fills lanes the per-element phases would leave idle; the
// wide schedules (4 elements, 2 wavefronts) give one wavefront the frames
// of all 4 elements while the other computes their tying points.  Elements
// of one colour share no node, so each half adds its 16 blocks into the BSR
// values without atomics.
#include "elem_common.h"
#include "shellgpu_internal.h"

namespace shellgpu {

// What each kernel mode computes.
template <int MODE>
struct ModeTraits {
  // residual and alpha K + gamma M in one pass (addJacobian)
  static constexpr bool kJac = (MODE == kModeJacLin || MODE == kModeJacNl);
  // (alpha K + gamma M) x per element (addJacobianVecProduct)
  static constexpr bool kJvp = (MODE == kModeJvpLin || MODE == kModeJvpNl);
  static constexpr bool kMassTerms = kJac || kJvp;
  static constexpr bool kMat = (MODE == kModeKLin || MODE == kModeKNl || MODE == kModeG || kJac || kJvp);
  static constexpr bool kState = (MODE != kModeKLin);
  static constexpr bool kNonlin = (MODE == kModeKNl || MODE == kModeRNl || MODE == kModeAdjResNl ||
                                   MODE == kModeOutNl || MODE == kModeJacNl || MODE == kModeJvpNl);
  // strains, stresses and failure values at the element nodes instead of
  // the quadrature points (getOutputData)
  static constexpr bool kOut = (MODE == kModeOutLin || MODE == kModeOutNl);
  static constexpr bool kRes = (MODE == kModeRLin || MODE == kModeRNl || MODE == kModeAdjResLin ||
                                MODE == kModeAdjResNl || kJac);
  static constexpr bool kGeo = (MODE == kModeKNl || MODE == kModeG || MODE == kModeJacNl || MODE == kModeJvpNl);
  // B1 = dB/dq along the state is formed too (geometric stiffness and the
  // thickness adjoint), which doubles the tying Jacobian and B storage
  static constexpr bool kTwo = (MODE == kModeG || MODE == kModeAdjDV);
  static constexpr bool kAdj =
      (MODE == kModeAdjDV || MODE == kModeAdjResLin || MODE == kModeAdjResNl);
  static constexpr bool kStrain = kRes || kGeo;  // stresses at the quadrature points
};

// LDS of one element, arrays sized to what the mode uses.  The union holds
// arrays of disjoint phases: the node frames and tying coefficients (P1-P3)
// and the B / C B columns (P5 on).
template <int MODE>
struct ElemSmem {
  using MT = ModeTraits<MODE>;
  static constexpr int kNB = MT::kTwo ? 2 : 1;
  double X[12];
  double s[24];
  double v2[MT::kAdj ? 24 : 1];  // second state of the adjoint modes: path u or psi
  double dd[MT::kJac ? 24 : 1];  // second time derivative of the state (addJacobian)
  double fn[12];
  double T[36], XT[36], XzT[36], wdet[4];
  double ety[MT::kState ? 9 : 1];
  // [qp][5 strain rows (0,1,2,6,7)][5 tying fields]: transformation of the
  // interpolated tying fields to the strain rows
  double Wf[100];
  double u0x[MT::kState ? 36 : 1], u1x[MT::kState ? 36 : 1];
  double drot[96];  // [node][24] d(drill strain at node)/dq
  double etn[4];
  double sig[MT::kStrain ? 36 : 1];  // [qp][9]
  double dety[MT::kGeo ? 9 : 1];
  struct Frames {
    double Tn[36], XTn[36], tXxi[54];
    double tcU[kNB][54], tcd[kNB][27];  // tying derivative coefficients
  };
  // [strain][24] for one qp.  The adjoint modes, which form no C B
  // products, keep in Y[0]: five strain vectors [0,45) and two weighted
  // stress vectors [45,63)
  struct Cols {
    double B[kNB][216], Y[kNB][216];
  };
  union {
    Frames f;
    Cols c;
  };
};

// Tying fields interpolated at a point with the weights Nt of tying_interp:
// each tying point feeds only its own field (g11: points 0,1; g22: 2,3;
// g12: 4; g23: 5,6; g13: 7,8)
__device__ __forceinline__ void tying_fields(const double Nt[9], const double x[9], double g[5]) {
  g[kG11] = Nt[0] * x[0] + Nt[1] * x[1];
  g[kG22] = Nt[2] * x[2] + Nt[3] * x[3];
  g[kG12] = Nt[4] * x[4];
  g[kG23] = Nt[5] * x[5] + Nt[6] * x[6];
  g[kG13] = Nt[7] * x[7] + Nt[8] * x[8];
}

// rows of the 9-strain vector fed by the tying strains
__device__ __forceinline__ int tying_row(int r) {
  return r < 3 ? r : (r == 3 ? 6 : 7);
}

// alpha/beta/gamma factors of dof k at a quadrature point:
// d(u0x)/dq_k = alpha beta^T, d(u1x)/dq_k = alpha gamma^T
// (from TacsShellComputeDispGrad, TACSShellUtilities.h:361-422, with the
// linearized-rotation director d = q x n, TACSDirector.h:215-242)
template <class SM>
__device__ __forceinline__ void dof_abg(const SM &sm, int iq, int k,
                                        const double N[4], const double Na[4],
                                        const double Nb[4], double al[3],
                                        double be[3], double ga[3]) {
  const int n = k / 6, r = k % 6;
  const double *T = &sm.T[9 * iq], *XT = &sm.XT[9 * iq],
               *XzT = &sm.XzT[9 * iq];
  if (r < 3) {
    al[0] = T[3 * r];
    al[1] = T[3 * r + 1];
    al[2] = T[3 * r + 2];
#pragma unroll
    for (int b = 0; b < 3; b++) {
      be[b] = Na[n] * XT[b] + Nb[n] * XT[3 + b];
      ga[b] = Na[n] * XzT[b] + Nb[n] * XzT[3 + b];
    }
  } else {
    const double *f = &sm.fn[3 * n];
    double c[3];
    const int rr = r - 3;
    // c = e_rr x f
    c[0] = rr == 0 ? 0.0 : (rr == 1 ? f[2] : -f[1]);
    c[1] = rr == 0 ? -f[2] : (rr == 1 ? 0.0 : f[0]);
    c[2] = rr == 0 ? f[1] : (rr == 1 ? -f[0] : 0.0);
#pragma unroll
    for (int a = 0; a < 3; a++) al[a] = T[a] * c[0] + T[3 + a] * c[1] + T[6 + a] * c[2];
#pragma unroll
    for (int b = 0; b < 3; b++) {
      be[b] = N[n] * XT[6 + b];
      ga[b] = Na[n] * XT[b] + Nb[n] * XT[3 + b] + N[n] * XzT[6 + b];
    }
  }
}

enum BVariant { kBLin = 0, kBFull = 1, kBQuad = 2 };

// one column of B_qp (9 strains) for dof k
// (J: this dof's column of the tying-strain Jacobian, one value per tying point)
template <int VAR, class SM>
__device__ __forceinline__ void build_bcol(const SM &sm, int iq, int k,
                                           const double J[9], double out[9]) {
  double u, v, N[4], Na[4], Nb[4];
  quad_point(iq, u, v);
  shape4(u, v, N, Na, Nb);
  double Nt[9], g[5];
  tying_interp(u, v, Nt);
  tying_fields(Nt, J, g);
  const double *Wq = &sm.Wf[25 * iq];
#pragma unroll
  for (int r = 0; r < 5; r++) {
    double acc = 0.0;
#pragma unroll
    for (int f = 0; f < 5; f++) acc += Wq[5 * r + f] * g[f];
    out[tying_row(r)] = acc;
  }
  double al[3], be[3], ga[3];
  dof_abg(sm, iq, k, N, Na, Nb, al, be, ga);
  double e3 = 0.0, e4 = 0.0, e5 = 0.0;
  if (VAR != kBQuad) {
    e3 = al[0] * ga[0];
    e4 = al[1] * ga[1];
    e5 = al[0] * ga[1] + al[1] * ga[0];
  }
  if (VAR != kBLin) {
    const double *U0 = &sm.u0x[9 * iq], *U1 = &sm.u1x[9 * iq];
    const double p00 = al[0] * U0[0] + al[1] * U0[3] + al[2] * U0[6];
    const double p01 = al[0] * U0[1] + al[1] * U0[4] + al[2] * U0[7];
    const double p10 = al[0] * U1[0] + al[1] * U1[3] + al[2] * U1[6];
    const double p11 = al[0] * U1[1] + al[1] * U1[4] + al[2] * U1[7];
    e3 += be[0] * p10 + ga[0] * p00;
    e4 += be[1] * p11 + ga[1] * p01;
    e5 += be[0] * p11 + ga[1] * p00 + ga[0] * p01 + be[1] * p10;
  }
  out[3] = e3;
  out[4] = e4;
  out[5] = e5;
  double e8 = 0.0;
  if (VAR != kBQuad) {
#pragma unroll
    for (int m = 0; m < 4; m++) e8 += N[m] * sm.drot[24 * m + k];
  }
  out[8] = e8;
}

// Isotropic section stiffness with the entries that are zero by construction
// dropped: A = [[a11 a12 0][a12 a11 0][0 0 a66]], D likewise, As = as I,
// drill.  make_section and make_section_dt produce exactly this pattern, and
// the dropped terms add exact zeros, so C b is unchanged bit for bit.
struct IsoSection {
  double a11, a12, a66, d11, d12, d66, as, drill;
};

__device__ __forceinline__ IsoSection iso_section(const Section &c) {
  return IsoSection{c.A[0], c.A[1], c.A[5], c.D[0], c.D[1], c.D[5], c.As[0], c.drill};
}

// Y = C B for one column (C: A membrane, D bending, As shear, drill)
__device__ __forceinline__ void apply_c(const IsoSection &c, double w,
                                        const double b[9], double y[9]) {
  y[0] = w * (c.a11 * b[0] + c.a12 * b[1]);
  y[1] = w * (c.a12 * b[0] + c.a11 * b[1]);
  y[2] = w * (c.a66 * b[2]);
  y[3] = w * (c.d11 * b[3] + c.d12 * b[4]);
  y[4] = w * (c.d12 * b[3] + c.d11 * b[4]);
  y[5] = w * (c.d66 * b[5]);
  y[6] = w * (c.as * b[6]);
  y[7] = w * (c.as * b[7]);
  y[8] = w * (c.drill * b[8]);
}

// second derivative of the tying strains summed with weights dety
// (the dety * d2(ety)/dq2 terms of addComputeTyingStrainHessian,
// TACSShellElementModel.h:782-1031)
template <class SM>
__device__ __forceinline__ double tying_geo(const SM &sm, int k, int l) {
  const int n = k / 6, r = k % 6, m = l / 6, rp = l % 6;
  if (r >= 3 && rp >= 3) return 0.0;
  if (r < 3 && rp < 3) {
    if (r != rp) return 0.0;
    double h = 0.0;
#pragma unroll
    for (int t = 0; t < 5; t++) {
      double u, v, N[4], Na[4], Nb[4];
      tying_point(t, u, v);
      shape4(u, v, N, Na, Nb);
      if (t < 2) {
        h += sm.dety[t] * (Na[n] * Na[m]);
      } else if (t < 4) {
        h += sm.dety[t] * (Nb[n] * Nb[m]);
      } else {
        h += sm.dety[t] * (0.5 * (Na[n] * Nb[m] + Nb[n] * Na[m]));
      }
    }
    return h;
  }
  // one displacement dof (node nd, comp cd) and one rotation dof (node nr, comp cr)
  const int nd = r < 3 ? n : m, cd = r < 3 ? r : rp;
  const int nr = r < 3 ? m : n, cr = (r < 3 ? rp : r) - 3;
  const double *f = &sm.fn[3 * nr];
  double c[3];
  c[0] = cr == 0 ? 0.0 : (cr == 1 ? f[2] : -f[1]);
  c[1] = cr == 0 ? -f[2] : (cr == 1 ? 0.0 : f[0]);
  c[2] = cr == 0 ? f[1] : (cr == 1 ? -f[0] : 0.0);
  double h = 0.0;
#pragma unroll
  for (int t = 5; t < 9; t++) {
    double u, v, N[4], Na[4], Nb[4];
    tying_point(t, u, v);
    shape4(u, v, N, Na, Nb);
    const double g = (t < 7) ? Nb[nd] : Na[nd];
    h += sm.dety[t] * (0.5 * N[nr] * c[cd] * g);
  }
  return h;
}

// Consistent mass of the linearized-rotation shell between nodes ni and nj
// (the derivation is given at mass_kernel below): with S = sum_qp detXd
// N_ni N_nj, translations couple by m0 S I and rotations by
// m2 S ((t_i . t_j) I - t_j t_i^T), t = node normals.  MassTile holds the
// factors of one 3 x 6 tile (rows rr..rr+2 of node ni, the 6 dofs of nj).
struct MassTile {
  double S, m0, m2, tt;
  const double *ti, *tj;
  int rr;
};

// S = sum_qp detXd N_ni N_nj (Gauss weights 1)
__device__ __forceinline__ double mass_weight(const double wdet[4], int ni, int nj) {
  const double ha = (ni & 1) ? 0.5 : -0.5, hb = (ni & 2) ? 0.5 : -0.5;
  const double ga = (nj & 1) ? 0.5 : -0.5, gb = (nj & 2) ? 0.5 : -0.5;
  double S = 0.0;
#pragma unroll
  for (int iq = 0; iq < 4; iq++) {
    double u, v;
    quad_point(iq, u, v);
    // N of nodes ni and nj at the point, (0.5 + h u)(0.5 + h v) as in shape4
    const double Ni = (0.5 + ha * u) * (0.5 + hb * v), Nj = (0.5 + ga * u) * (0.5 + gb * v);
    S += wdet[iq] * (Ni * Nj);
  }
  return S;
}

__device__ __forceinline__ MassTile mass_tile(const double wdet[4], const double fn[12], int ni, int nj, int rr,
                                              double m0, double m2) {
  MassTile m;
  m.S = mass_weight(wdet, ni, nj);
  m.m0 = m0;
  m.m2 = m2;
  m.ti = &fn[3 * ni];
  m.tj = &fn[3 * nj];
  m.tt = m.ti[0] * m.tj[0] + m.ti[1] * m.tj[1] + m.ti[2] * m.tj[2];
  m.rr = rr;
  return m;
}

// entry (i, j) of the tile, i < 3, j < 6
__device__ __forceinline__ double mass_entry(const MassTile &m, int i, int j) {
  if (m.rr == 0) return (j == i) ? m.m0 * m.S : 0.0;
  return (j < 3) ? 0.0 : m.m2 * m.S * ((j - 3 == i ? m.tt : 0.0) - m.tj[i] * m.ti[j - 3]);
}

// Synchronisation of one element's 32 lanes once the geometry phases are
// done (those share data across elements and take the workgroup barrier).
// With one wavefront per workgroup: the workgroup barrier.  With wide
// workgroups the 32 lanes are half of one wavefront, whose LDS operations
// complete in issue order: waiting for its own LDS operations and keeping
// the compiler from moving memory operations across suffices, and the
// workgroup's wavefronts then run their quadrature loops independently.
template <int EPW>
__device__ __forceinline__ void elem_sync() {
  if constexpr (EPW == 2) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// EPW elements per workgroup: lanes 32 h .. 32 h + 31 work on element
// EPW blockIdx.x + h of the launch list, each with its own LDS.  Slots past
// the end of the list compute a copy of its last element with their writes
// suppressed.
//
// VAR selects a compiled schedule of the same arithmetic (chosen per mode
// and mesh by shellgpu_tune): bit 0 reads the lane's BSR block entries at
// the scatter instead of before the quadrature loop (fewer live registers
// through the loop).
template <int MODE, int VAR, int EPW>
__device__ __forceinline__ void elem_body(const ElemArgs &a) {
  using MT = ModeTraits<MODE>;
  constexpr bool kTwo = MT::kTwo;
  constexpr int kNB = ElemSmem<MODE>::kNB;
  static_assert(EPW == 2 || EPW == 4 || EPW == 8, "2, 4 or 8 elements per workgroup");
  __shared__ ElemSmem<MODE> smem[EPW];
  const int half = threadIdx.x >> 5, lane = threadIdx.x & 31;
  ElemSmem<MODE> &sm = smem[half];
  const int slot = EPW * blockIdx.x + half;
  const bool valid = slot < a.count;
  const int e = a.elems[valid ? slot : a.count - 1];
  // The geometry phases (P1-P2) run per node, quadrature point or tying
  // point: their tasks go to the workgroup's threads by (element, point),
  // not by each element's own 32 lanes, so with 4 elements per workgroup
  // one wavefront computes the frames of all 4 while the other computes
  // their tying points (with 8, two others).  Task classes of one phase
  // start at thread 0 and kG2.
  const int gt = threadIdx.x;
  constexpr int kG2 = EPW >= 4 ? 64 : 16;
  static_assert(kG2 >= 8 * EPW && kG2 + 9 * EPW <= 32 * EPW, "geometry task classes overlap");
  const long design = blockIdx.y;
  // output modes: the element's 4 x 31 output values
  [[maybe_unused]] double *obase = a.out ? a.out + (design * a.nelems + e) * 124 : nullptr;
  constexpr bool kMat = MT::kMat, kState = MT::kState, kNonlin = MT::kNonlin;
  constexpr bool kRes = MT::kRes, kGeo = MT::kGeo;

  const double thick = a.thick[design * a.nelems + e];
  double temp = a.temp ? a.temp[design * a.nelems + e] : 0.0;
  double alpha = a.alpha;
  const IsoSection sec = iso_section(make_section(a.mat, thick));
  const int *cn = a.conn + 4 * e;

  // ---- P0: gather coordinates and state ----
  if (lane < 12) sm.X[lane] = a.Xpts[3 * cn[lane / 3] + lane % 3];
  if (lane < 24) {
    sm.s[lane] = (kState && a.state)
                     ? a.state[design * a.state_stride + 6 * cn[lane / 6] + lane % 6]
                     : 0.0;
    if (MODE == kModeAdjDV)
      sm.v2[lane] = a.u2 ? a.u2[design * a.u2_stride + 6 * cn[lane / 6] + lane % 6] : 0.0;
    if (MODE == kModeAdjResLin || MODE == kModeAdjResNl)
      sm.v2[lane] = a.psi[design * a.psi_stride + 6 * cn[lane / 6] + lane % 6];
    if (MT::kJac) sm.dd[lane] = a.ddu ? a.ddu[design * a.ddu_stride + 6 * cn[lane / 6] + lane % 6] : 0.0;
  }
  __syncthreads();

  // central-difference step of getMatType(GEOMETRIC) (TACSShellElement.h:
  // 705-766): state h u/n, temperature T (1 + h/n), scale n/(2h)
  if (MODE == kModeKNl && a.fd_sign != 0) {
    const double dh = 1e-4;
    double nrm = 0.0;
#pragma unroll
    for (int i = 0; i < 24; i++) nrm += sm.s[i] * sm.s[i];
    nrm += temp * temp;
    nrm = (nrm == 0.0) ? 1.0 : sqrt(nrm);
    __syncthreads();
    if (lane < 24) sm.s[lane] = (a.fd_sign * dh) * sm.s[lane] / nrm;
    temp = temp + (a.fd_sign * dh) * temp / nrm;
    alpha = a.alpha * (a.fd_sign * (0.5 * nrm / dh));
    __syncthreads();
  }

  // ---- P1: node normals; tying-point X,xi ----
  // (TacsShellComputeNodeNormals, TACSShellUtilities.h:301-343)
  auto p1_normal = [&](ElemSmem<MODE> &sm, int n) {
    double u, v, N[4], Na[4], Nb[4];
    node_point(n, u, v);
    shape4(u, v, N, Na, Nb);
    double Xxi[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int d = 0; d < 3; d++) {
        Xxi[2 * d] += Na[m] * sm.X[3 * m + d];
        Xxi[2 * d + 1] += Nb[m] * sm.X[3 * m + d];
      }
    double a3[3] = {Xxi[0], Xxi[2], Xxi[4]}, b3[3] = {Xxi[1], Xxi[3], Xxi[5]};
    double f[3];
    cross3(a3, b3, f);
    const double nrm = sqrt(dot3(f, f));
    if (nrm != 0.0) {
      const double inv = 1.0 / nrm;
      f[0] *= inv;
      f[1] *= inv;
      f[2] *= inv;
    }
    sm.fn[3 * n] = f[0];
    sm.fn[3 * n + 1] = f[1];
    sm.fn[3 * n + 2] = f[2];
  };
  auto p1_tying = [&](ElemSmem<MODE> &sm, int t) {
    double u, v, N[4], Na[4], Nb[4];
    tying_point(t, u, v);
    shape4(u, v, N, Na, Nb);
    double Xxi[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int d = 0; d < 3; d++) {
        Xxi[2 * d] += Na[m] * sm.X[3 * m + d];
        Xxi[2 * d + 1] += Nb[m] * sm.X[3 * m + d];
      }
#pragma unroll
    for (int i = 0; i < 6; i++) sm.f.tXxi[6 * t + i] = Xxi[i];
  };
  if (gt < 4 * EPW)
    p1_normal(smem[gt >> 2], gt & 3);
  else if (gt >= kG2 && gt < kG2 + 9 * EPW)
    p1_tying(smem[(gt - kG2) / 9], (gt - kG2) % 9);
  __syncthreads();

  // ---- P1b: frames at the nodes (points 0-3: the drilling-strain frames)
  // and at the evaluation points (4-7: the quadrature points, or the nodes
  // in the output modes) in one pass of the same code; tying-point state
  // and derivative coefficients ----
  auto p1b_frame = [&](ElemSmem<MODE> &sm, int p) {
    const bool at_node = p < 4;
    const int iq = p & 3;
    double u, v, N[4], Na[4], Nb[4];
    if (at_node || MT::kOut)
      node_point(iq, u, v);
    else
      quad_point(iq, u, v);
    shape4(u, v, N, Na, Nb);
    // at a node the interpolated normal is that node's normal exactly
    double Xxi[6] = {0, 0, 0, 0, 0, 0}, n0[3] = {0, 0, 0};
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int d = 0; d < 3; d++) {
        Xxi[2 * d] += Na[m] * sm.X[3 * m + d];
        Xxi[2 * d + 1] += Nb[m] * sm.X[3 * m + d];
        n0[d] += N[m] * sm.fn[3 * m + d];
      }
    double T[9], Xd[9], Xdi[9], XT[9];
    natural_transform(Xxi, n0, T);
    frame3(Xxi, n0, Xd);
    const double det = inv3(Xd, Xdi);
    mm3(Xdi, T, XT);
    if (at_node) {
#pragma unroll
      for (int i = 0; i < 9; i++) {
        sm.f.Tn[9 * iq + i] = T[i];
        sm.f.XTn[9 * iq + i] = XT[i];
      }
    } else {
      double nxi[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int m = 0; m < 4; m++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
          nxi[2 * d] += Na[m] * sm.fn[3 * m + d];
          nxi[2 * d + 1] += Nb[m] * sm.fn[3 * m + d];
        }
      double Xdz[9], nXdz[9], XzT[9];
      frame3z(nxi, Xdz);
      mm3(Xdi, Xdz, nXdz);
#pragma unroll
      for (int i = 0; i < 9; i++) nXdz[i] = -nXdz[i];
      mm3(nXdz, XT, XzT);
#pragma unroll
      for (int i = 0; i < 9; i++) {
        sm.T[9 * iq + i] = T[i];
        sm.XT[9 * iq + i] = XT[i];
        sm.XzT[9 * iq + i] = XzT[i];
      }
      sm.wdet[iq] = det;  // Gauss weight 1
    }
  };
  auto p1b_tying = [&](ElemSmem<MODE> &sm, int t) {
    double u, v, N[4], Na[4], Nb[4];
    tying_point(t, u, v);
    shape4(u, v, N, Na, Nb);
    double n0[3] = {0, 0, 0}, U[6] = {0, 0, 0, 0, 0, 0}, d0[3] = {0, 0, 0};
#pragma unroll
    for (int m = 0; m < 4; m++) {
#pragma unroll
      for (int d = 0; d < 3; d++) n0[d] += N[m] * sm.fn[3 * m + d];
      if (kState) {  // the linear stiffness has no state: U = d0 = 0
        double dm[3];
        cross3(&sm.s[6 * m + 3], &sm.fn[3 * m], dm);
#pragma unroll
        for (int d = 0; d < 3; d++) {
          d0[d] += N[m] * dm[d];
          U[2 * d] += Na[m] * sm.s[6 * m + d];
          U[2 * d + 1] += Nb[m] * sm.s[6 * m + d];
        }
      }
    }
    const double *X = &sm.f.tXxi[6 * t];
    const int f = tying_field(t);
    // tying strain value (TACSShellElementModel.h:32-77 and 643-697)
    double lin = 0.0, quad = 0.0;
    if (f == kG11) {
      lin = U[0] * X[0] + U[2] * X[2] + U[4] * X[4];
      quad = 0.5 * (U[0] * U[0] + U[2] * U[2] + U[4] * U[4]);
    } else if (f == kG22) {
      lin = U[1] * X[1] + U[3] * X[3] + U[5] * X[5];
      quad = 0.5 * (U[1] * U[1] + U[3] * U[3] + U[5] * U[5]);
    } else if (f == kG12) {
      lin = 0.5 * (U[0] * X[1] + U[2] * X[3] + U[4] * X[5] + U[1] * X[0] + U[3] * X[2] + U[5] * X[4]);
      quad = 0.5 * (U[0] * U[1] + U[2] * U[3] + U[4] * U[5]);
    } else if (f == kG23) {
      lin = 0.5 * (X[1] * d0[0] + X[3] * d0[1] + X[5] * d0[2] + n0[0] * U[1] + n0[1] * U[3] + n0[2] * U[5]);
      quad = 0.5 * (d0[0] * U[1] + d0[1] * U[3] + d0[2] * U[5]);
    } else {
      lin = 0.5 * (X[0] * d0[0] + X[2] * d0[1] + X[4] * d0[2] + n0[0] * U[0] + n0[1] * U[2] + n0[2] * U[4]);
      quad = 0.5 * (d0[0] * U[0] + d0[1] * U[2] + d0[2] * U[4]);
    }
    if (kState) sm.ety[t] = kNonlin ? lin + quad : lin;
    // derivative coefficients with a = xfac X + ufac U, nd = xfac n0 + ufac d0
    for (int pass = 0; pass < 2; pass++) {
      if (pass == 1 && MODE != kModeG && MODE != kModeAdjDV) break;
      const double xf = (pass == 0) ? 1.0 : 0.0;
      const double uf = (pass == 0) ? (kNonlin ? 1.0 : 0.0) : 1.0;
      double av[6], nd[3], dU[6] = {0, 0, 0, 0, 0, 0}, dd[3] = {0, 0, 0};
#pragma unroll
      for (int i = 0; i < 6; i++) av[i] = xf * X[i] + uf * U[i];
#pragma unroll
      for (int i = 0; i < 3; i++) nd[i] = xf * n0[i] + uf * d0[i];
      if (f == kG11) {
        dU[0] = av[0];
        dU[2] = av[2];
        dU[4] = av[4];
      } else if (f == kG22) {
        dU[1] = av[1];
        dU[3] = av[3];
        dU[5] = av[5];
      } else if (f == kG12) {
        dU[0] = 0.5 * av[1];
        dU[1] = 0.5 * av[0];
        dU[2] = 0.5 * av[3];
        dU[3] = 0.5 * av[2];
        dU[4] = 0.5 * av[5];
        dU[5] = 0.5 * av[4];
      } else if (f == kG23) {
        dU[1] = 0.5 * nd[0];
        dU[3] = 0.5 * nd[1];
        dU[5] = 0.5 * nd[2];
        dd[0] = 0.5 * av[1];
        dd[1] = 0.5 * av[3];
        dd[2] = 0.5 * av[5];
      } else {
        dU[0] = 0.5 * nd[0];
        dU[2] = 0.5 * nd[1];
        dU[4] = 0.5 * nd[2];
        dd[0] = 0.5 * av[0];
        dd[1] = 0.5 * av[2];
        dd[2] = 0.5 * av[4];
      }
      double *cU = sm.f.tcU[pass];
      double *cd = sm.f.tcd[pass];
#pragma unroll
      for (int i = 0; i < 6; i++) cU[6 * t + i] = dU[i];
#pragma unroll
      for (int i = 0; i < 3; i++) cd[3 * t + i] = dd[i];
    }
  };
  if (gt < 8 * EPW)
    p1b_frame(smem[gt >> 3], gt & 7);
  else if (gt >= kG2 && gt < kG2 + 9 * EPW)
    p1b_tying(smem[(gt - kG2) / 9], (gt - kG2) % 9);
  __syncthreads();

  // ---- P2: tying-field weights and the state's displacement gradients at
  // the evaluation points (TacsShellComputeDispGrad, TACSShellUtilities.h:
  // 361-422) from the frames P1b stored -- a task of its own, so that it
  // runs beside the drilling rows rather than on the frame wavefront's
  // critical path -- and the drilling rows at the nodes ----
  auto p2_point = [&](ElemSmem<MODE> &sm, int iq) {
    double u, v, N[4], Na[4], Nb[4];
    if (MT::kOut)
      node_point(iq, u, v);
    else
      quad_point(iq, u, v);
    shape4(u, v, N, Na, Nb);
    double T[9], XT[9], XzT[9];
#pragma unroll
    for (int i = 0; i < 9; i++) {
      T[i] = sm.T[9 * iq + i];
      XT[i] = sm.XT[9 * iq + i];
      XzT[i] = sm.XzT[9 * iq + i];
    }
    // tying field -> strain weights: e0ty = XT^T gty XT, gty the tying
    // fields interpolated at the point (mat3x3SymmTransformTranspose,
    // TACSElementAlgebra.h:1094); the interpolation weights are applied
    // where the fields are formed (tying_fields)
    const int kk[5] = {0, 3, 1, 4, 2};
    const double ff[5] = {1.0, 1.0, 2.0, 2.0, 2.0};
    const int pa[5] = {0, 0, 0, 1, 1}, pb[5] = {0, 1, 2, 1, 2};
#pragma unroll
    for (int r = 0; r < 5; r++) {
      const int ka = pa[kk[r]], kb = pb[kk[r]];
#pragma unroll
      for (int f = 0; f < 5; f++) {
        const int sl = field_slot(f);
        const int c = pa[sl], d = pb[sl];
        const double M = (c == d) ? XT[3 * c + ka] * XT[3 * c + kb]
                                  : XT[3 * c + ka] * XT[3 * d + kb] + XT[3 * d + ka] * XT[3 * c + kb];
        sm.Wf[25 * iq + 5 * r + f] = ff[r] * M;
      }
    }
    if (kState) {
      // displacement-gradient state at the point: u0x, u1x
      double d0[3] = {0, 0, 0}, d0xi[6] = {0, 0, 0, 0, 0, 0}, u0xi[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int m = 0; m < 4; m++) {
        double dm[3];
        cross3(&sm.s[6 * m + 3], &sm.fn[3 * m], dm);
#pragma unroll
        for (int d = 0; d < 3; d++) {
          d0[d] += N[m] * dm[d];
          d0xi[2 * d] += Na[m] * dm[d];
          d0xi[2 * d + 1] += Nb[m] * dm[d];
          u0xi[2 * d] += Na[m] * sm.s[6 * m + d];
          u0xi[2 * d + 1] += Nb[m] * sm.s[6 * m + d];
        }
      }
      double u0d[9], u1d[9], tmp[9], tmp2[9], o0[9], o1[9];
      frame3(u0xi, d0, u0d);
      frame3z(d0xi, u1d);
      mm3(u1d, XT, tmp);
      mm3(u0d, XzT, tmp2);
#pragma unroll
      for (int i = 0; i < 9; i++) tmp[i] += tmp2[i];
      mtm3(T, tmp, o1);
      mm3(u0d, XT, tmp);
      mtm3(T, tmp, o0);
#pragma unroll
      for (int i = 0; i < 9; i++) {
        sm.u0x[9 * iq + i] = o0[i];
        sm.u1x[9 * iq + i] = o1[i];
      }
    }
  };
  auto p2_drill = [&](ElemSmem<MODE> &sm, int n) {
    // TacsShellAddDrillStrainSens map (TACSShellUtilities.h:787-835) with
    // TACSLinearizedRotation: et = 0.5 (Ct[3] + u0x[3] - Ct[1] - u0x[1])
    const double *Tn = &sm.f.Tn[9 * n], *XTn = &sm.f.XTn[9 * n];
    const double dA[9] = {0, -0.5, 0, 0.5, 0, 0, 0, 0, 0};
    double tmp[9], dC[9], du0d[9];
    // dC = Tn dA Tn^T ; du0d = Tn dA XTn^T
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        tmp[3 * i + j] = dA[3 * i] * Tn[3 * j] + dA[3 * i + 1] * Tn[3 * j + 1] +
                         dA[3 * i + 2] * Tn[3 * j + 2];
      }
    mm3(Tn, tmp, dC);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        tmp[3 * i + j] = dA[3 * i] * XTn[3 * j] + dA[3 * i + 1] * XTn[3 * j + 1] +
                         dA[3 * i + 2] * XTn[3 * j + 2];
      }
    mm3(Tn, tmp, du0d);
    // rotation dofs of node n enter through C, displacements of every node
    // through u0x
    const double rq[3] = {-(dC[7] - dC[5]), -(dC[2] - dC[6]), -(dC[3] - dC[1])};
    double u, v, N[4], Na[4], Nb[4];
    node_point(n, u, v);
    shape4(u, v, N, Na, Nb);
    const double g[6] = {du0d[0], du0d[1], du0d[3], du0d[4], du0d[6], du0d[7]};
    double et = 0.0;
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int d = 0; d < 6; d++) {
        const double val = (d < 3) ? Na[m] * g[2 * d] + Nb[m] * g[2 * d + 1] : (m == n ? rq[d - 3] : 0.0);
        sm.drot[24 * n + 6 * m + d] = val;
        et += val * sm.s[6 * m + d];
      }
    sm.etn[n] = et;
  };
  if (gt < 4 * EPW)
    p2_drill(smem[gt >> 2], gt & 3);
  else if (gt >= kG2 && gt < kG2 + 4 * EPW)
    p2_point(smem[(gt - kG2) >> 2], (gt - kG2) & 3);
  __syncthreads();

  // ---- P3: tying-strain Jacobian rows; strains and stresses at the
  // quadrature points ----
  // lane k < 24 keeps column k of the tying-strain Jacobian (one value per
  // tying point) in registers: only it reads that column when building B
  double Jr[kNB][9];
  {
    if (!MT::kOut && lane < 24) {
      const int k = lane;
      const int n = k / 6, r = k % 6;
      // shape values of node n at the tying points (coordinates in {-1, 0, 1},
      // so every value below is exact and equals shape4's)
      const double ha = (n & 1) ? 0.5 : -0.5, hb = (n & 2) ? 0.5 : -0.5;
      // rotation rows: component c of fn[n] x dd
      const int c = r < 3 ? 0 : r - 3, i1 = (c + 1) % 3, i2 = (c + 2) % 3;
      const double fa = sm.fn[3 * n + i1], fb = sm.fn[3 * n + i2];
#pragma unroll
      for (int pass = 0; pass < kNB; pass++) {
        const double *cU = sm.f.tcU[pass];
        const double *cd = sm.f.tcd[pass];
        double *J = Jr[pass];
#pragma unroll
        for (int t = 0; t < 9; t++) {
          double u, v;
          tying_point(t, u, v);
          const double na = 0.5 + ha * u, nb = 0.5 + hb * v;
          double val;
          if (r < 3) {  // N,a = ha nb and N,b = na hb
            val = cU[6 * t + 2 * r] * (ha * nb) + cU[6 * t + 2 * r + 1] * (na * hb);
          } else {
            val = (na * nb) * (fa * cd[3 * t + i2] - fb * cd[3 * t + i1]);
          }
          J[t] = val;
        }
      }
    }
    if ((kRes || kGeo || MT::kOut) && lane >= 24 && lane < 28) {
      const int iq = lane - 24;
      double u, v, N[4], Na[4], Nb[4];
      if (MT::kOut)
        node_point(iq, u, v);
      else
        quad_point(iq, u, v);
      shape4(u, v, N, Na, Nb);
      double e[9], Nt[9], g[5], ety[9];
      tying_interp(u, v, Nt);
#pragma unroll
      for (int t = 0; t < 9; t++) ety[t] = sm.ety[t];
      tying_fields(Nt, ety, g);
      const double *Wq = &sm.Wf[25 * iq];
#pragma unroll
      for (int r = 0; r < 5; r++) {
        double acc = 0.0;
#pragma unroll
        for (int f = 0; f < 5; f++) acc += Wq[5 * r + f] * g[f];
        e[tying_row(r)] = acc;
      }
      const double *U0 = &sm.u0x[9 * iq], *U1 = &sm.u1x[9 * iq];
      e[3] = U1[0];
      e[4] = U1[4];
      e[5] = U1[1] + U1[3];
      if (kNonlin) {
        e[3] += U0[0] * U1[0] + U0[3] * U1[3] + U0[6] * U1[6];
        e[4] += U0[1] * U1[1] + U0[4] * U1[4] + U0[7] * U1[7];
        e[5] += U0[0] * U1[1] + U0[3] * U1[4] + U0[6] * U1[7] + U1[0] * U0[1] + U1[3] * U0[4] +
                U1[6] * U0[7];
      }
      e[8] = N[0] * sm.etn[0] + N[1] * sm.etn[1] + N[2] * sm.etn[2] + N[3] * sm.etn[3];
      if constexpr (MT::kOut) {
        // getOutputData (TACSShellElement.h:1402-1500): total strain, stress
        // s = C e (TACSIsoShellConstitutive::evalStress, no thermal strain),
        // failure0 = evalFailure (TACSIsoShellConstitutive.cpp:254-297: KS,
        // weight 100, of the top and bottom von Mises indices), dv1 = t
        double s9[9];
        apply_c(sec, 1.0, e, s9);
        const double Q0 = a.mat.q11, Q1 = a.mat.q12, Q5 = a.mat.q66;
        const double ht = 0.5 * thick, hb = -0.5 * thick;
        double et[3], eb[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
          et[i] = e[i] + ht * e[3 + i];
          eb[i] = e[i] + hb * e[3 + i];
        }
        const double st0 = Q0 * et[0] + Q1 * et[1], st1 = Q1 * et[0] + Q0 * et[1], st2 = Q5 * et[2];
        const double sb0 = Q0 * eb[0] + Q1 * eb[1], sb1 = Q1 * eb[0] + Q0 * eb[1], sb2 = Q5 * eb[2];
        const double top = sqrt(st0 * st0 + st1 * st1 - st0 * st1 + 3.0 * st2 * st2) / a.ys;
        const double bottom = sqrt(sb0 * sb0 + sb1 * sb1 - sb0 * sb1 + 3.0 * sb2 * sb2) / a.ys;
        const double mx = top > bottom ? top : bottom;
        const double fail = mx + log(exp(100.0 * (top - mx)) + exp(100.0 * (bottom - mx))) / 100.0;
        if (valid) {
          double *o = obase + 31 * iq;
#pragma unroll
          for (int i = 0; i < 9; i++) {
            o[i] = e[i];
            o[9 + i] = s9[i];
          }
#pragma unroll
          for (int i = 18; i < 31; i++) o[i] = 0.0;
          o[18] = fail;
          o[24] = thick;
        }
      } else {
        // mechanical strain (TACSIsoShellConstitutive.cpp:438-456)
        e[0] -= a.mat.cte * temp;
        e[1] -= a.mat.cte * temp;
        double s9[9];
        if (MODE == kModeAdjResLin || MODE == kModeAdjResNl) {
#pragma unroll
          for (int i = 0; i < 9; i++) s9[i] = e[i];  // mechanical strain
        } else {
          apply_c(sec, sm.wdet[iq], e, s9);
        }
#pragma unroll
        for (int i = 0; i < 9; i++) sm.sig[9 * iq + i] = s9[i];
      }
    }
  }
  elem_sync<EPW>();

  if (MT::kOut) return;

  // ---- P4: stress resultants at the tying points ----
  if (kGeo && lane >= 8 && lane < 17) {
    const int t = lane - 8, f = tying_field(t);
    double acc = 0.0;
#pragma unroll
    for (int iq = 0; iq < 4; iq++) {
      double u, v, Nt[9];
      quad_point(iq, u, v);
      tying_interp(u, v, Nt);
      double w = 0.0;
#pragma unroll
      for (int r = 0; r < 5; r++) w += sm.sig[9 * iq + tying_row(r)] * sm.Wf[25 * iq + 5 * r + f];
      acc += Nt[t] * w;
    }
    sm.dety[t] = acc;
  }

  // ---- thickness derivative of phi^T K phi and phi^T G(u) phi, and the
  // path-adjoint right-hand side d(phi^T G(u) phi)/du ----
  // With e(q) = B0 q + Q(q, q)/2 (Q symmetric bilinear, Bq(a) b = Q(a, b)):
  //   phi^T K phi    = sum w e1(phi)^T C e1(phi),            e1 = B0 .
  //   phi^T G(u) phi = sum w [2 Q(u, phi)^T C e1(phi) + (e1(u) - e_th)^T C Q(phi, phi)]
  if constexpr (MODE == kModeAdjDV) {
    const IsoSection dsec = iso_section(make_section_dt(a.mat, thick));
    const double *su = sm.v2;
    double *E = &sm.c.Y[0][0], *cv = &sm.c.Y[0][45];
    const double ca = a.coef[2 * design], cb = a.coef[2 * design + 1];
    double gdv = 0.0, gu = 0.0;
    for (int iq = 0; iq < 4; iq++) {
      if (lane < 24) {
        double b[9];
        build_bcol<kBLin>(sm, iq, lane, Jr[0], b);
#pragma unroll
        for (int i = 0; i < 9; i++) sm.c.B[0][24 * i + lane] = b[i];
        build_bcol<kBQuad>(sm, iq, lane, Jr[kNB - 1], b);
#pragma unroll
        for (int i = 0; i < 9; i++) sm.c.B[1][24 * i + lane] = b[i];
      }
      elem_sync<EPW>();
      // E = [e1(phi), e1(u), Q(phi, u), Q(phi, phi)]
      for (int it = lane; it < 36; it += 32) {
        const int v = it / 9, i = it % 9;
        const double *M = (v < 2) ? &sm.c.B[0][24 * i] : &sm.c.B[1][24 * i];
        const double *x = (v == 0 || v == 3) ? sm.s : su;
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 24; k++) acc += M[k] * x[k];
        E[it] = acc;
      }
      elem_sync<EPW>();
      if (lane == 0) {
        const double w = sm.wdet[iq];
        const double *e1p = &E[0], *e1u = &E[9], *e2up = &E[18], *e2pp = &E[27];
        double eu[9], t1[9], t2[9], c1[9], c2[9];
#pragma unroll
        for (int i = 0; i < 9; i++) eu[i] = e1u[i];
        eu[0] -= a.mat.cte * temp;
        eu[1] -= a.mat.cte * temp;
        apply_c(dsec, w, e1p, t1);  // w C' e1(phi)
        apply_c(dsec, w, eu, t2);   // w C' (e1(u) - e_th)
        double da = 0.0, db = 0.0;
#pragma unroll
        for (int i = 0; i < 9; i++) {
          da += e1p[i] * t1[i];
          db += 2.0 * e2up[i] * t1[i] + e2pp[i] * t2[i];
        }
        gdv += ca * da + cb * db;
        apply_c(sec, w, e1p, c1);
        apply_c(sec, w, e2pp, c2);
#pragma unroll
        for (int i = 0; i < 9; i++) {
          cv[i] = 2.0 * cb * c1[i];
          cv[9 + i] = cb * c2[i];
        }
      }
      elem_sync<EPW>();
      // d/du: Bq(phi)^T (2 w C e1(phi)) + B0^T (w C Q(phi, phi))
      if (lane < 24) {
#pragma unroll
        for (int i = 0; i < 9; i++) gu += sm.c.B[1][24 * i + lane] * cv[i] + sm.c.B[0][24 * i + lane] * cv[9 + i];
      }
      elem_sync<EPW>();
    }
    if (valid && lane == 0 && a.dv) a.dv[design * a.nelems + e] += gdv;
    if (valid && lane < 24 && a.res) a.res[design * a.res_stride + 6 * cn[lane / 6] + lane % 6] += gu;
    return;
  }

  // ---- psi^T dR_e/dt (TACSShellElement::addAdjResProduct, SE.h:774-868,
  // with TACSIsoShellConstitutive::addStressDVSens) ----
  if constexpr (MODE == kModeAdjResLin || MODE == kModeAdjResNl) {
    const IsoSection dsec = iso_section(make_section_dt(a.mat, thick));
    const double *sp = sm.v2;
    double *E = &sm.c.Y[0][0];
    double g = 0.0;
    for (int iq = 0; iq < 4; iq++) {
      if (lane < 24) {
        double b[9];
        if (kNonlin)
          build_bcol<kBFull>(sm, iq, lane, Jr[0], b);
        else
          build_bcol<kBLin>(sm, iq, lane, Jr[0], b);
#pragma unroll
        for (int i = 0; i < 9; i++) sm.c.B[0][24 * i + lane] = b[i];
      }
      elem_sync<EPW>();
      if (lane < 9) {  // dR/dq^T psi strain: B(u) psi
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 24; k++) acc += sm.c.B[0][24 * lane + k] * sp[k];
        E[lane] = acc;
      }
      elem_sync<EPW>();
      if (lane == 0) {
        double t2[9];
        apply_c(dsec, sm.wdet[iq], &sm.sig[9 * iq], t2);  // sig holds e(u) - e_th
        double d = 0.0;
#pragma unroll
        for (int i = 0; i < 9; i++) d += E[i] * t2[i];
        g += d;
      }
      elem_sync<EPW>();
    }
    if (valid && lane == 0) a.dv[design * a.nelems + e] += a.coef[design] * g;
    return;
  }

  // ---- P5: quadrature loop: B columns, residual, B^T C B tile ----
  // lane tile: rows row0..row0+2 (one node's translations or rotations),
  // columns col0..col0+5 (all dofs of node col0 / 6)
  const int row0 = 3 * (lane >> 2), col0 = 6 * (lane & 3);
  // the lane's 18 entries of its BSR block, read now so that the read
  // completes during the quadrature loop (no other element of this launch
  // writes the block): 6 % faster than reading them at the scatter for the
  // linear stiffness of the 200 x 400 cylinder, more for the nonlinear modes
  // (schedule bit 0 reads them at the scatter instead: fewer live registers
  // through the loop)
  constexpr bool kEarly = kMat && !MT::kJvp && !(VAR & 1);
  double prev[18];
  double *dst = nullptr;
  // first touch of the block in this assembly: written, not added to (the
  // matrix is not zeroed first)
  bool fresh = false;
  if (kMat && !MT::kJvp && valid) {
    const int slot = 4 * (row0 / 6) + col0 / 6;
    const long blk = a.blk[16 * e + slot];
    dst = a.vals + design * a.vals_stride + 36 * blk + 6 * (row0 % 6);
    fresh = a.first && ((a.first[e] >> slot) & 1);
    if (kEarly) {
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) prev[6 * i + j] = dst[6 * i + j];
    }
  }
  double acc[18];
#pragma unroll
  for (int i = 0; i < 18; i++) acc[i] = 0.0;
  double racc = 0.0;

  for (int iq = 0; iq < 4; iq++) {
    if (lane < 24) {
      double b[9], y[9];
      if (kNonlin) {
        build_bcol<kBFull>(sm, iq, lane, Jr[0], b);
      } else {
        build_bcol<kBLin>(sm, iq, lane, Jr[0], b);
      }
      if (kRes) {
#pragma unroll
        for (int i = 0; i < 9; i++) racc += b[i] * sm.sig[9 * iq + i];
      }
      if (kMat) {
        apply_c(sec, sm.wdet[iq], b, y);
#pragma unroll
        for (int i = 0; i < 9; i++) {
          sm.c.B[0][24 * i + lane] = b[i];
          sm.c.Y[0][24 * i + lane] = y[i];
        }
      }
      if constexpr (MODE == kModeG) {
        build_bcol<kBQuad>(sm, iq, lane, Jr[kNB - 1], b);
        apply_c(sec, sm.wdet[iq], b, y);
#pragma unroll
        for (int i = 0; i < 9; i++) {
          sm.c.B[kTwo ? 1 : 0][24 * i + lane] = b[i];
          sm.c.Y[kTwo ? 1 : 0][24 * i + lane] = y[i];
        }
      }
    }
    if (kMat) {
      elem_sync<EPW>();
#pragma unroll
      for (int s = 0; s < 9; s++) {
        double b[3], y[6];
#pragma unroll
        for (int i = 0; i < 3; i++) b[i] = sm.c.B[0][24 * s + row0 + i];
        if constexpr (MODE == kModeG) {
          // B0^T C B1 + B1^T C B0
          double c[3], z[6];
#pragma unroll
          for (int i = 0; i < 3; i++) c[i] = sm.c.B[kTwo ? 1 : 0][24 * s + row0 + i];
#pragma unroll
          for (int j = 0; j < 6; j++) {
            y[j] = sm.c.Y[kTwo ? 1 : 0][24 * s + col0 + j];
            z[j] = sm.c.Y[0][24 * s + col0 + j];
          }
#pragma unroll
          for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 6; j++) acc[6 * i + j] += b[i] * y[j] + c[i] * z[j];
        } else {
#pragma unroll
          for (int j = 0; j < 6; j++) y[j] = sm.c.Y[0][24 * s + col0 + j];
#pragma unroll
          for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 6; j++) acc[6 * i + j] += b[i] * y[j];
        }
      }
      elem_sync<EPW>();
    }
  }

  // ---- P6: geometric (stress) terms ----
  if (kGeo) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 6; j++) acc[6 * i + j] += tying_geo(sm, row0 + i, col0 + j);
    for (int iq = 0; iq < 4; iq++) {
      double u, v, N[4], Na[4], Nb[4];
      quad_point(iq, u, v);
      shape4(u, v, N, Na, Nb);
      const double S00 = sm.sig[9 * iq + 3], S11 = sm.sig[9 * iq + 4], S01 = sm.sig[9 * iq + 5];
      double ar[3][3], br[3][3], gr[3][3];
#pragma unroll
      for (int i = 0; i < 3; i++) dof_abg(sm, iq, row0 + i, N, Na, Nb, ar[i], br[i], gr[i]);
#pragma unroll
      for (int j = 0; j < 6; j++) {
        double ac[3], bc[3], gc[3];
        dof_abg(sm, iq, col0 + j, N, Na, Nb, ac, bc, gc);
#pragma unroll
        for (int i = 0; i < 3; i++) {
          const double aa = ar[i][0] * ac[0] + ar[i][1] * ac[1] + ar[i][2] * ac[2];
          const double h = S00 * (br[i][0] * gc[0] + bc[0] * gr[i][0]) +
                           S11 * (br[i][1] * gc[1] + bc[1] * gr[i][1]) +
                           S01 * (br[i][0] * gc[1] + bc[0] * gr[i][1] + br[i][1] * gc[0] +
                                  bc[1] * gr[i][0]);
          acc[6 * i + j] += aa * h;
        }
      }
    }
  }

  // ---- inertial terms of addJacobian (TACSShellElement.h:613-648 and the
  // gamma d2Tdotd part of TACSLinearizedRotation::addDirectorJacobian,
  // TACSDirector.h:368-430): gamma M into the matrix and M u'' into the
  // residual, with the mass moments m0 = rho t, m1 = 0, m2 = rho t^3 / 12
  // (TACSIsoShellConstitutive.cpp:120-129).  The residual is not scaled by
  // alpha or gamma. ----
  if constexpr (MT::kMassTerms) {
    const double m0 = a.mat.rho * thick, m2 = a.mat.rho * thick * thick * thick * (1.0 / 12.0);
    if (MT::kJac && a.ddu && lane < 24) {
      const int n = lane / 6, r = lane % 6;
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const MassTile mt = mass_tile(sm.wdet, sm.fn, n, m, r < 3 ? 0 : 3, m0, m2);
#pragma unroll
        for (int j = 0; j < 6; j++) racc += mass_entry(mt, r % 3, j) * sm.dd[6 * m + j];
      }
    }
    // the tile becomes alpha K_e + gamma M_e
    const MassTile mt = mass_tile(sm.wdet, sm.fn, row0 / 6, col0 / 6, row0 % 6, m0, m2);
    const double gamma = a.gamma;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 6; j++) acc[6 * i + j] = alpha * acc[6 * i + j] + gamma * mass_entry(mt, i, j);
  }

  // ---- element matrix times a vector instead of the scatter
  // (TACSAssembler::addJacobianVecProduct, TACSAssembler.cpp:4334-4390) ----
  if constexpr (MT::kJvp) {
    const double *x = a.mvx + design * a.mv_stride;
    double *y = a.mvy + design * a.mv_stride;
    const int nr = cn[row0 / 6], nc = cn[col0 / 6], rr = row0 % 6;
    if (!a.mv_trans) {
      // rows row0..row0+2: sum over the 4 column nodes (lanes 4k..4k+3)
      double xc[6], p[3];
#pragma unroll
      for (int j = 0; j < 6; j++) xc[j] = x[6 * nc + j];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        p[i] = 0.0;
#pragma unroll
        for (int j = 0; j < 6; j++) p[i] += acc[6 * i + j] * xc[j];
      }
#pragma unroll
      for (int i = 0; i < 3; i++) {
        p[i] += __shfl_xor(p[i], 1);
        p[i] += __shfl_xor(p[i], 2);
      }
      if (valid && (lane & 3) == 0) {
#pragma unroll
        for (int i = 0; i < 3; i++) y[6 * nr + rr + i] += a.mv_scale * p[i];
      }
    } else {
      // columns of node col0 / 6: sum over the 8 row tiles (lanes with the
      // same lane & 3)
      double xr[3], p[6];
#pragma unroll
      for (int i = 0; i < 3; i++) xr[i] = x[6 * nr + rr + i];
#pragma unroll
      for (int j = 0; j < 6; j++) {
        p[j] = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) p[j] += acc[6 * i + j] * xr[i];
      }
#pragma unroll
      for (int j = 0; j < 6; j++) {
        p[j] += __shfl_xor(p[j], 4);
        p[j] += __shfl_xor(p[j], 8);
        p[j] += __shfl_xor(p[j], 16);
      }
      if (valid && lane < 4) {
#pragma unroll
        for (int j = 0; j < 6; j++) y[6 * nc + j] += a.mv_scale * p[j];
      }
    }
    return;
  }

  // ---- P7: scatter (elements of one colour share no node) ----
  if (kMat && valid) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 6; j++) {
        // read even where the value is discarded: a load only for blocks
        // already reached measured slower than loading them all
        const double old = kEarly ? prev[6 * i + j] : dst[6 * i + j];
        // (the mass-term modes have applied alpha already)
        dst[6 * i + j] = (fresh ? 0.0 : old) + (MT::kMassTerms ? acc[6 * i + j] : alpha * acc[6 * i + j]);
      }
  }
  // the residual is unscaled (alpha scales only the matrix)
  if (kRes && valid && lane < 24 && a.res) {
    double *dst = a.res + design * a.res_stride;
    dst[6 * cn[lane / 6] + lane % 6] += racc;
  }
}

template <int MODE, int VAR, int EPW>
__global__ __launch_bounds__(32 * EPW) void elem_kernel(ElemArgs a) {
  elem_body<MODE, VAR, EPW>(a);
}

// the same body within the register budget of two wavefronts per SIMD (the
// compiler spills to scratch memory instead of holding one wavefront with
// accumulation registers)
template <int MODE, int VAR, int EPW>
__global__ __launch_bounds__(32 * EPW) __attribute__((amdgpu_waves_per_eu(2))) void elem_kernel_w2(ElemArgs a) {
  elem_body<MODE, VAR, EPW>(a);
}

// ---- mass matrix ----
// getMatType(TACS_MASS_MATRIX) (TACSShellElement.h:675-704, gamma = 1) keeps
// only the inertial terms of addJacobian (TACSShellElement.h:613-648) and the
// gamma d2Tdotd part of TACSLinearizedRotation::addDirectorJacobian
// (TACSDirector.h:368-430).  With the mass moments of TACSIsoShellConstitutive
// (TACSIsoShellConstitutive.cpp:120-129, tOffset = 0) m0 = rho t, m1 = 0,
// m2 = rho t^3 / 12, and S_ij = sum_qp detXd N_i N_j:
//   M[u_i, u_j] = m0 S_ij I
//   M[q_i, q_j] = m2 S_ij ((t_i . t_j) I - t_j t_i^T)
// since the director d = q x t moves by -[t]x dq and [t_i]x^T [t_j]x =
// (t_i . t_j) I - t_j t_i^T (t = node normals).  Translations and rotations do
// not couple (m1 = 0).  kModeMassDV contracts dM/dt (dm0/dt = rho,
// dm2/dt = rho t^2 / 4) with the state phi: dv[e] += coef[d] phi^T dM_e/dt phi.
// Same layout as elem_kernel: two elements per wavefront, a 3 x 6 tile per lane.
template <int MODE>
__global__ __launch_bounds__(64) void mass_kernel(ElemArgs a) {
  struct Smem {
    double X[12], fn[12], wdet[4], part[32];
  };
  __shared__ Smem smem[2];
  const int half = threadIdx.x >> 5, lane = threadIdx.x & 31;
  Smem &sm = smem[half];
  const int slot = 2 * blockIdx.x + half;
  const bool valid = slot < a.count;
  const int e = a.elems[valid ? slot : slot - 1];
  const long design = blockIdx.y;
  const int *cn = a.conn + 4 * e;
  const double t = a.thick[design * a.nelems + e];

  if (lane < 12) sm.X[lane] = a.Xpts[3 * cn[lane / 3] + lane % 3];
  __syncthreads();
  // node normals (TacsShellComputeNodeNormals, TACSShellUtilities.h:301-343)
  if (lane < 4) {
    double u, v, N[4], Na[4], Nb[4];
    node_point(lane, u, v);
    shape4(u, v, N, Na, Nb);
    double Xxi[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int d = 0; d < 3; d++) {
        Xxi[2 * d] += Na[m] * sm.X[3 * m + d];
        Xxi[2 * d + 1] += Nb[m] * sm.X[3 * m + d];
      }
    const double a3[3] = {Xxi[0], Xxi[2], Xxi[4]}, b3[3] = {Xxi[1], Xxi[3], Xxi[5]};
    double f[3];
    cross3(a3, b3, f);
    const double nrm = sqrt(dot3(f, f));
    if (nrm != 0.0) {
      const double inv = 1.0 / nrm;
      f[0] *= inv;
      f[1] *= inv;
      f[2] *= inv;
    }
    sm.fn[3 * lane] = f[0];
    sm.fn[3 * lane + 1] = f[1];
    sm.fn[3 * lane + 2] = f[2];
  }
  __syncthreads();
  // detXd at the quadrature points (Gauss weight 1)
  if (lane < 4) {
    double u, v, N[4], Na[4], Nb[4];
    quad_point(lane, u, v);
    shape4(u, v, N, Na, Nb);
    double Xxi[6] = {0, 0, 0, 0, 0, 0}, n0[3] = {0, 0, 0};
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int d = 0; d < 3; d++) {
        Xxi[2 * d] += Na[m] * sm.X[3 * m + d];
        Xxi[2 * d + 1] += Nb[m] * sm.X[3 * m + d];
        n0[d] += N[m] * sm.fn[3 * m + d];
      }
    double Xd[9], Xdi[9];
    frame3(Xxi, n0, Xd);
    sm.wdet[lane] = inv3(Xd, Xdi);
  }
  __syncthreads();

  const int row0 = 3 * (lane >> 2), col0 = 6 * (lane & 3);
  const int ni = row0 / 6, nj = col0 / 6, rr = row0 % 6;
  const double rho = a.mat.rho;
  const double m0 = (MODE == kModeMass) ? rho * t : rho;
  const double m2 = (MODE == kModeMass) ? rho * t * t * t * (1.0 / 12.0) : rho * t * t * 0.25;
  const MassTile mt = mass_tile(sm.wdet, sm.fn, ni, nj, rr, m0, m2);
  double acc[18];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 6; j++) acc[6 * i + j] = mass_entry(mt, i, j);

  if (MODE == kModeMass) {
    if (valid) {
      const long blk = a.blk[16 * e + 4 * ni + nj];
      double *dst = a.vals + design * a.vals_stride + 36 * blk;
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) dst[6 * (rr + i) + j] += a.alpha * acc[6 * i + j];
    }
  } else {
    const double *phi = a.state + design * a.state_stride;
    double pr[3], pc[6];
#pragma unroll
    for (int i = 0; i < 3; i++) pr[i] = phi[6 * cn[ni] + rr + i];
#pragma unroll
    for (int j = 0; j < 6; j++) pc[j] = phi[6 * cn[nj] + j];
    double p = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 6; j++) p += pr[i] * acc[6 * i + j] * pc[j];
    sm.part[lane] = p;
    __syncthreads();
    if (lane == 0 && valid) {
      double sum = 0.0;
      for (int k = 0; k < 32; k++) sum += sm.part[k];
      a.dv[design * a.nelems + e] += a.coef[design] * sum;
    }
  }
}

// explicit instantiation + launch wrappers
template <int MODE>
static hipError_t launch_mode(const ElemArgs &args, int count, int nbatch, hipStream_t stream) {
  if (count <= 0 || nbatch <= 0) return hipSuccess;
  ElemArgs a = args;
  a.count = count;
  hipLaunchKernelGGL((elem_kernel<MODE, 0, 2>), dim3((count + 1) / 2, nbatch), dim3(64), 0, stream, a);
  return hipGetLastError();
}

// elements per workgroup of the wide schedules: 4 (two wavefronts).  On the
// 200x400 cylinder the stiffness took 90.2 us per colour launch with 4
// against 91.9 with 8 (four wavefronts, 51 KB of LDS); the tangent, G and
// the Jacobians, whose LDS for 8 would not fit twice per CU, take 4 either
// way.
template <int MODE>
constexpr int wide_epw() {
  return 4;
}

template <int MODE, int EPW>
static void launch_var(int variant, const dim3 &grid, hipStream_t stream, const ElemArgs &a) {
  const dim3 block(32 * EPW);
  switch (variant & 3) {
    case 1:
      hipLaunchKernelGGL((elem_kernel<MODE, 1, EPW>), grid, block, 0, stream, a);
      break;
    case 2:
      hipLaunchKernelGGL((elem_kernel_w2<MODE, 0, EPW>), grid, block, 0, stream, a);
      break;
    case 3:
      hipLaunchKernelGGL((elem_kernel_w2<MODE, 1, EPW>), grid, block, 0, stream, a);
      break;
    default:
      hipLaunchKernelGGL((elem_kernel<MODE, 0, EPW>), grid, block, 0, stream, a);
  }
}

// the matrix modes in every schedule: variant bit 0 = VAR bit 0 (BSR block
// reads at the scatter instead of before the quadrature loop), bit 1 = the
// two-wavefront register budget, bit 2 = wide workgroups (wide_epw elements,
// the geometry tasks spread over the workgroup) instead of two elements
template <int MODE>
static hipError_t launch_tuned(const ElemArgs &args, int count, int nbatch, int variant, hipStream_t stream) {
  if (count <= 0 || nbatch <= 0) return hipSuccess;
  ElemArgs a = args;
  a.count = count;
  if (variant & 4) {
    constexpr int W = wide_epw<MODE>();
    launch_var<MODE, W>(variant, dim3((count + W - 1) / W, nbatch), stream, a);
  } else {
    launch_var<MODE, 2>(variant, dim3((count + 1) / 2, nbatch), stream, a);
  }
  return hipGetLastError();
}

hipError_t launch_elem_kernel(int mode, int variant, const ElemArgs &args, int count, int nbatch,
                              hipStream_t stream) {
  switch (mode) {
    case kModeKLin:
      return launch_tuned<kModeKLin>(args, count, nbatch, variant, stream);
    case kModeKNl:
      return launch_tuned<kModeKNl>(args, count, nbatch, variant, stream);
    case kModeG:
      return launch_tuned<kModeG>(args, count, nbatch, variant, stream);
    case kModeRLin:
      return launch_mode<kModeRLin>(args, count, nbatch, stream);
    case kModeRNl:
      return launch_mode<kModeRNl>(args, count, nbatch, stream);
    case kModeAdjDV:
      return launch_mode<kModeAdjDV>(args, count, nbatch, stream);
    case kModeAdjResLin:
      return launch_mode<kModeAdjResLin>(args, count, nbatch, stream);
    case kModeAdjResNl:
      return launch_mode<kModeAdjResNl>(args, count, nbatch, stream);
    case kModeOutLin:
      return launch_mode<kModeOutLin>(args, count, nbatch, stream);
    case kModeOutNl:
      return launch_mode<kModeOutNl>(args, count, nbatch, stream);
    case kModeJacLin:
      return launch_tuned<kModeJacLin>(args, count, nbatch, variant, stream);
    case kModeJacNl:
      return launch_tuned<kModeJacNl>(args, count, nbatch, variant, stream);
    case kModeJvpLin:
      return launch_mode<kModeJvpLin>(args, count, nbatch, stream);
    case kModeJvpNl:
      return launch_mode<kModeJvpNl>(args, count, nbatch, stream);
    case kModeMass:
    case kModeMassDV: {
      if (count <= 0 || nbatch <= 0) return hipSuccess;
      ElemArgs m = args;
      m.count = count;
      const dim3 grid((count + 1) / 2, nbatch);
      if (mode == kModeMass)
        hipLaunchKernelGGL(mass_kernel<kModeMass>, grid, dim3(64), 0, stream, m);
      else
        hipLaunchKernelGGL(mass_kernel<kModeMassDV>, grid, dim3(64), 0, stream, m);
      return hipGetLastError();
    }
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace shellgpu
