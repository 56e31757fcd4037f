// shellgpu_internal.h -- types shared by the HIP kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

namespace shellgpu {

// Isotropic material (TACSMaterialProperties, TACSMaterialProperties.cpp:24-63)
// with its plane-stress stiffness Q = [[q11 q12 0][q12 q11 0][0 0 q66]]
// (evalTangentStiffness2D, TACSMaterialProperties.cpp:295-313) formed once
struct Material {
  double E, nu, rho, cte, kdrill;
  double q11, q12, q66;
};

inline Material make_material(double E, double nu, double rho, double cte, double kdrill) {
  const double G = 0.5 * E / (1.0 + nu);
  const double Dm = E / (1.0 - nu * nu);
  return Material{E, nu, rho, cte, kdrill, Dm, nu * Dm, G};
}

// Section stiffness of TACSIsoShellConstitutive with tOffset = 0
// (TACSIsoShellConstitutive.cpp:192-226): A = t Q, B = 0, D = t^3/12 Q,
// As = 5/6 G t I, drill = 0.5 kdrill (As0 + As2).
struct Section {
  double A[6], D[6], As[3], drill;
};

__host__ __device__ inline Section make_section(const Material &m, double t) {
  Section s;
  const double Q[6] = {m.q11, m.q12, 0.0, m.q11, 0.0, m.q66};
  const double I = t * t * t / 12.0;
  for (int i = 0; i < 6; i++) {
    s.D[i] = I * Q[i];
    s.A[i] = Q[i] * t;
  }
  s.As[0] = s.As[2] = (5.0 / 6.0) * s.A[5];
  s.As[1] = 0.0;
  s.drill = 0.5 * m.kdrill * (s.As[0] + s.As[2]);
  return s;
}

enum ElemMode {
  kModeKLin = 0,  // linear-model stiffness (getMatType STIFFNESS, TACSQuad4Shell)
  kModeKNl = 1,   // nonlinear-model tangent at the state
  kModeG = 2,     // geometric stiffness, exact directional derivative
  kModeRLin = 3,  // linear-model residual
  kModeRNl = 4,   // nonlinear-model residual
  kModeAdjDV = 5,     // thickness terms of a buckling eigenpair (state = phi)
  kModeAdjResLin = 6, // psi^T dR/dt, linear model (state = u)
  kModeAdjResNl = 7,  // psi^T dR/dt, nonlinear model (state = u)
  kModeMass = 8,      // mass matrix (getMatType MASS)
  kModeMassDV = 9,    // dv[e] += coef[d] phi^T dM_e/dt phi (state = phi)
  kModeOutLin = 10,   // output data at the element nodes, linear model (getOutputData)
  kModeOutNl = 11,    // the same, nonlinear model
  kModeJacLin = 12,   // addJacobian(alpha, beta, gamma): R and alpha K + gamma M, linear model
  kModeJacNl = 13,    // the same, nonlinear model (alpha K_t)
  kModeJvpLin = 14,   // (alpha K + gamma M) x per element, nothing assembled, linear model
  kModeJvpNl = 15,    // the same, nonlinear model
};

struct ElemArgs {
  const int *elems;  // element ids of this launch (one colour)
  int count;         // number of them (set by the launcher)
  const int *conn;   // 4 node ids per element
  const int *blk;    // 16 BSR block ids per element, (row node i, col node j)
  const double *Xpts;
  const double *state;  // 6 per node, may be null
  long state_stride;    // per design
  const double *thick;  // per element, stride nelems per design
  const double *temp;   // per element, may be null
  int nelems;
  Material mat;
  double alpha;
  // kModeJac*: coefficient of the mass matrix and the second time
  // derivative of the state (may be null: zero), 6 per node
  double gamma;
  const double *ddu;
  long ddu_stride;
  int fd_sign;   // kModeKNl only: +-1 selects the reference's central-difference
                 // geometric stiffness step (TACSShellElement.h:705-766), 0 = off
  double *vals;  // BSR values
  long vals_stride;
  // matrix modes: per element the blocks it writes rather than adds to
  // (first-touch bits, see shellgpu_ctx::d_first); null: add to every block
  const int *first;
  double *res;
  long res_stride;
  // thickness-derivative modes (' = d/dt_e, all outputs accumulate):
  //  kModeAdjDV (state = phi, u2 = path u):
  //    dv[e] += coef[2d] phi^T K' phi + coef[2d+1] phi^T G'(u) phi
  //    res   += coef[2d+1] d(phi^T G(u) phi)/du
  //  kModeAdjRes* (state = u): dv[e] += coef[d] psi^T dR_e/dt
  const double *u2;
  const double *psi;
  long u2_stride, psi_stride;
  const double *coef;
  double *dv;
  // output modes: out[((d nelems + e) 4 + node) 31 + k], ys = yield stress
  double *out;
  double ys;
  // kModeJvp*: no matrix is written; instead mvy += mv_scale A_e mvx
  // (mv_trans = 0) or A_e^T mvx (1) per element, A_e = alpha K_e + gamma
  // M_e (6 per node, stride mv_stride per design)
  const double *mvx;
  double *mvy;
  long mv_stride;
  double mv_scale;
  int mv_trans;
};

// d/dt of the section stiffness (A' = Q, D' = t^2/4 Q, As' = 5/6 G,
// drill' = kdrill 5/6 G), the derivative of make_section
__host__ __device__ inline Section make_section_dt(const Material &m, double t) {
  Section s;
  const double G = m.q66;
  const double Q[6] = {m.q11, m.q12, 0.0, m.q11, 0.0, m.q66};
  for (int i = 0; i < 6; i++) {
    s.A[i] = Q[i];
    s.D[i] = 0.25 * t * t * Q[i];
  }
  s.As[0] = s.As[2] = (5.0 / 6.0) * G;
  s.As[1] = 0.0;
  s.drill = 0.5 * m.kdrill * (s.As[0] + s.As[2]);
  return s;
}

// coordinate-derivative kernel: dfdX[3 node + x] += d/dX of
// coef[4d] phi^T K_e phi + coef[4d+1] phi^T G_e(u) phi + coef[4d+2] psi^T R_e(u)
// + coef[4d+3] phi^T M_e phi (residual of the nonlinear model when
// nonlin != 0); phi, u, psi may be null.  terms (kXpt* bits) names the
// scalars whose coefficients may be non-zero: the launch evaluates only the
// vectors and bilinear pairs those need.
enum { kXptK = 1, kXptG = 2, kXptR = 4, kXptM = 8 };
struct XptArgs {
  const int *elems;
  const int *conn;
  const double *Xpts;
  const double *thick;
  const double *temp;
  int nelems;
  Material mat;
  const double *phi, *u, *psi;
  long stride;
  const double *coef;
  int nonlin;
  double *dfdX;
  long dX_stride;
  int terms;
};
hipError_t launch_xpt_kernel(const XptArgs &args, int nel, int nbatch, hipStream_t stream);

// variant: the compiled schedule of the matrix modes (kModeKLin, kModeKNl,
// kModeG, kModeJac*): bit 0 reads the BSR blocks at the scatter (default:
// before the quadrature loop), bit 1 two-wavefront register budget, bit 2
// wide workgroups (4 elements, the geometry spread over them, instead of
// 2); other modes ignore it
constexpr int kElemVariants = 8;
hipError_t launch_elem_kernel(int mode, int variant, const ElemArgs &args, int count, int nbatch,
                              hipStream_t stream);

// Batched FP64 GEMM on the matrix cores (gemm_f64.hip): C_b = alpha op(A_b)
// op(B_b) + beta C_b, column-major, op(A) M x K, op(B) K x N; entry b's
// matrices from the pointer arrays Ap/Bp/Cp when set, else at A + b sA etc.
// Triangular operands: tri_a = 1 (2) takes op(A) as lower (upper)
// triangular, tri_b likewise for op(B) -- the other triangle is not read and
// counts as zero, and the products skip its K ranges; lower_c = 1 computes
// only the tiles of C that hold entries on or below the diagonal (entries
// above it in those tiles are written too: C's upper triangle is scratch).
enum { kGemmFull = 0, kGemmLower = 1, kGemmUpper = 2 };
struct GemmArgs {
  int M, N, K;
  double alpha, beta;
  const double *A;
  long sA;
  const double *const *Ap;
  int lda;
  const double *B;
  long sB;
  const double *const *Bp;
  int ldb;
  double *C;
  long sC;
  double *const *Cp;
  int ldc;
  int tri_a = kGemmFull, tri_b = kGemmFull, lower_c = 0;
  // optional workspace (ws_len doubles) for products with at most 16
  // columns: a launch of few tiles splits K over workgroups into it and
  // adds the parts in a fixed order
  double *ws = nullptr;
  long ws_len = 0;
};
hipError_t launch_gemm_f64(bool transA, bool transB, const GemmArgs &g, int batch, hipStream_t stream);

}  // namespace shellgpu
