// solver.hip -- direct solver, GMRES and shift-invert Lanczos buckling.
//
// Direct solver.  The reference factors K and K + sigma G with an
// AMD-ordered complete block LU on the CPU (TACSSchurPc with levFill = 1e6,
// src/bpmat/TACSSchurMat.cpp:944-1652).  Here the nodes are ordered by
// breadth-first level sets from a seed set (a whole end ring of the
// cylinder, or a pseudo-peripheral node), which makes every shell mesh
// block-tridiagonal in level blocks: a node couples only to nodes of its own
// and the neighbouring levels.  Boundary conditions are eliminated
// symmetrically (BC rows and columns zero, unit diagonal), so the stiffness
// and the shifted operator K + sigma G (sigma below the first buckling
// load) are symmetric positive definite.  Two factorisations of the
// level-block-tridiagonal matrix (diagonal blocks D_l, couplings E_l of
// level l to level l - 1, which keep the mesh's sparsity: a node couples to
// a few nodes of the previous level):
//
// Sequential (block LDL^T), level by level, batched over designs, keeping
// the couplings sparse and the Schur-complement inverses explicit:
//
//   S_0 = D_0,  S_l = D_l - E_l P_{l-1} E_l^T,  P_l = S_l^{-1}
//
// where P_l comes from the Cholesky factor's inverse X_l = C_l^{-1}
// (P_l = X_l^T X_l) and E_l P E_l^T costs two sparse products instead of
// dense GEMMs.  A solve is a chain of sparse coupling products and
// bandwidth-bound symmetric matrix-vector products reading half of P_l:
//
//   forward  t_l = P_l (b_l - E_l t_{l-1})
//   backward x_l = t_l - P_l E_{l+1}^T x_{l+1}
//
// An explicit inverse has an error that grows with cond(S_l), not with its
// square root as the triangular inverses do: the solve error is ~1e-10
// relative on shell stiffness matrices of condition 1e9 (about ten times
// smaller than a partial-pivoting LU solve).
//
// Block cyclic reduction (below) for a few designs of long level chains.
//
// A failed Cholesky (non-positive pivot) reports a shift above the first
// eigenvalue.  The solve is exact for right-hand sides that vanish on BC
// dofs (all solves of the buckling path); GMRES corrects the remainder.
//
// GMRES.  Right-preconditioned restarted GMRES with Givens rotations
// (GMRES::solve, src/bpmat/KSM.cpp:785-955), FP64, preconditioned by the
// direct solver.
//
// Lanczos.  Shift-invert Lanczos on (K + sigma G)^{-1} K in the K inner
// product with full re-orthogonalisation (SEP::solve + EPBucklingShiftInvert,
// src/bpmat/GSEP.cpp:202-244, 485-599), run in lock-step for every design of
// the batch.  Re-orthogonalisation is classical Gram-Schmidt applied twice
// (two GEMV pairs per step against the stored K q_j), the cached K q_j
// removing the reference's one K-SpMV per inner product.  Ritz values are
// converted by lambda = -theta sigma / (1 - theta) and sorted by magnitude
// (SMALLEST_MAGNITUDE), with a residual-based stopping test (the reference's
// test is disabled, GSEP.cpp:754-789).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "shellgpu.h"
#include "shellgpu_ctx.h"

using namespace shellgpu;

// one round of block cyclic reduction: levels s, 3s, 5s, ... eliminated
struct FactorRound {
  int s, nE, nR;  // stride, levels eliminated, of which with a right neighbour
  long off;       // first pointer of this round in the device pointer arrays
};

// One product of a batch of dense-times-coupling products (first round of
// cyclic reduction): Y = beta Y + alpha X E^T (RIGHT: E couples level ylev
// to ylev - 1, Y's columns are ylev's nodes, X's columns ylev - 1's) or
// Y = beta Y + alpha X E (LEFT: E couples level ylev + 1 to ylev, Y's
// columns are ylev's nodes, X's columns ylev + 1's), rows 0 .. rows - 1;
// E points at the design's coupling blocks.  Transposes and copies of the
// same round use X, Y, rows and cols.
struct CoupleJob {
  const double *X;
  double *Y;
  const double *E;
  int rows, cols, ylev, pad;
};

// one step of the recursive Cholesky-inverse of a diagonal block, in
// execution order: a leaf (n2 == 0: factor and invert rows/cols [r0, r0 +
// n1) in LDS), the update after the leading half (first == 1) or the
// off-diagonal inverse after the trailing half (first == 0)
struct CholNode {
  int r0, n1, n2, first;
};

struct shellgpu_factor {
  shellgpu_ctx *ctx = nullptr;
  int nbatch = 0;
  int nlevels = 0;
  int n = 0;  // scalar dofs
  std::vector<int> m;        // dofs per level
  std::vector<long> voff;    // offset of level l in a level-ordered vector
  std::vector<int> noff;     // first level-ordered node of level l (nlevels + 1)
  std::vector<long> S_off;   // D_l -> P_l in a design's buffer (m_l x m_l)
  long E_off = 0;            // first coupling block (6 x 6, column-major) in a design's buffer
  long ncoup = 0;            // coupling blocks (level l to level l - 1)
  int *d_cp_rowp = nullptr;  // level-ordered node -> its coupling blocks to the previous level
  int *d_cp_col = nullptr;   // per coupling block: local index of the previous-level node
  int *d_ct_rowp = nullptr;  // level-ordered node -> coupling blocks from next-level nodes
  int *d_ct_src = nullptr;   // per entry: local index of the next-level node
  int *d_ct_blk = nullptr;   // per entry: the coupling block
  long per_design = 0;       // doubles per design
  int maxm = 0;
  long tmp_elems = 0;        // largest m_l * m_{l-1}
  double *buf = nullptr;     // nbatch * per_design
  double *tmp = nullptr;     // 2 * nbatch * tmp_elems (E P and its transpose)
  double *wbuf = nullptr;    // SPD-inverse scratch: nbatch * wstride
  std::vector<long> woff;    // per recursion depth, offset in a design's scratch
  long wstride = 0;
  int *info = nullptr;       // nbatch
  int *d_bcflag = nullptr;   // per level-ordered dof: 1 if a BC dof
  int *d_pos2dof = nullptr;  // level-ordered position -> original dof
  long *d_blk_base = nullptr;  // per BSR block: dense offset of entry (0,0)
  int *d_blk_ld = nullptr;     // per BSR block: leading dimension of its target
  double *vb = nullptr, *vz = nullptr, *vx = nullptr;  // level-ordered work vectors
  double *part = nullptr;    // GEMV partial sums
  int rhs_cap = 1;           // vectors per design that vb, vz, vx and part hold
  hipGraphExec_t solve_graph = nullptr;  // the single-vector solve body, captured
  bool graph_off = false;
  // single-vector solves of the sequential form: only designs d_act[0 ..
  // nact) when nact >= 0 (Lanczos stops solving for converged designs)
  int *d_act = nullptr;
  int nact = -1;
  // several vectors through a cyclic-reduction factor of one design: work
  // vectors of crm_cap vectors each (level-pair stride), the symmetric
  // products' partial sums, and the (vector, odd level) entries they
  // reduce
  int crm_cap = 0;
  double *crm_b = nullptr, *crm_z = nullptr, *crm_x = nullptr, *crm_part = nullptr;
  int *d_crm_act = nullptr;
  bool factored = false;
  long nv = 0;               // level-ordered vector length per design
  // block cyclic reduction over uniform (padded) levels of M dofs: per
  // design, blocks D_i (-> C_i^{-1}), B_i (coupling to the left active
  // neighbour -> Y_i = C_i^{-1} B_i), Z_i (-> W_i = B_{i+s} C_i^{-T})
  bool cr = false;
  int M = 0;
  int nkmax = 1;             // largest number of levels eliminated in one round
  int *d_mlev = nullptr;     // dofs per level (padding below)
  double **d_ptr = nullptr;  // device pointer arrays of all rounds
  int *cr_info = nullptr;    // potrf / trtri status per (round entry)
  std::vector<FactorRound> rounds;
  long root_off = 0;         // pointer group of the root level D_0
  // recursive Cholesky-inverse of the M x M diagonal blocks: the node
  // sequence (shared by all rounds) and, per round (and the root, last),
  // the offset of each node's pointer arrays in d_ptr
  std::vector<CholNode> chol;
  // pointer-array groups of the Cholesky-inverse: cyclic reduction, one per
  // round and the root (last); sequential factorisation, one per level
  struct CholGroup {
    int n, lda, cnt;
    std::vector<CholNode> nodes;
    std::vector<long> offs;  // per node: first pointer (A11, A21, A22 arrays of cnt)
  };
  std::vector<CholGroup> cgroups;
  // first round of cyclic reduction on the sparse couplings: the level
  // node offsets on the device and the job lists of its batched products
  // (offsets into d_jobs, entries q = k nb + d of the round)
  int *d_noff = nullptr;
  CoupleJob *d_jobs = nullptr;
  long job_q = 0, job_t1 = 0, job_dl = 0, job_v = 0, job_t2 = 0, job_dr = 0, job_br = 0;
};

namespace {

int ferr(shellgpu_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

#define HIPCALL(ctx, call)                                                      \
  do {                                                                          \
    hipError_t _e = (call);                                                     \
    if (_e != hipSuccess) return ferr(ctx, 2, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define BLASCALL(ctx, call)                                                      \
  do {                                                                           \
    rocblas_status _s = (call);                                                  \
    if (_s != rocblas_status_success)                                            \
      return ferr(ctx, 4, std::string(#call) + ": " + rocblas_status_to_string(_s)); \
  } while (0)

// Dense updates of the sequential factorisation: rocBLAS, or with
// SHELLGPU_GEMM=own the project's FP64 matrix-core GEMM (gemm_f64.hip,
// which runs every dense update of cyclic reduction).
bool own_gemm() {
  static const bool own = [] {
    const char *e = getenv("SHELLGPU_GEMM");
    return e && strcmp(e, "own") == 0;
  }();
  return own;
}

// strided batch: C_b = alpha op(A_b) op(B_b) + beta C_b
int gemm_sb(shellgpu_ctx *ctx, bool ta, bool tb, int M, int N, int K, double alpha, const double *A, int lda, long sA,
            const double *B, int ldb, long sB, double beta, double *C, int ldc, long sC, int batch) {
  if (own_gemm()) {
    GemmArgs g{M, N, K, alpha, beta, A, sA, nullptr, lda, B, sB, nullptr, ldb, C, sC, nullptr, ldc};
    HIPCALL(ctx, launch_gemm_f64(ta, tb, g, batch, ctx->stream));
    return 0;
  }
  BLASCALL(ctx, rocblas_dgemm_strided_batched((rocblas_handle)ctx->blas,
                                              ta ? rocblas_operation_transpose : rocblas_operation_none,
                                              tb ? rocblas_operation_transpose : rocblas_operation_none, M, N, K,
                                              &alpha, A, lda, sA, B, ldb, sB, &beta, C, ldc, sC, batch));
  return 0;
}

// pointer-array batch on the matrix cores, with triangular operands and
// lower-only results (gemm_f64.hip): the dense updates of cyclic reduction
int mgemm(shellgpu_ctx *ctx, bool ta, bool tb, int M, int N, int K, double alpha, double *const *A, int lda,
          int tri_a, double *const *B, int ldb, int tri_b, double beta, double *const *C, int ldc, int lower_c,
          int batch) {
  if (batch <= 0 || M <= 0 || N <= 0) return 0;
  GemmArgs g{M, N, K, alpha, beta, nullptr, 0, (const double *const *)A, lda, nullptr, 0,
             (const double *const *)B, ldb, nullptr, 0, C, ldc};
  g.tri_a = tri_a;
  g.tri_b = tri_b;
  g.lower_c = lower_c;
  HIPCALL(ctx, launch_gemm_f64(ta, tb, g, batch, ctx->stream));
  return 0;
}

// BSR values -> dense level blocks (D_l and L_l; blocks above the diagonal
// level are not stored).  BC columns are zeroed (the BSR rows of BC dofs are
// already zero with a unit diagonal), giving the symmetric elimination.
__global__ void scatter_dense_kernel(int nnzb, const double *__restrict__ vals, long vstride,
                                     const long *__restrict__ base, const int *__restrict__ ld,
                                     const int *__restrict__ colbc, double *__restrict__ dst, long dstride) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 36L * nnzb) return;
  const int p = (int)(i / 36), e = (int)(i % 36), r = e / 6, c = e % 6;
  if (base[p] < 0) return;
  vals += (long)blockIdx.y * vstride;
  dst += (long)blockIdx.y * dstride;
  const int bcmask = colbc[p];  // bits r (row dof BC) and 6 + c (column dof BC), bit 12 diagonal block
  double v = vals[i];
  if ((bcmask >> (6 + c)) & 1) v = ((bcmask >> 12) & 1) && r == c ? v : 0.0;
  dst[base[p] + (long)c * ld[p] + r] = v;
}

// 16-byte load of factor data, marked non-temporal: a solve streams every
// stored inverse once, with no reuse to keep in the caches (9 % faster
// level solves at 64 designs)
__device__ inline double2 load_stream2(const double *p) {
  typedef double dvec2 __attribute__((ext_vector_type(2)));
  const dvec2 t = __builtin_nontemporal_load(reinterpret_cast<const dvec2 *>(p));
  return make_double2(t.x, t.y);
}

// Batch of GEMVs: entry q = k * nd + d (k: level entry, d: design) uses
// A + d sAd + k sAk, x + d sxd + k sxk, y + d syd + k syk.
struct GemvBatch {
  int nd;
  long sAd, sAk, sxd, sxk, syd, syk;
};

// y = beta y + alpha A x, A column-major m x n (lda, even; A and x 16-byte
// aligned); LOWER restricts to the lower triangle (col <= row).
// Pass 1: a block of 4 wavefronts covers 128 rows (two per lane, one
// 16-byte load per column) of one column chunk (gridDim.y chunks), the
// wavefronts taking every fourth column, 4 columns in flight per lane;
// partial sums go to part[entry][chunk][row].
template <int LOWER>
__global__ __launch_bounds__(256) void gemv_n_pass1(int m, int n, const double *__restrict__ A, int lda,
                                                    const double *__restrict__ x, GemvBatch g, int chunk,
                                                    double *__restrict__ part) {
  __shared__ double2 red[256];
  const int q = blockIdx.z, d = q % g.nd, k = q / g.nd;
  A += d * g.sAd + k * g.sAk;
  x += d * g.sxd + k * g.sxk;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 128 + 2 * lane;
  const int c0 = blockIdx.y * chunk;
  int c1 = min(n, c0 + chunk);
  double a0 = 0.0, a1 = 0.0;
  if (row < m) {
    if (LOWER) c1 = min(c1, row + 2);
    int c = c0 + wave;
    // 8 columns' loads in flight per lane (about 10 % more bandwidth than 4)
    constexpr int kU = 8;
    for (; c + 4 * (kU - 1) < c1; c += 4 * kU) {
      double2 v[kU];
      double xc[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) v[u] = load_stream2(A + row + (long)(c + 4 * u) * lda);
#pragma unroll
      for (int u = 0; u < kU; u++) xc[u] = x[c + 4 * u];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        a0 += ((!LOWER || c + 4 * u <= row) ? v[u].x : 0.0) * xc[u];
        a1 += v[u].y * xc[u];
      }
    }
    for (; c < c1; c += 4) {
      const double2 v = load_stream2(A + row + (long)c * lda);
      const double xc = x[c];
      a0 += ((!LOWER || c <= row) ? v.x : 0.0) * xc;
      a1 += v.y * xc;
    }
  }
  red[threadIdx.x] = make_double2(a0, a1);
  __syncthreads();
  if (wave == 0 && row < m) {
    const double2 r0 = red[lane], r1 = red[lane + 64], r2 = red[lane + 128], r3 = red[lane + 192];
    double2 out;
    out.x = (r0.x + r1.x) + (r2.x + r3.x);
    out.y = (r0.y + r1.y) + (r2.y + r3.y);
    *reinterpret_cast<double2 *>(part + ((long)q * gridDim.y + blockIdx.y) * m + row) = out;
  }
}

// Pass 2: y = beta y + alpha sum over chunks, 8 partials in flight per thread.
__global__ void gemv_n_pass2(int m, int nsplit, const double *__restrict__ part, double alpha, double beta,
                             double *__restrict__ y, GemvBatch g) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= m) return;
  const int q = blockIdx.y, d = q % g.nd, k = q / g.nd;
  const double *p = part + (long)q * nsplit * m + row;
  double s = 0.0;
  int c = 0;
  for (; c + 8 <= nsplit; c += 8) {
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; u++) t[u] = p[(long)(c + u) * m];
    s += ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
  }
  for (; c < nsplit; c++) s += p[(long)c * m];
  double *yy = y + d * g.syd + k * g.syk;
  yy[row] = (beta == 0.0 ? 0.0 : beta * yy[row]) + alpha * s;
}

// y = beta y + alpha A^T x, A column-major m x n (lda, even; 16-byte
// aligned): one wavefront per output entry j, reading column j with one
// 16-byte load per lane and 4 loads in flight; LOWER restricts to rows
// i >= j.
template <int LOWER>
__global__ __launch_bounds__(256) void gemv_t_kernel(int m, int n, const double *__restrict__ A, int lda,
                                                     const double *__restrict__ x, GemvBatch g, double alpha,
                                                     double beta, double *__restrict__ y) {
  const int q = blockIdx.y, d = q % g.nd, k = q / g.nd;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n) return;
  const double *col = A + d * g.sAd + k * g.sAk + (long)j * lda;
  const double *xx = x + d * g.sxd + k * g.sxk;
  const int i0 = LOWER ? (j & ~1) : 0;  // even start keeps the 16-byte alignment
  double a0 = 0.0, a1 = 0.0;
  int i = i0 + 2 * lane;
  // 8 loads in flight per lane (about 8 % more bandwidth than 4)
  constexpr int kU = 8;
  for (; i + 128 * (kU - 1) < m; i += 128 * kU) {
    double2 v[kU], xv[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) v[u] = load_stream2(col + i + 128 * u);
#pragma unroll
    for (int u = 0; u < kU; u++) xv[u] = *reinterpret_cast<const double2 *>(xx + i + 128 * u);
    // in a LOWER column only the first pair can hold a row above the
    // diagonal (row j - 1 when j is odd)
#pragma unroll
    for (int u = 0; u < kU; u++) {
      a0 += (LOWER && u == 0 && i < j) ? 0.0 : v[u].x * xv[u].x;
      a1 += v[u].y * xv[u].y;
    }
  }
  for (; i < m; i += 128) {
    const double2 v = load_stream2(col + i);
    const double2 xv = *reinterpret_cast<const double2 *>(xx + i);
    a0 += (LOWER && i < j) ? 0.0 : v.x * xv.x;
    a1 += v.y * xv.y;
  }
  double acc = a0 + a1;
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) {
    double *yy = y + d * g.syd + k * g.syk;
    yy[j] = (beta == 0.0 ? 0.0 : beta * yy[j]) + alpha * acc;
  }
}

// Cholesky factor and its inverse of one n x n SPD block (n <= 96) per
// workgroup, in LDS.  Right-looking with the scaling of the pivot column
// deferred: step j subtracts L_ij L_kj / a_j (a_j the pivot) from the
// trailing triangle and, for the inverse, Y_ic -= (L_ij / a_j) Y_jc (c <= j,
// Y = I at the start), so one barrier per step suffices; at the end
// X = chol^{-1} has rows X_ic = Y_ic / sqrt(a_i).  A wavefront takes every
// sixteenth row, its lanes the columns; within a step a wavefront reads all
// of its rows before it writes any (the step reads only column j and row j
// besides the wavefront's own rows), so the LDS latency is paid once per
// step instead of once per row.
constexpr int kCholBase = 96;
constexpr int kCholThreads = 1024;
constexpr int kCholLd = kCholBase + 1;  // odd row stride: column walks spread over the banks
constexpr int kCholWaves = kCholThreads / 64;
constexpr int kCholRows = (kCholBase + kCholWaves - 1) / kCholWaves;  // rows per wavefront and step

// A (column-major, lda) -> L (scratch) and Y (rows of chol^{-1} before
// scaling), piv (pivots); returns true for a non-positive pivot
__device__ bool chol_inverse_lds(const double *A, int n, int lda, double *L, double *Y, double *piv) {
  constexpr int ld = kCholLd;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  for (int e = t; e < n * n; e += kCholThreads) {
    const int i = e % n, j = e / n;
    L[i * ld + j] = A[i + (long)j * lda];
    Y[i * ld + j] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  bool bad = false;
  for (int j = 0; j < n; j++) {
    double a = L[j * ld + j];
    if (!(a > 0.0)) {
      bad = true;
      a = 1.0;
    }
    if (t == 0) piv[j] = a;
    const double ra = 1.0 / a;
    const int k0 = j + 1 + lane, k1 = k0 + 64;  // trailing columns of this lane
    const double l0 = k0 < n ? L[k0 * ld + j] : 0.0, l1 = k1 < n ? L[k1 * ld + j] : 0.0;
    const bool c0 = lane <= j, c1 = lane + 64 <= j;
    const double y0 = c0 ? Y[j * ld + lane] : 0.0, y1 = c1 ? Y[j * ld + lane + 64] : 0.0;
    double r[kCholRows], v0[kCholRows], v1[kCholRows], w0[kCholRows], w1[kCholRows];
#pragma unroll
    for (int s = 0; s < kCholRows; s++) {
      const int i = j + 1 + wave + kCholWaves * s;
      r[s] = v0[s] = v1[s] = w0[s] = w1[s] = 0.0;
      if (i < n) {
        r[s] = L[i * ld + j] * ra;
        if (k0 <= i) v0[s] = L[i * ld + k0];
        if (k1 <= i) v1[s] = L[i * ld + k1];
        if (c0) w0[s] = Y[i * ld + lane];
        if (c1) w1[s] = Y[i * ld + lane + 64];
      }
    }
#pragma unroll
    for (int s = 0; s < kCholRows; s++) {
      const int i = j + 1 + wave + kCholWaves * s;
      if (i < n) {
        if (k0 <= i) L[i * ld + k0] = v0[s] - r[s] * l0;
        if (k1 <= i) L[i * ld + k1] = v1[s] - r[s] * l1;
        if (c0) Y[i * ld + lane] = w0[s] - r[s] * y0;
        if (c1) Y[i * ld + lane + 64] = w1[s] - r[s] * y1;
      }
    }
    __syncthreads();
  }
  return bad;
}

// X = chol(A)^{-1}, lower, with zeros above the diagonal of the block, in
// place (A: column-major, lda; blocks ptrs[blockIdx.x]); info[entry] = 1
// for a non-positive pivot
__global__ __launch_bounds__(kCholThreads) void cholinv_base_kernel(double *const *ptrs, int n, int lda, int *info) {
  __shared__ double L[kCholBase * kCholLd];
  __shared__ double Y[kCholBase * kCholLd];
  __shared__ double piv[kCholBase];
  double *A = ptrs[blockIdx.x];
  const bool bad = chol_inverse_lds(A, n, lda, L, Y, piv);
  for (int e = threadIdx.x; e < n * n; e += kCholThreads) {
    const int i = e % n, c = e / n;
    A[i + (long)c * lda] = i >= c ? Y[i * kCholLd + c] / sqrt(piv[i]) : 0.0;
  }
  if (threadIdx.x == 0 && bad) info[blockIdx.x] = 1;
}

// Inverse of one n x n SPD block (n <= 96) per workgroup (design
// blockIdx.x, blocks A + blockIdx.x sA), in place and in full (both
// triangles): X = chol(A)^{-1} as above, then A^{-1} = X^T X,
// P_ij = sum_{k >= i} X_ki X_kj.  info[blockIdx.x] = 1 for a non-positive
// pivot.
__global__ __launch_bounds__(kCholThreads) void spdinv_base_kernel(double *A, long sA, int n, int lda, int *info) {
  constexpr int ld = kCholLd;
  __shared__ double L[kCholBase * ld];
  __shared__ double Y[kCholBase * ld];
  __shared__ double piv[kCholBase];
  A += blockIdx.x * sA;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const bool bad = chol_inverse_lds(A, n, lda, L, Y, piv);
  // X = Y scaled by rows (lower), into Y
  for (int e = t; e < n * n; e += kCholThreads) {
    const int i = e / n, c = e % n;
    Y[i * ld + c] = c <= i ? Y[i * ld + c] / sqrt(piv[i]) : 0.0;
  }
  __syncthreads();
  // P = X^T X into L (lower; four partial sums per entry), then written in full
  for (int i = wave; i < n; i += kCholWaves)
    for (int j = lane; j <= i; j += 64) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      int k = i;
      for (; k + 3 < n; k += 4) {
        s0 += Y[k * ld + i] * Y[k * ld + j];
        s1 += Y[(k + 1) * ld + i] * Y[(k + 1) * ld + j];
        s2 += Y[(k + 2) * ld + i] * Y[(k + 2) * ld + j];
        s3 += Y[(k + 3) * ld + i] * Y[(k + 3) * ld + j];
      }
      for (; k < n; k++) s0 += Y[k * ld + i] * Y[k * ld + j];
      L[i * ld + j] = (s0 + s1) + (s2 + s3);
    }
  __syncthreads();
  for (int e = t; e < n * n; e += kCholThreads) {
    const int i = e % n, j = e / n;
    A[i + (long)j * lda] = i >= j ? L[i * ld + j] : L[j * ld + i];
  }
  if (t == 0 && bad) info[blockIdx.x] = 1;
}

// B = A^T (A: rows x cols, column-major lda; B: cols x rows, ldb), 32 x 32
// tiles through LDS, batched over designs (blockIdx.z)
__global__ __launch_bounds__(256) void transpose_kernel(int rows, int cols, const double *__restrict__ A, int lda,
                                                        long sAd, double *__restrict__ B, int ldb, long sBd) {
  __shared__ double tile[32][33];
  A += blockIdx.z * sAd;
  B += blockIdx.z * sBd;
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8)
    if (r0 + tx < rows && c0 + k < cols) tile[k][tx] = A[(r0 + tx) + (long)(c0 + k) * lda];
  __syncthreads();
  for (int k = ty; k < 32; k += 8)
    if (c0 + tx < cols && r0 + k < rows) B[(c0 + tx) + (long)(r0 + k) * ldb] = tile[tx][k];
}

// lower-triangle tile index t -> (I, J), I >= J
__device__ inline void tri_tile(int t, int &I, int &J) {
  I = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
  while ((I + 1) * (I + 2) / 2 <= t) I++;
  while (I * (I + 1) / 2 > t) I--;
  J = t - I * (I + 1) / 2;
}

// Y[:, 6j + a] = beta Y[:, 6j + a] + alpha sum_q sum_r X[:, 6 col_q + r] E_q(a, r)
// over the coupling blocks q of node j of one level, rows 0..rows-1 of the
// column-major X and Y; E_q(a, r) = E[36 q + 6 r + a].  A workgroup takes
// 256 rows of kCoupleNodes consecutive nodes (blockIdx.y), which share most
// of their neighbour columns of X; per design (blockIdx.z) strides of X, Y
// and E.
constexpr int kCoupleNodes = 4;
__global__ __launch_bounds__(256) void couple_right_kernel(int rows, int nnodes, const double *__restrict__ X,
                                                           int ldx, long sXd, double *__restrict__ Y, int ldy,
                                                           long sYd, const int *__restrict__ rowp,
                                                           const int *__restrict__ colx, const double *__restrict__ E,
                                                           long sEd, double alpha, double beta) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= rows) return;
  const int d = blockIdx.z;
  X += d * sXd + i;
  Y += d * sYd + i;
  E += d * sEd;
  const int j1 = min(nnodes, (int)(blockIdx.y + 1) * kCoupleNodes);
  for (int j = blockIdx.y * kCoupleNodes; j < j1; j++) {
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q = rowp[j]; q < rowp[j + 1]; q++) {
      const double *xk = X + 6L * colx[q] * ldx;
      const double *e = E + 36L * q;
      double xv[6];
#pragma unroll
      for (int r = 0; r < 6; r++) xv[r] = xk[(long)r * ldx];
#pragma unroll
      for (int r = 0; r < 6; r++)
#pragma unroll
        for (int a = 0; a < 6; a++) acc[a] += xv[r] * e[6 * r + a];
    }
    double *y = Y + 6L * j * ldy;
#pragma unroll
    for (int a = 0; a < 6; a++) y[(long)a * ldy] = (beta == 0.0 ? 0.0 : beta * y[(long)a * ldy]) + alpha * acc[a];
  }
}

// Y[:, 6j + a] over the nodes j of level ylev of job blockIdx.z, a
// workgroup taking 256 rows of kCoupleNodes consecutive nodes (blockIdx.y);
// X and Y have leading dimension ld
template <bool LEFT>
__global__ __launch_bounds__(256) void couple_job_kernel(const CoupleJob *__restrict__ jobs, int ld,
                                                         const int *__restrict__ noff, const int *__restrict__ rowp,
                                                         const int *__restrict__ colx, const int *__restrict__ blkx,
                                                         double alpha, double beta) {
  const CoupleJob J = jobs[blockIdx.z];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= J.rows) return;
  const int u0 = noff[J.ylev], nn = noff[J.ylev + 1] - u0;
  const double *X = J.X + i;
  double *Y = J.Y + i;
  const int j1 = min(nn, (int)(blockIdx.y + 1) * kCoupleNodes);
  for (int j = blockIdx.y * kCoupleNodes; j < j1; j++) {
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q = rowp[u0 + j]; q < rowp[u0 + j + 1]; q++) {
      const double *xk = X + 6L * colx[q] * ld;
      const double *e = J.E + 36L * (LEFT ? blkx[q] : q);
      double xv[6];
#pragma unroll
      for (int r = 0; r < 6; r++) xv[r] = xk[(long)r * ld];
      // RIGHT: block (row a, column r) = e[6 r + a]; LEFT: the transpose
#pragma unroll
      for (int r = 0; r < 6; r++)
#pragma unroll
        for (int a = 0; a < 6; a++) acc[a] += xv[r] * (LEFT ? e[6 * a + r] : e[6 * r + a]);
    }
    double *y = Y + 6L * j * ld;
#pragma unroll
    for (int a = 0; a < 6; a++) y[(long)a * ld] = (beta == 0.0 ? 0.0 : beta * y[(long)a * ld]) + alpha * acc[a];
  }
}

// Y = X^T (X: rows x cols, Y: cols x rows, leading dimension ld) for job
// blockIdx.z, 32 x 32 tiles through LDS
__global__ __launch_bounds__(256) void transpose_job_kernel(const CoupleJob *__restrict__ jobs, int ld) {
  __shared__ double tile[32][33];
  const CoupleJob J = jobs[blockIdx.z];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (r0 >= J.rows || c0 >= J.cols) return;
  for (int k = ty; k < 32; k += 8)
    if (r0 + tx < J.rows && c0 + k < J.cols) tile[k][tx] = J.X[(r0 + tx) + (long)(c0 + k) * ld];
  __syncthreads();
  for (int k = ty; k < 32; k += 8)
    if (c0 + tx < J.cols && r0 + k < J.rows) J.Y[(c0 + tx) + (long)(r0 + k) * ld] = tile[tx][k];
}

// dst_k = beta dst_k + alpha (E_k src_{k-1} + E_{k+1}^T src_{k+1}) over the
// node dofs of levels k = k0 + 2 blockIdx.y (level k at offset k M of a
// vector; designs blockIdx.z with vector stride svd and coupling stride sEd)
__global__ __launch_bounds__(256) void cr_gather_kernel(int k0, int nl, int M, const int *__restrict__ noff,
                                                        const int *__restrict__ cp_rowp,
                                                        const int *__restrict__ cp_col,
                                                        const int *__restrict__ ct_rowp,
                                                        const int *__restrict__ ct_src,
                                                        const int *__restrict__ ct_blk, const double *__restrict__ E,
                                                        long sEd, const double *__restrict__ src,
                                                        double *__restrict__ dst, long svd, double alpha,
                                                        double beta) {
  const int k = k0 + 2 * blockIdx.y, d = blockIdx.z;
  const int u0 = noff[k], nn = noff[k + 1] - u0;
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= 6 * nn) return;
  const int j = g / 6, a = g - 6 * j, u = u0 + j;
  E += d * sEd;
  src += d * svd;
  double s = 0.0;
  if (k > 0) {
    const double *t = src + (long)(k - 1) * M;
    for (int q = cp_rowp[u]; q < cp_rowp[u + 1]; q++) {
      const double *tk = t + 6 * cp_col[q];
      const double *ep = E + 36L * q + a;
#pragma unroll
      for (int r = 0; r < 6; r++) s += ep[6 * r] * tk[r];
    }
  }
  if (k + 1 < nl) {
    const double *t = src + (long)(k + 1) * M;
    for (int e = ct_rowp[u]; e < ct_rowp[u + 1]; e++) {
      const double *tk = t + 6 * ct_src[e];
      const double *ep = E + 36L * ct_blk[e] + 6 * a;
#pragma unroll
      for (int r = 0; r < 6; r++) s += ep[r] * tk[r];
    }
  }
  double *y = dst + d * svd + (long)k * M + g;
  *y = (beta == 0.0 ? 0.0 : beta * *y) + alpha * s;
}

// z[6i + a] -= sum_q sum_r E_q(a, r) t[6 col_q + r] for the nodes i of one
// level (t: the previous level's segment); vector entry e = blockIdx.y (of
// stride svd) uses the couplings of design e % nbf
__global__ void couple_fwd_kernel(int nnodes, const int *__restrict__ rowp, const int *__restrict__ colx,
                                  const double *__restrict__ E, long sEd, const double *__restrict__ t,
                                  double *__restrict__ z, long svd, int nbf, const int *__restrict__ act) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= 6 * nnodes) return;
  const int i = g / 6, a = g - 6 * i, e = act ? act[blockIdx.y] : blockIdx.y;
  E += (e % nbf) * sEd + a;
  t += e * svd;
  double s = 0.0;
  for (int q = rowp[i]; q < rowp[i + 1]; q++) {
    const double *tk = t + 6 * colx[q];
    const double *ep = E + 36L * q;
#pragma unroll
    for (int r = 0; r < 6; r++) s += ep[6 * r] * tk[r];
  }
  z[e * svd + g] -= s;
}

// y[6v + r] = sum_e sum_a E_{blk_e}(a, r) x[6 src_e + a] for the nodes v of
// one level (x: the next level's segment); vector entries as above
__global__ void couple_bwd_kernel(int nnodes, const int *__restrict__ rowp, const int *__restrict__ src,
                                  const int *__restrict__ blk, const double *__restrict__ E, long sEd,
                                  const double *__restrict__ x, double *__restrict__ y, long svd, int nbf,
                                  const int *__restrict__ act) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= 6 * nnodes) return;
  const int v = g / 6, r = g - 6 * v, e = act ? act[blockIdx.y] : blockIdx.y;
  E += (e % nbf) * sEd + 6 * r;
  x += e * svd;
  double s = 0.0;
  for (int k = rowp[v]; k < rowp[v + 1]; k++) {
    const double *xu = x + 6 * src[k];
    const double *ep = E + 36L * blk[k];
#pragma unroll
    for (int a = 0; a < 6; a++) s += ep[a] * xu[a];
  }
  y[e * svd + g] = s;
}

// y = P x for a symmetric m x m P stored in full (column-major, lda; m
// even, 16-byte aligned), reading the lower triangle only, in 128 x 128
// tiles (I, J), I >= J (blockIdx.x), designs blockIdx.y.  Pass 1: tile
// (I, J) writes the partial sums P_IJ x_J of block row I into slot J and,
// for I > J, P_IJ^T x_I of block row J into slot I: every (block row, slot)
// pair is written once.  Lanes take two rows each (one 16-byte load per
// column), each wavefront every fourth column; the column sums of the
// transposed product are reduced over the lanes by a halving exchange (32
// values in 32 shuffles).
constexpr int kSymvTile = 128;
// Entry q of a batch of symmetric products (blockIdx.y, or act[blockIdx.y]):
// level entry k = q / nd, design q % nd; matrix, vector and result offsets
// (q % nd) s*d + k s*k (the sequential form: nd = entries, one stride).
struct SymvBatch {
  int nd;
  long sAd, sAk, sxd, sxk, syd, syk;
  __device__ long a(int q) const { return (q % nd) * sAd + (q / nd) * sAk; }
  __device__ long x(int q) const { return (q % nd) * sxd + (q / nd) * sxk; }
  __device__ long y(int q) const { return (q % nd) * syd + (q / nd) * syk; }
};
__global__ __launch_bounds__(256) void symv_pass1_kernel(int m, int nT, const double *__restrict__ A, int lda,
                                                         const double *__restrict__ x, SymvBatch sb,
                                                         double *__restrict__ part, const int *__restrict__ act) {
  __shared__ double2 red[4][64];
  int I, J;
  tri_tile(blockIdx.x, I, J);
  const int d = act ? act[blockIdx.y] : blockIdx.y;
  A += sb.a(d);
  x += sb.x(d);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int row = I * kSymvTile + 2 * lane;
  const bool rv = row < m;  // m is even: row + 1 < m as well
  const double2 xi = rv ? *reinterpret_cast<const double2 *>(x + row) : make_double2(0.0, 0.0);
  double a0 = 0.0, a1 = 0.0, tp[32];
#pragma unroll
  for (int p = 0; p < 32; p++) {
    const int col = J * kSymvTile + wave + 4 * p;
    const double2 v = (rv && col < m) ? load_stream2(A + row + (long)col * lda) : make_double2(0.0, 0.0);
    const double xc = col < m ? x[col] : 0.0;
    a0 += v.x * xc;
    a1 += v.y * xc;
    tp[p] = v.x * xi.x + v.y * xi.y;
  }
  red[wave][lane] = make_double2(a0, a1);
  if (I > J) {
#pragma unroll
    for (int h = 16, b = 32; h >= 1; h >>= 1, b >>= 1) {
      const bool hi = (lane & b) != 0;
#pragma unroll
      for (int k = 0; k < h; k++) {
        const double send = hi ? tp[k] : tp[k + h];
        const double keep = hi ? tp[k + h] : tp[k];
        tp[k] = keep + __shfl_xor(send, b, 64);
      }
    }
    tp[0] += __shfl_xor(tp[0], 1, 64);
    if ((lane & 1) == 0)
      part[(((long)d * nT + J) * nT + I) * kSymvTile + wave + 4 * ((lane >> 1) & 31)] = tp[0];
  }
  __syncthreads();
  if (wave == 0) {
    const double2 r0 = red[0][lane], r1 = red[1][lane], r2 = red[2][lane], r3 = red[3][lane];
    *reinterpret_cast<double2 *>(part + (((long)d * nT + I) * nT + J) * kSymvTile + 2 * lane) =
        make_double2((r0.x + r1.x) + (r2.x + r3.x), (r0.y + r1.y) + (r2.y + r3.y));
  }
}

// Y_r = P X_r for NR vectors per design (vector entries e = r nbf + d of
// stride sxe), P read in full once for all of them: 128 x 128 tiles (I, J)
// over the whole matrix, each split into CS column slabs of 128 / CS
// columns (blockIdx.x = (I nT + J) CS + slab), designs blockIdx.y; the
// partial sums of block row I go to slot J CS + slab of nT CS slots.
// Splitting keeps the launch wider than the GPU when one design is
// factored.  The slab's NR x (128 / CS) vector entries are staged in LDS
// first (read there as wave-uniform broadcasts); columns past m read the
// last column against zero entries, so every wave issues all its loads
// without a bound test.
constexpr int kMaxSplit = 2;
template <int NR, int CS>
__global__ __launch_bounds__(256) void symv_multi_pass1_kernel(int m, int nT, const double *__restrict__ A, int lda,
                                                               long sAd, const double *__restrict__ x, long sxe,
                                                               int nbf, double *__restrict__ part) {
  // the 4 waves' sums are added through LDS, 8 vectors at a time, in the
  // space that held the staged entries
  constexpr int RH = NR < 8 ? NR : 8, NC = kSymvTile / CS;
  constexpr int kRed = 4 * RH * 64 * 2, kX = NC * NR;
  __shared__ double lds[kRed > kX ? kRed : kX];
  double2(*red)[RH][64] = reinterpret_cast<double2(*)[RH][64]>(lds);
  const int tile = blockIdx.x / CS, slab = blockIdx.x - tile * CS;
  const int I = tile / nT, J = tile - I * nT, d = blockIdx.y, ns = nT * CS;
  A += d * sAd;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int row = I * kSymvTile + 2 * lane;
  const bool rv = row < m;
  const int cbase = J * kSymvTile + slab * NC;
  for (int q = threadIdx.x; q < kX; q += 256) {
    const int c = q % NC, r = q / NC;
    lds[c * NR + r] = cbase + c < m ? x[((long)r * nbf + d) * sxe + cbase + c] : 0.0;
  }
  __syncthreads();
  double2 acc[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) acc[r] = make_double2(0.0, 0.0);
  // a wave's columns w, w + 4, ...: CH loads issued before their products
  constexpr int CPW = NC / 4, CH = CPW < 8 ? CPW : 8;
#pragma unroll 1
  for (int p0 = 0; p0 < CPW; p0 += CH) {
    double2 v[CH];
#pragma unroll
    for (int q = 0; q < CH; q++) {
      const int col = min(cbase + wave + 4 * (p0 + q), m - 1);
      v[q] = rv ? load_stream2(A + row + (long)col * lda) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < CH; q++) {
      const int c = wave + 4 * (p0 + q);
#pragma unroll
      for (int r = 0; r < NR; r++) {
        const double xc = lds[c * NR + r];
        acc[r].x += v[q].x * xc;
        acc[r].y += v[q].y * xc;
      }
      // keep each column's LDS reads in its own iteration (hoisted, they
      // would hold every column's entries in registers at once)
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int h = 0; h < NR; h += RH) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RH; r++) red[wave][r][lane] = acc[h + r];
    __syncthreads();
    for (int q = threadIdx.x; q < 64 * RH; q += 256) {
      const int ln = q & 63, r = q >> 6;
      const double2 a0 = red[0][r][ln], a1 = red[1][r][ln], a2 = red[2][r][ln], a3 = red[3][r][ln];
      const long e = (long)(h + r) * nbf + d;
      *reinterpret_cast<double2 *>(part + ((e * nT + I) * ns + J * CS + slab) * kSymvTile + 2 * ln) =
          make_double2((a0.x + a1.x) + (a2.x + a3.x), (a0.y + a1.y) + (a2.y + a3.y));
    }
  }
}

// Pass 2: y = beta c + alpha (sum over the ns slots of the row's tile).
// 64 rows per workgroup; the slots are dealt round-robin to G waves, whose
// sums are added in LDS in a fixed order, so the result does not depend on
// timing.  Spreading the slots keeps a launch of few rows (one factored
// design, many slots) from being a chain of dependent loads per thread.
template <int G>
__global__ __launch_bounds__(64 * G) void symv_pass2_kernel(int m, int nT, int ns, const double *__restrict__ part,
                                                            double alpha, const double *__restrict__ c, double beta,
                                                            double *__restrict__ y, SymvBatch sb,
                                                            const int *__restrict__ act) {
  __shared__ double red[G][64];
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int row = blockIdx.x * 64 + lane, d = act ? act[blockIdx.y] : blockIdx.y;
  double s = 0.0;
  if (row < m) {
    const double *p = part + ((long)d * nT + row / kSymvTile) * ns * kSymvTile + row % kSymvTile;
    for (int J = g; J < ns; J += G) s += p[(long)J * kSymvTile];
  }
  if (G > 1) {
    red[g][lane] = s;
    __syncthreads();
    if (g != 0) return;
    for (int k = 1; k < G; k++) s += red[k][lane];
  }
  if (row < m) y[sb.y(d) + row] = (beta == 0.0 ? 0.0 : beta * c[sb.y(d) + row]) + alpha * s;
}

// identity on the padded diagonal entries of the cyclic-reduction levels
__global__ void pad_identity_kernel(int M, long M2, const int *__restrict__ mlev, double *__restrict__ buf,
                                    long per_design) {
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  const int l = blockIdx.y;
  if (kk >= M || kk < mlev[l]) return;
  buf[(long)blockIdx.z * per_design + (long)l * M2 + (long)kk * M + kk] = 1.0;
}

// level order <-> dof order; pos2dof[pos] < 0 marks a padding position
// (zero in level order); vector blockIdx.y, or act[blockIdx.y]
__global__ void permute_kernel(long nv, long n, const int *__restrict__ pos2dof, const double *__restrict__ src,
                               double *__restrict__ dst, int to_level, const int *__restrict__ act) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int p = pos2dof[i];
  const long v = act ? act[blockIdx.y] : blockIdx.y;
  const long lo = v * nv, doff = v * n;
  if (to_level) {
    dst[lo + i] = p >= 0 ? src[doff + p] : 0.0;
  } else if (p >= 0) {
    dst[doff + p] = src[lo + i];
  }
}

// level order <-> dof order for nrhs vectors (blockIdx.y) of one design,
// level-ordered vectors at a stride of their own
__global__ void permute_rhs_kernel(long nv, long n, long stride, const int *__restrict__ pos2dof,
                                   const double *__restrict__ src, double *__restrict__ dst, int to_level) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int p = pos2dof[i];
  const long v = blockIdx.y;
  if (to_level)
    dst[v * stride + i] = p >= 0 ? src[v * n + p] : 0.0;
  else if (p >= 0)
    dst[v * n + p] = src[v * stride + i];
}

// x[b] *= s[b]  (per design scale)
__global__ void scale_batched_kernel(int n, double *x, const double *s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[(long)blockIdx.y * n + i] *= s[blockIdx.y];
}

// y[b] += a[b] x[b]
__global__ void axpy_batched_kernel(int n, const double *a, const double *x, double *y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long o = (long)blockIdx.y * n + i;
  y[o] += a[blockIdx.y] * x[o];
}

// Batched dot products x[b] . y[b] in two passes with a fixed summation
// order (bit-reproducible from run to run): pass 1 writes one partial sum
// per (chunk, b) to part[b * gridDim.x + chunk], pass 2 adds a design's
// partials in chunk order.  Both passes use 256-thread blocks and the same
// pairwise tree.
constexpr int kDotThreads = 256;
constexpr int kDotChunks = 256;  // largest pass-1 grid.x

__device__ __forceinline__ double block_tree_sum(double v, double *red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kDotThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

__global__ __launch_bounds__(kDotThreads) void dot_batched_kernel(int n, const double *x, long xs, const double *y,
                                                                  long ys, double *part) {
  __shared__ double red[kDotThreads];
  const int b = blockIdx.y;
  double acc = 0.0;
  for (int i = blockIdx.x * kDotThreads + threadIdx.x; i < n; i += gridDim.x * kDotThreads)
    acc += x[b * xs + i] * y[b * ys + i];
  const double s = block_tree_sum(acc, red);
  if (threadIdx.x == 0) part[(long)b * gridDim.x + blockIdx.x] = s;
}

__global__ __launch_bounds__(kDotThreads) void dot_finish_kernel(int nchunks, const double *part, double *out) {
  __shared__ double red[kDotThreads];
  const int b = blockIdx.x;
  const double v = threadIdx.x < nchunks ? part[(long)b * nchunks + threadIdx.x] : 0.0;
  const double s = block_tree_sum(v, red);
  if (threadIdx.x == 0) out[b] = s;
}

// recursive halving of an n-block at r0 down to leaves of <= kCholBase
void chol_plan(int r0, int n, std::vector<CholNode> &seq) {
  if (n <= kCholBase) {
    seq.push_back({r0, n, 0, 0});
    return;
  }
  int n1 = n / 2;
  n1 -= n1 & 1;  // even split keeps 16-byte alignment of the sub-blocks
  const int n2 = n - n1;
  chol_plan(r0, n1, seq);
  seq.push_back({r0, n1, n2, 1});
  chol_plan(r0 + n1, n2, seq);
  seq.push_back({r0, n1, n2, 0});
}

int setup_plan(shellgpu_factor *F) {
  shellgpu_ctx *ctx = F->ctx;
  const int nn = ctx->nnodes;
  const std::vector<int> &rowp = ctx->h_rowp, &cols = ctx->h_cols;
  // breadth-first levels from the seed set
  std::vector<int> level(nn, -1), queue;
  queue.reserve(nn);
  std::vector<int> seeds = ctx->order_seed;
  if (seeds.empty()) {
    // pseudo-peripheral node: repeat BFS from the farthest node
    int start = 0;
    for (int it = 0; it < 4; it++) {
      std::fill(level.begin(), level.end(), -1);
      queue.clear();
      level[start] = 0;
      queue.push_back(start);
      for (size_t h = 0; h < queue.size(); h++) {
        int v = queue[h];
        for (int p = rowp[v]; p < rowp[v + 1]; p++)
          if (level[cols[p]] < 0) {
            level[cols[p]] = level[v] + 1;
            queue.push_back(cols[p]);
          }
      }
      start = queue.back();
    }
    seeds.push_back(queue.front());
  }
  std::fill(level.begin(), level.end(), -1);
  queue.clear();
  for (int s : seeds) {
    if (s < 0 || s >= nn) return ferr(ctx, 1, "ordering seed out of range");
    if (level[s] < 0) {
      level[s] = 0;
      queue.push_back(s);
    }
  }
  for (size_t h = 0; h < queue.size(); h++) {
    int v = queue[h];
    for (int p = rowp[v]; p < rowp[v + 1]; p++)
      if (level[cols[p]] < 0) {
        level[cols[p]] = level[v] + 1;
        queue.push_back(cols[p]);
      }
  }
  // unreached components start new level chains after the last level
  int nl = 0;
  for (int i = 0; i < nn; i++) nl = std::max(nl, level[i] + 1);
  for (int i = 0; i < nn; i++)
    if (level[i] < 0) return ferr(ctx, 1, "mesh graph is disconnected from the ordering seed");
  F->nlevels = nl;
  std::vector<int> cnt(nl, 0), node_pos(nn);
  for (int i = 0; i < nn; i++) cnt[level[i]]++;
  std::vector<int> noff(nl + 1, 0);
  for (int l = 0; l < nl; l++) noff[l + 1] = noff[l] + cnt[l];
  {
    std::vector<int> fill(noff.begin(), noff.end() - 1);
    for (int i = 0; i < nn; i++) node_pos[i] = fill[level[i]]++;
  }
  F->n = 6 * nn;
  F->m.assign(nl, 0);
  F->voff.assign(nl + 1, 0);
  for (int l = 0; l < nl; l++) {
    F->m[l] = 6 * cnt[l];  // even: the level GEMVs load 16-byte pairs
    F->voff[l + 1] = F->voff[l] + F->m[l];
    F->maxm = std::max(F->maxm, F->m[l]);
  }

  // Block cyclic reduction when the padded factor (three M x M blocks per
  // level and design) fits in 85 % of the free device memory, padding the
  // levels to M adds at most half the unknowns, and the factor holds few
  // designs; a factor of one design shared by the context's designs (load
  // cases) solves their vectors together (up to 8, or 16) and reads
  // itself once per solve.  Otherwise the sequential factorisation.
  // SHELLGPU_FACTOR=thomas / =cr selects either (cr only where it fits).
  {
    const long M = F->maxm;
    size_t freeb = 0, totalb = 0;
    HIPCALL(ctx, hipMemGetInfo(&freeb, &totalb));
    const double need = 3.0 * nl * (double)M * (double)M * 8.0 * F->nbatch;
    const char *env = getenv("SHELLGPU_FACTOR");
    const bool fits = nl >= 3 && need <= 0.85 * (double)freeb;
    const bool uniform = (double)nl * M <= 1.5 * F->n;
    // cyclic reduction pays off while one level's work is too small to fill
    // the GPU and the level chain is long (its rounds batch the levels);
    // with many vectors per solve the sequential factorisation is parallel
    // enough and reads fewer bytes per solve (half of one inverse per
    // level, where cyclic reduction reads three padded blocks)
    const int per = ctx->nbatch / F->nbatch;  // vectors per design of one solve
    const bool multi_ok = per == 1 || (F->nbatch == 1 && (per <= 8 || per == 16));
    F->cr = fits && uniform && (long)F->nbatch * M <= 8192 && multi_ok;
    if (env && strcmp(env, "thomas") == 0) F->cr = false;
    if (env && strcmp(env, "cr") == 0) F->cr = fits;
  }
  if (F->cr) {
    const long M = F->maxm, M2 = M * M;
    F->M = (int)M;
    F->nv = (long)nl * M;
    F->per_design = 3L * nl * M2;
    F->E_off = F->per_design;  // the sparse couplings follow the dense blocks
    F->rounds.clear();
    F->nkmax = 1;
    for (int st = 1; st < nl; st *= 2) {
      FactorRound r;
      r.s = st;
      r.nE = r.nR = 0;
      for (int i = st; i < nl; i += 2 * st) {
        r.nE++;
        if (i + st < nl) r.nR++;
      }
      r.off = 0;
      F->rounds.push_back(r);
      F->nkmax = std::max(F->nkmax, r.nE);
    }
    F->chol.clear();
    chol_plan(0, (int)M, F->chol);
  } else {
    F->nv = F->n;
    F->S_off.assign(nl, 0);
    long off = 0;
    F->tmp_elems = 1;
    for (int l = 0; l < nl; l++) {
      F->S_off[l] = off;
      off += (long)F->m[l] * F->m[l];
      if (l > 0) F->tmp_elems = std::max(F->tmp_elems, (long)F->m[l] * F->m[l - 1]);
    }
    F->E_off = off;
    // SPD-inverse scratch: per recursion depth the largest n2 x n1 block
    std::vector<long> need;
    struct Rec {
      static void go(int n, int depth, std::vector<long> &need) {
        if (n <= kCholBase) return;
        int n1 = n / 2;
        n1 -= n1 & 1;
        if ((int)need.size() <= depth) need.resize(depth + 1, 0);
        need[depth] = std::max(need[depth], (long)n1 * (n - n1));
        go(n1, depth + 1, need);
        go(n - n1, depth + 1, need);
      }
    };
    for (int l = 0; l < nl; l++) Rec::go(F->m[l], 0, need);
    F->woff.assign(need.size(), 0);
    long w = 0;
    for (size_t k = 0; k < need.size(); k++) {
      F->woff[k] = w;
      w += (need[k] + 1) & ~1L;
    }
    F->wstride = std::max(2L, w);
  }
  F->noff = noff;

  // coupling blocks (the sequential factorisation and the first round of
  // cyclic reduction), numbered in level order of their row node: CSR over
  // level-ordered nodes (to the previous level) and its transpose (from the
  // next level)
  std::vector<int> blockq(ctx->nnzb, -1);
  {
    std::vector<int> pos2node(nn);
    for (int i = 0; i < nn; i++) pos2node[node_pos[i]] = i;
    std::vector<int> cp_rowp(nn + 1, 0), cp_col, ct_cnt(nn + 1, 0);
    std::vector<int> cq_row;  // per block: level-ordered row node
    for (int u = 0; u < nn; u++) {
      const int i = pos2node[u], li = level[i];
      for (int p = rowp[i]; p < rowp[i + 1]; p++) {
        const int j = cols[p];
        if (level[j] != li - 1) continue;
        blockq[p] = (int)cp_col.size();
        cp_col.push_back(node_pos[j] - noff[li - 1]);
        cq_row.push_back(u);
        ct_cnt[node_pos[j] + 1]++;
      }
      cp_rowp[u + 1] = (int)cp_col.size();
    }
    F->ncoup = (long)cp_col.size();
    for (int u = 0; u < nn; u++) ct_cnt[u + 1] += ct_cnt[u];
    std::vector<int> ct_src(F->ncoup), ct_blk(F->ncoup), fill(ct_cnt.begin(), ct_cnt.end() - 1);
    for (int u = 0; u < nn; u++) {
      const int i = pos2node[u];
      for (int p = rowp[i]; p < rowp[i + 1]; p++) {
        const int q = blockq[p];
        if (q < 0) continue;
        const int v = node_pos[cols[p]], k = fill[v]++;
        ct_src[k] = u - noff[level[i]];
        ct_blk[k] = q;
      }
    }
    F->per_design = F->E_off + 36 * F->ncoup;
    const size_t nc = std::max<size_t>(1, (size_t)F->ncoup);
    HIPCALL(ctx, hipMalloc((void **)&F->d_cp_rowp, sizeof(int) * (nn + 1)));
    HIPCALL(ctx, hipMalloc((void **)&F->d_cp_col, sizeof(int) * nc));
    HIPCALL(ctx, hipMalloc((void **)&F->d_ct_rowp, sizeof(int) * (nn + 1)));
    HIPCALL(ctx, hipMalloc((void **)&F->d_ct_src, sizeof(int) * nc));
    HIPCALL(ctx, hipMalloc((void **)&F->d_ct_blk, sizeof(int) * nc));
    HIPCALL(ctx, hipMemcpy(F->d_cp_rowp, cp_rowp.data(), sizeof(int) * (nn + 1), hipMemcpyHostToDevice));
    HIPCALL(ctx, hipMemcpy(F->d_ct_rowp, ct_cnt.data(), sizeof(int) * (nn + 1), hipMemcpyHostToDevice));
    if (F->ncoup > 0) {
      HIPCALL(ctx, hipMemcpy(F->d_cp_col, cp_col.data(), sizeof(int) * F->ncoup, hipMemcpyHostToDevice));
      HIPCALL(ctx, hipMemcpy(F->d_ct_src, ct_src.data(), sizeof(int) * F->ncoup, hipMemcpyHostToDevice));
      HIPCALL(ctx, hipMemcpy(F->d_ct_blk, ct_blk.data(), sizeof(int) * F->ncoup, hipMemcpyHostToDevice));
    }
  }

  // BC dofs (symmetric elimination)
  std::vector<int> isbc(F->n, 0);
  for (size_t b = 0; b < ctx->h_bc_nodes.size(); b++)
    for (int r = 0; r < 6; r++)
      if (ctx->h_bc_mask[b] & (1 << r)) isbc[6 * ctx->h_bc_nodes[b] + r] = 1;

  // dof permutation (level order, padded to M per level under cyclic
  // reduction) and block scatter map
  std::vector<int> pos2dof(F->nv, -1);
  for (int i = 0; i < nn; i++) {
    const int li = level[i];
    const long p0 = F->cr ? (long)li * F->M + 6L * (node_pos[i] - noff[li]) : 6L * node_pos[i];
    for (int d = 0; d < 6; d++) pos2dof[p0 + d] = 6 * i + d;
  }
  std::vector<long> base(ctx->nnzb);
  std::vector<int> ld(ctx->nnzb), colbc(ctx->nnzb);
  const long M2 = (long)F->M * F->M;
  for (int i = 0; i < nn; i++) {
    const int li = level[i];
    const long rl = 6L * (node_pos[i] - noff[li]);
    for (int p = rowp[i]; p < rowp[i + 1]; p++) {
      const int j = cols[p], lj = level[j];
      const long cl = 6L * (node_pos[j] - noff[lj]);
      int mask = 0;
      for (int d = 0; d < 6; d++) mask |= (isbc[6 * i + d] << d) | (isbc[6 * j + d] << (6 + d));
      if (i == j) mask |= 1 << 12;
      colbc[p] = mask;
      ld[p] = F->cr ? F->M : F->m[li];
      if (lj == li) {
        base[p] = F->cr ? (long)li * M2 + cl * F->M + rl : F->S_off[li] + cl * F->m[li] + rl;
      } else if (lj == li - 1) {
        base[p] = F->E_off + 36L * blockq[p];  // a 6 x 6 column-major block
        ld[p] = 6;
      } else if (lj == li + 1) {
        base[p] = -1;  // upper block: the transpose of the next level's coupling
      } else {
        return ferr(ctx, 1, "level ordering is not block tridiagonal");
      }
    }
  }
  if (F->cr) {
    HIPCALL(ctx, hipMalloc((void **)&F->d_mlev, sizeof(int) * nl));
    HIPCALL(ctx, hipMemcpy(F->d_mlev, F->m.data(), sizeof(int) * nl, hipMemcpyHostToDevice));
    HIPCALL(ctx, hipMalloc((void **)&F->d_noff, sizeof(int) * (nl + 1)));
    HIPCALL(ctx, hipMemcpy(F->d_noff, noff.data(), sizeof(int) * (nl + 1), hipMemcpyHostToDevice));
  }
  HIPCALL(ctx, hipMalloc((void **)&F->d_bcflag, sizeof(int) * ctx->nnzb));
  HIPCALL(ctx, hipMemcpyAsync(F->d_bcflag, colbc.data(), sizeof(int) * ctx->nnzb, hipMemcpyHostToDevice,
                              ctx->stream));
  HIPCALL(ctx, hipMalloc((void **)&F->d_pos2dof, sizeof(int) * F->nv));
  HIPCALL(ctx, hipMalloc((void **)&F->d_blk_base, sizeof(long) * ctx->nnzb));
  HIPCALL(ctx, hipMalloc((void **)&F->d_blk_ld, sizeof(int) * ctx->nnzb));
  HIPCALL(ctx, hipMemcpyAsync(F->d_pos2dof, pos2dof.data(), sizeof(int) * F->nv, hipMemcpyHostToDevice, ctx->stream));
  HIPCALL(ctx, hipMemcpyAsync(F->d_blk_base, base.data(), sizeof(long) * ctx->nnzb, hipMemcpyHostToDevice,
                              ctx->stream));
  HIPCALL(ctx, hipMemcpyAsync(F->d_blk_ld, ld.data(), sizeof(int) * ctx->nnzb, hipMemcpyHostToDevice, ctx->stream));
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int ensure_blas(shellgpu_ctx *ctx) {
  if (!ctx->blas) {
    rocblas_handle h;
    BLASCALL(ctx, rocblas_create_handle(&h));
    ctx->blas = h;
  }
  BLASCALL(ctx, rocblas_set_stream((rocblas_handle)ctx->blas, ctx->stream));
  BLASCALL(ctx, rocblas_set_pointer_mode((rocblas_handle)ctx->blas, rocblas_pointer_mode_host));
  return 0;
}

// ---- symmetric tridiagonal eigenproblem (implicit QL with Wilkinson
// shifts, the EISPACK tql2 algorithm); z[k*n + i] = component i of vector k.
// last_only: z[k] = component n - 1 of vector k only (each component is
// rotated independently, so the last one alone costs O(n^2): enough for
// the Ritz residual estimates)
int tridiag_ql(int n, std::vector<double> d, std::vector<double> e, std::vector<double> &w,
               std::vector<double> &z, bool last_only = false) {
  e.resize(n + 1, 0.0);
  const int nz = last_only ? 1 : n, k0 = last_only ? n - 1 : 0;
  z.assign((size_t)n * nz, 0.0);
  for (int i = 0; i < n; i++)
    if (!last_only || i == n - 1) z[(size_t)i * nz + i - k0] = 1.0;
  for (int l = 0; l < n; l++) {
    int iter = 0, mm;
    do {
      for (mm = l; mm < n - 1; mm++) {
        const double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
        if (std::fabs(e[mm]) <= 1e-16 * dd) break;
      }
      if (mm != l) {
        if (iter++ == 60) return -1;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = mm - 1; i >= l; i--) {
          const double f = s * e[i], b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[mm] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          for (int k = 0; k < nz; k++) {
            const double t = z[(size_t)(i + 1) * nz + k];
            z[(size_t)(i + 1) * nz + k] = s * z[(size_t)i * nz + k] + c * t;
            z[(size_t)i * nz + k] = c * z[(size_t)i * nz + k] - s * t;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[mm] = 0.0;
      }
    } while (mm != l);
  }
  w = d;
  return 0;
}

// Cholesky-inverse node sequence of an n-block and, per node, the pointer
// arrays (A11, A21, A22, A12 and the scratch W[q]; leaves A11 only) of the
// blocks D[q] (lda); W[q] holds n^2 / 4 doubles
void add_chol_group(shellgpu_factor *F, std::vector<double *> &h, const std::vector<double *> &D,
                    const std::vector<double *> &W, int n, int lda) {
  shellgpu_factor::CholGroup grp;
  grp.n = n;
  grp.lda = lda;
  grp.cnt = (int)D.size();
  chol_plan(0, n, grp.nodes);
  for (const CholNode &c : grp.nodes) {
    grp.offs.push_back((long)h.size());
    const long a11 = (long)c.r0 * (lda + 1), a21 = a11 + c.n1, a22 = a11 + (long)c.n1 * (lda + 1),
               a12 = a11 + (long)c.n1 * lda;
    for (double *p : D) h.push_back(p + a11);
    if (c.n2 > 0) {
      for (double *p : D) h.push_back(p + a21);
      for (double *p : D) h.push_back(p + a22);
      for (double *p : D) h.push_back(p + a12);
      for (double *p : W) h.push_back(p);
    }
  }
  F->cgroups.push_back(grp);
}

// Pointer arrays of every cyclic-reduction round, entry q = k nb + d
// (level entry k, design d): D_i, B_i, Z_i, D_{i-s}, D_{i+s}, B_{i+s},
// B_{i-1} for i = s (2k + 1); D_{i+s} and B_{i+s} exist for the first nR nb
// entries only.  B_{i-1} and D_{i-1} (odd levels: free after the first
// round) are the scratch of the later rounds: Y_i is stored in B_{i-1}.
// A final group of nb entries points at D_0 (the root).
int build_cr_pointers(shellgpu_factor *F) {
  shellgpu_ctx *ctx = F->ctx;
  const int nb = F->nbatch, nl = F->nlevels;
  const long M2 = (long)F->M * F->M, PD = F->per_design;
  std::vector<double *> h;
  for (FactorRound &r : F->rounds) {
    r.off = (long)h.size();
    const long cnt = (long)nb * r.nE;
    h.resize(h.size() + 7 * cnt, nullptr);
    double **P = h.data() + r.off;
    for (int k = 0; k < r.nE; k++) {
      const int i = r.s * (2 * k + 1);
      for (int d = 0; d < nb; d++) {
        const long q = (long)k * nb + d;
        double *base = F->buf + d * PD;
        P[0 * cnt + q] = base + (long)i * M2;
        P[1 * cnt + q] = base + (long)(nl + i) * M2;
        P[2 * cnt + q] = base + (long)(2 * nl + i) * M2;
        P[3 * cnt + q] = base + (long)(i - r.s) * M2;
        if (i + r.s < nl) {
          P[4 * cnt + q] = base + (long)(i + r.s) * M2;
          P[5 * cnt + q] = base + (long)(nl + i + r.s) * M2;
        } else {
          P[4 * cnt + q] = P[3 * cnt + q];  // unused (entries beyond nR nb)
          P[5 * cnt + q] = P[1 * cnt + q];
        }
        P[6 * cnt + q] = base + (long)(nl + i - 1) * M2;
      }
    }
  }
  const long root = (long)h.size();
  for (int d = 0; d < nb; d++) h.push_back(F->buf + d * PD);
  // Cholesky-inverse scratch: B_i in the first round (its coupling products
  // come after), D_{i-1} in the later rounds, B_0 (never used) for the root
  F->cgroups.clear();
  for (size_t g = 0; g <= F->rounds.size(); g++) {
    const bool is_root = g == F->rounds.size();
    const long cnt = is_root ? nb : (long)nb * F->rounds[g].nE;
    const long src = is_root ? root : F->rounds[g].off;  // D pointers of the group
    std::vector<double *> D(h.begin() + src, h.begin() + src + cnt), W(cnt);
    for (long q = 0; q < cnt; q++) {
      if (is_root)
        W[q] = D[q] + (long)nl * M2;
      else if (F->rounds[g].s == 1)
        W[q] = h[src + cnt + q];
      else
        W[q] = D[q] - M2;
    }
    add_chol_group(F, h, D, W, F->M, F->M);
  }
  HIPCALL(ctx, hipMalloc((void **)&F->d_ptr, sizeof(double *) * h.size()));
  HIPCALL(ctx, hipMemcpy(F->d_ptr, h.data(), sizeof(double *) * h.size(), hipMemcpyHostToDevice));
  F->root_off = root;

  // jobs of the first round (odd levels i, left neighbour l = i - 1, right
  // r = i + 1): the D_i slot holds X_i, then the transposes; Z_i holds P_i;
  // the B_i slot holds Q_i = P_i E_i, then V_i = P_i E_r^T
  const FactorRound &r1 = F->rounds[0];
  std::vector<CoupleJob> jobs;
  auto add = [&](long &off, bool right_only, auto make) {
    off = (long)jobs.size();
    for (int k = 0; k < (right_only ? r1.nR : r1.nE); k++) {
      const int i = 2 * k + 1;
      for (int d = 0; d < nb; d++) {
        double *base = F->buf + d * PD;
        double *Di = base + (long)i * M2, *Bi = base + (long)(nl + i) * M2, *Zi = base + (long)(2 * nl + i) * M2;
        double *Dl = base + (long)(i - 1) * M2;
        double *Dr = right_only ? base + (long)(i + 1) * M2 : nullptr;
        double *Br = right_only ? base + (long)(nl + i + 1) * M2 : nullptr;
        const double *E = base + F->E_off;
        const int mi = F->m[i], ml = F->m[i - 1], mr = right_only ? F->m[i + 1] : 0;
        jobs.push_back(make(i, Di, Bi, Zi, Dl, Dr, Br, E, mi, ml, mr));
      }
    }
  };
  add(F->job_q, false, [&](int i, double *, double *Bi, double *Zi, double *, double *, double *, const double *E,
                           int mi, int ml, int) { return CoupleJob{Zi, Bi, E, mi, ml, i - 1, 0}; });
  add(F->job_t1, false, [&](int, double *Di, double *Bi, double *, double *, double *, double *, const double *E,
                            int mi, int ml, int) { return CoupleJob{Bi, Di, E, mi, ml, 0, 0}; });
  add(F->job_dl, false, [&](int i, double *Di, double *, double *, double *Dl, double *, double *, const double *E,
                            int mi, int ml, int) { return CoupleJob{Di, Dl, E, ml, mi, i - 1, 0}; });
  add(F->job_v, true, [&](int i, double *, double *Bi, double *Zi, double *, double *, double *, const double *E,
                          int mi, int, int mr) { return CoupleJob{Zi, Bi, E, mi, mr, i + 1, 0}; });
  add(F->job_t2, true, [&](int, double *Di, double *Bi, double *, double *, double *, double *, const double *E,
                           int mi, int, int mr) { return CoupleJob{Bi, Di, E, mi, mr, 0, 0}; });
  add(F->job_dr, true, [&](int i, double *Di, double *, double *, double *, double *Dr, double *, const double *E,
                           int mi, int, int mr) { return CoupleJob{Di, Dr, E, mr, mi, i + 1, 0}; });
  add(F->job_br, true, [&](int i, double *Di, double *, double *, double *, double *, double *Br, const double *E,
                           int mi, int, int mr) { return CoupleJob{Di, Br, E, mr, mi, i - 1, 0}; });
  HIPCALL(ctx, hipMalloc((void **)&F->d_jobs, sizeof(CoupleJob) * jobs.size()));
  HIPCALL(ctx, hipMemcpy(F->d_jobs, jobs.data(), sizeof(CoupleJob) * jobs.size(), hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

extern "C" {

int shellgpu_set_ordering_seed(shellgpu_ctx *ctx, int n, const int *nodes) {
  if (!ctx || n < 0) return 1;
  ctx->order_seed.assign(nodes, nodes + n);
  return 0;
}

}  // extern "C"

namespace {

// a factor of nbatch designs: the context's (nbatch = ctx->nbatch) or, for
// identical designs, one shared by all (nbatch = 1)
int factor_create(shellgpu_ctx *ctx, int nbatch, shellgpu_factor **out) {
  if (!ctx || !out) return 1;
  if (ctx->nnzb <= 0 || ctx->nbatch <= 0) return ferr(ctx, 1, "set the mesh and designs before the factor");
  int rc = ensure_blas(ctx);
  if (rc) return rc;
  shellgpu_factor *F = new shellgpu_factor();
  F->ctx = ctx;
  F->nbatch = nbatch;
  if ((rc = setup_plan(F))) {
    delete F;
    return rc;
  }
  const int nb = F->nbatch;
  const size_t bytes = sizeof(double) * (size_t)F->per_design * nb;
  hipError_t e = hipMalloc((void **)&F->buf, bytes);
  if (e == hipSuccess && !F->cr)
    e = hipMalloc((void **)&F->tmp, sizeof(double) * 2 * std::max(1L, F->tmp_elems) * nb);
  if (e == hipSuccess && !F->cr) e = hipMalloc((void **)&F->wbuf, sizeof(double) * F->wstride * nb);
  if (e == hipSuccess) e = hipMalloc((void **)&F->info, sizeof(int) * nb);
  if (e == hipSuccess) e = hipMalloc((void **)&F->d_act, sizeof(int) * nb);
  if (e == hipSuccess) e = hipMalloc((void **)&F->vb, sizeof(double) * (size_t)F->nv * nb);
  if (e == hipSuccess) e = hipMalloc((void **)&F->vz, sizeof(double) * (size_t)F->nv * nb);
  if (e == hipSuccess) e = hipMalloc((void **)&F->vx, sizeof(double) * (size_t)F->nv * nb);
  if (e == hipSuccess) {
    // GEMV partial sums of cyclic reduction; symmetric-product tile sums of
    // the sequential factorisation (64 x 64 tiles)
    const long nT = (F->maxm + kSymvTile - 1) / kSymvTile;
    // (cyclic reduction: GEMV splits of any round, and the first round's
    // symmetric products of its nE entries per design)
    const long parts = F->cr ? std::max((long)F->maxm * 32 * nb * F->nkmax, nT * nT * kSymvTile * nb * F->rounds[0].nE)
                             : nT * nT * kSymvTile * nb;
    e = hipMalloc((void **)&F->part, sizeof(double) * (size_t)parts);
  }
  if (e == hipSuccess) e = hipMalloc((void **)&F->cr_info, sizeof(int) * 2 * nb * (F->nlevels + 1));
  if (e != hipSuccess) {
    shellgpu_factor_destroy(F);
    return ferr(ctx, 2, std::string("factor allocation (") + std::to_string(bytes >> 20) +
                            " MiB): " + hipGetErrorString(e));
  }
  if (F->cr && (rc = build_cr_pointers(F))) {
    shellgpu_factor_destroy(F);
    return rc;
  }
  *out = F;
  return 0;
}

}  // namespace

extern "C" {

int shellgpu_factor_create(shellgpu_ctx *ctx, shellgpu_factor **out) {
  if (!ctx) return 1;
  return factor_create(ctx, ctx->nbatch, out);
}

void shellgpu_factor_destroy(shellgpu_factor *F) {
  if (!F) return;
  (void)hipStreamSynchronize(F->ctx->stream);
  if (F->solve_graph) (void)hipGraphExecDestroy(F->solve_graph);
  void *bufs[] = {F->buf,      F->tmp,       F->info,      F->d_bcflag,  F->d_pos2dof, F->d_blk_base,
                  F->d_blk_ld, F->vb,        F->vz,        F->vx,        F->part,      F->d_mlev,
                  F->d_ptr,    F->cr_info,   F->d_cp_rowp, F->d_cp_col,  F->d_ct_rowp, F->d_ct_src,
                  F->d_ct_blk, F->wbuf,      F->d_act,     F->d_noff,    F->d_jobs,    F->crm_b,
                  F->crm_z,    F->crm_x,     F->crm_part,  F->d_crm_act};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  delete F;
}

int shellgpu_factor_info(shellgpu_factor *F, int *nlevels, int *max_level_dofs, double *gib) {
  if (!F) return 1;
  if (nlevels) *nlevels = F->nlevels;
  if (max_level_dofs) *max_level_dofs = F->maxm;
  if (gib) *gib = sizeof(double) * (double)F->per_design * F->nbatch / (1024.0 * 1024.0 * 1024.0);
  return 0;
}

static int factor_cr(shellgpu_factor *F);
static int apply_cr(shellgpu_factor *F);
static int factor_schur(shellgpu_factor *F);
static int apply_schur(shellgpu_factor *F, int nrhs);

int shellgpu_factor_compute(shellgpu_factor *F, shellgpu_mat *A) {
  shellgpu_ctx *ctx = F->ctx;
  // a shared factor (nbatch 1) takes the first design's matrix
  if (A->nbatch != F->nbatch && F->nbatch != 1) return ferr(ctx, 1, "matrix and factor batch sizes differ");
  int rc = ensure_blas(ctx);
  if (rc) return rc;
  const int nb = F->nbatch;
  const long ps = F->per_design;
  F->factored = false;
  HIPCALL(ctx, hipMemsetAsync(F->buf, 0, sizeof(double) * (size_t)ps * nb, ctx->stream));
  {
    const long nthreads = 36L * ctx->nnzb;
    hipLaunchKernelGGL(scatter_dense_kernel, dim3((unsigned)((nthreads + 255) / 256), nb), dim3(256), 0, ctx->stream,
                       ctx->nnzb, A->vals, 36L * ctx->nnzb, F->d_blk_base, F->d_blk_ld, F->d_bcflag,
                       F->buf, ps);
    HIPCALL(ctx, hipGetLastError());
  }
  if (F->cr) return factor_cr(F);
  if ((rc = factor_schur(F))) return rc;
  std::vector<int> all((long)F->nlevels * nb), info(nb, 0);
  HIPCALL(ctx, hipMemcpyAsync(all.data(), F->cr_info, sizeof(int) * all.size(), hipMemcpyDeviceToHost, ctx->stream));
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  for (size_t q = 0; q < all.size(); q++)
    if (all[q] != 0) info[q % nb] = 1;
  HIPCALL(ctx, hipMemcpy(F->info, info.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
  for (int b = 0; b < nb; b++)
    if (info[b] != 0)
      return ferr(ctx, 5, "matrix not positive definite in the level Cholesky (design " + std::to_string(b) +
                              "): for K + sigma G the shift is above the first buckling load");
  F->factored = true;
  return 0;
}

// y = beta y + alpha op(A) x over nq batch entries (GemvBatch strides)
static int gemv_batch(shellgpu_factor *F, bool trans, int m, int n, const double *A, int lda, const double *x,
                      double alpha, double beta, double *y, int lower, const GemvBatch &g, int nq) {
  shellgpu_ctx *ctx = F->ctx;
  if (nq <= 0) return 0;
  if (!trans) {
    // column chunks per row block: enough workgroups to fill the GPU
    // (about 2,048), and as few as that allows (longer column loops)
    const long rb = (m + 127) / 128;
    const int nsplit = (int)std::max<long>(1, std::min<long>({32, n / 64, (2048 + rb * nq - 1) / (rb * nq)}));
    const int chunk = (n + nsplit - 1) / nsplit;
    const dim3 grid((m + 127) / 128, nsplit, nq);
    if (lower)
      hipLaunchKernelGGL(gemv_n_pass1<1>, grid, dim3(256), 0, ctx->stream, m, n, A, lda, x, g, chunk, F->part);
    else
      hipLaunchKernelGGL(gemv_n_pass1<0>, grid, dim3(256), 0, ctx->stream, m, n, A, lda, x, g, chunk, F->part);
    HIPCALL(ctx, hipGetLastError());
    hipLaunchKernelGGL(gemv_n_pass2, dim3((m + 255) / 256, nq), dim3(256), 0, ctx->stream, m, nsplit, F->part, alpha,
                       beta, y, g);
    HIPCALL(ctx, hipGetLastError());
  } else {
    if (lower)
      hipLaunchKernelGGL(gemv_t_kernel<1>, dim3((n + 3) / 4, nq), dim3(256), 0, ctx->stream, m, n, A, lda, x, g,
                         alpha, beta, y);
    else
      hipLaunchKernelGGL(gemv_t_kernel<0>, dim3((n + 3) / 4, nq), dim3(256), 0, ctx->stream, m, n, A, lda, x, g,
                         alpha, beta, y);
    HIPCALL(ctx, hipGetLastError());
  }
  return 0;
}

// y = beta c + alpha P x for nrhs vectors of every design (symmetric P of
// level l of the sequential factorisation; vector entries e = r nb + d of
// stride nv): one vector per design reads half of P, several read it in
// full once for all (nrhs <= 8 or 16)
constexpr int kMaxRhs = 16;
// vectors per pass when a factor of one design solves n vectors: 16 while
// there are that many (P read once for 16), else up to 8
inline int rhs_chunk(int n) { return n >= kMaxRhs ? kMaxRhs : std::min(8, n); }
static int symv(shellgpu_factor *F, int l, const double *x, double alpha, const double *c, double beta, double *y,
                int nrhs) {
  shellgpu_ctx *ctx = F->ctx;
  const int m = F->m[l], nT = (m + kSymvTile - 1) / kSymvTile, nb = F->nbatch;
  const double *P = F->buf + F->S_off[l];
  const long ps = F->per_design, nv = F->nv;
  // active designs of a single-vector solve (Lanczos skipping converged ones)
  const int *act = (nrhs == 1 && F->nact >= 0) ? F->d_act : nullptr;
  // column slabs per tile when the tiles alone would not fill the GPU
  const int cs = (nrhs > 1 && (long)nT * nT * nb < 512) ? kMaxSplit : 1;
  // one stride: entry e (design, or vector r nb + d) at e ps / e nv
  const SymvBatch sb{1 << 30, ps, 0, nv, 0, nv, 0};
  switch (nrhs * 8 + (cs > 1)) {
#define SYMV_MULTI(R)                                                                                               \
  case R * 8:                                                                                                       \
    hipLaunchKernelGGL((symv_multi_pass1_kernel<R, 1>), dim3(nT * nT, nb), dim3(256), 0, ctx->stream, m, nT, P, m,  \
                       ps, x, nv, nb, F->part);                                                                     \
    break;                                                                                                          \
  case R * 8 + 1:                                                                                                   \
    hipLaunchKernelGGL((symv_multi_pass1_kernel<R, kMaxSplit>), dim3(nT * nT * kMaxSplit, nb), dim3(256), 0,       \
                       ctx->stream, m, nT, P, m, ps, x, nv, nb, F->part);                                          \
    break;
    case 8:
      hipLaunchKernelGGL(symv_pass1_kernel, dim3(nT * (nT + 1) / 2, act ? F->nact : nb), dim3(256), 0, ctx->stream,
                         m, nT, P, m, x, sb, F->part, act);
      break;
      SYMV_MULTI(2)
      SYMV_MULTI(3)
      SYMV_MULTI(4)
      SYMV_MULTI(5)
      SYMV_MULTI(6)
      SYMV_MULTI(7)
      SYMV_MULTI(8)
      SYMV_MULTI(16)
#undef SYMV_MULTI
    default:
      return ferr(ctx, 1, "too many right-hand sides for one level product");
  }
  HIPCALL(ctx, hipGetLastError());
  // about 4 slots per thread
  const int ns = nT * cs;
  const dim3 g2((m + 63) / 64, act ? F->nact : nrhs * nb);
  if (ns >= 64)
    hipLaunchKernelGGL(symv_pass2_kernel<16>, g2, dim3(1024), 0, ctx->stream, m, nT, ns, F->part, alpha, c, beta, y,
                       sb, act);
  else if (ns >= 32)
    hipLaunchKernelGGL(symv_pass2_kernel<8>, g2, dim3(512), 0, ctx->stream, m, nT, ns, F->part, alpha, c, beta, y, sb,
                       act);
  else if (ns >= 16)
    hipLaunchKernelGGL(symv_pass2_kernel<4>, g2, dim3(256), 0, ctx->stream, m, nT, ns, F->part, alpha, c, beta, y, sb,
                       act);
  else if (ns >= 8)
    hipLaunchKernelGGL(symv_pass2_kernel<2>, g2, dim3(128), 0, ctx->stream, m, nT, ns, F->part, alpha, c, beta, y, sb,
                       act);
  else
    hipLaunchKernelGGL(symv_pass2_kernel<1>, g2, dim3(64), 0, ctx->stream, m, nT, ns, F->part, alpha, c, beta, y, sb,
                       act);
  HIPCALL(ctx, hipGetLastError());
  return 0;
}

// y_q = beta c_q + alpha P_q x_q for the first round of cyclic reduction:
// entries q = k nb + d (odd level 2k + 1, design d), P_q symmetric in the
// Z slot of the level (P_i = D_i^{-1}), half of it read
static int symv_cr(shellgpu_factor *F, const double *x, double alpha, const double *c, double beta, double *y) {
  shellgpu_ctx *ctx = F->ctx;
  const int M = F->M, nT = (M + kSymvTile - 1) / kSymvTile, nb = F->nbatch, cnt = nb * F->rounds[0].nE;
  const long M2 = (long)M * M, PD = F->per_design, NV = F->nv;
  const double *P = F->buf + 2L * F->nlevels * M2 + M2;  // Z slot of level 1
  const SymvBatch sb{nb, PD, 2 * M2, NV, 2L * M, NV, 2L * M};
  hipLaunchKernelGGL(symv_pass1_kernel, dim3(nT * (nT + 1) / 2, cnt), dim3(256), 0, ctx->stream, M, nT, P, M, x, sb,
                     F->part, nullptr);
  HIPCALL(ctx, hipGetLastError());
  const dim3 g2((M + 63) / 64, cnt);
  if (nT >= 16)
    hipLaunchKernelGGL(symv_pass2_kernel<4>, g2, dim3(256), 0, ctx->stream, M, nT, nT, F->part, alpha, c, beta, y, sb,
                       nullptr);
  else
    hipLaunchKernelGGL(symv_pass2_kernel<1>, g2, dim3(64), 0, ctx->stream, M, nT, nT, F->part, alpha, c, beta, y, sb,
                       nullptr);
  HIPCALL(ctx, hipGetLastError());
  return 0;
}

// X = chol(A)^{-1} (lower) in place for the cnt diagonal blocks of group g:
// leaves in LDS; internal nodes on the matrix cores, out of place through
// the block's upper triangle A12 (unused: X is lower triangular) and the
// scratch W:
//   A12 = X11 A21^T (= C21^T),  A22 -= A12^T A12 (lower),  (trailing half),
//   W = X11^T A12 (= (C21 X11)^T),  A21 = -X22 W^T (= X21).
static int cholinv(shellgpu_factor *F, size_t g, int *info) {
  shellgpu_ctx *ctx = F->ctx;
  const shellgpu_factor::CholGroup &grp = F->cgroups[g];
  const int M = grp.lda, cnt = grp.cnt;
  for (size_t k = 0; k < grp.nodes.size(); k++) {
    const CholNode &c = grp.nodes[k];
    double **P = F->d_ptr + grp.offs[k];
    double **A11 = P, **A21 = P + cnt, **A22 = P + 2 * (long)cnt, **A12 = P + 3 * (long)cnt,
           **W = P + 4 * (long)cnt;
    const int n1 = c.n1, n2 = c.n2;
    int rc = 0;
    if (n2 == 0) {
      hipLaunchKernelGGL(cholinv_base_kernel, dim3(cnt), dim3(kCholThreads), 0, ctx->stream, A11, n1, M, info);
      HIPCALL(ctx, hipGetLastError());
    } else if (c.first) {
      if ((rc = mgemm(ctx, false, true, n1, n2, n1, 1.0, A11, M, kGemmLower, A21, M, kGemmFull, 0.0, A12, M, 0,
                      cnt)) ||
          (rc = mgemm(ctx, true, false, n2, n2, n1, -1.0, A12, M, kGemmFull, A12, M, kGemmFull, 1.0, A22, M, 1,
                      cnt)))
        return rc;
    } else {
      if ((rc = mgemm(ctx, true, false, n1, n2, n1, 1.0, A11, M, kGemmUpper, A12, M, kGemmFull, 0.0, W, n1, 0,
                      cnt)) ||
          (rc = mgemm(ctx, false, true, n2, n1, n2, -1.0, A22, M, kGemmLower, W, n1, kGemmFull, 0.0, A21, M, 0,
                      cnt)))
        return rc;
    }
  }
  return 0;
}

// In-place inverse of the SPD diagonal block [r0, r0 + n) of level l's
// S_l (every design, strided), by recursive halving down to LDS leaves:
//   A11 <- A11^{-1},  W = A21 A11,  A22 <- (A22 - W A21^T)^{-1},
//   A21 <- -A22 W,  A11 <- A11 - W^T A21,  A12 <- A21^T
// (block elimination of an SPD matrix: every step is a GEMM; W of each
// recursion depth has its own scratch)
static int spdinv(shellgpu_factor *F, int l, int r0, int n, int depth, int *info) {
  shellgpu_ctx *ctx = F->ctx;
  const int nb = F->nbatch, m = F->m[l];
  const long ps = F->per_design;
  double *A11 = F->buf + F->S_off[l] + (long)r0 * (m + 1);
  if (n <= kCholBase) {
    hipLaunchKernelGGL(spdinv_base_kernel, dim3(nb), dim3(kCholThreads), 0, ctx->stream, A11, ps, n, m, info);
    HIPCALL(ctx, hipGetLastError());
    return 0;
  }
  int n1 = n / 2;
  n1 -= n1 & 1;
  const int n2 = n - n1;
  double *A21 = A11 + n1, *A22 = A11 + (long)n1 * (m + 1), *A12 = A11 + (long)n1 * m;
  double *W = F->wbuf + F->woff[depth];
  const long ws = F->wstride;
  const double one = 1.0, mone = -1.0, zero = 0.0;
  int rc = 0;
  if ((rc = spdinv(F, l, r0, n1, depth + 1, info))) return rc;
  if ((rc = gemm_sb(ctx, false, false, n2, n1, n1, one, A21, m, ps, A11, m, ps, zero, W, n2, ws, nb))) return rc;
  if ((rc = gemm_sb(ctx, false, true, n2, n2, n1, mone, W, n2, ws, A21, m, ps, one, A22, m, ps, nb))) return rc;
  if ((rc = spdinv(F, l, r0 + n1, n2, depth + 1, info))) return rc;
  if ((rc = gemm_sb(ctx, false, false, n2, n1, n2, mone, A22, m, ps, W, n2, ws, zero, A21, m, ps, nb))) return rc;
  if ((rc = gemm_sb(ctx, true, false, n1, n1, n2, mone, W, n2, ws, A21, m, ps, one, A11, m, ps, nb))) return rc;
  hipLaunchKernelGGL(transpose_kernel, dim3((n2 + 31) / 32, (n1 + 31) / 32, nb), dim3(256), 0, ctx->stream, n2, n1,
                     A21, m, ps, A12, m, ps);
  HIPCALL(ctx, hipGetLastError());
  return 0;
}

// The sequential factorisation after the scatter: per level,
//   T = P_{l-1} E_l^T,  U = T^T (= E_l P_{l-1}),  S_l = D_l - U E_l^T,
//   P_l = S_l^{-1} (in full)
static int factor_schur(shellgpu_factor *F) {
  shellgpu_ctx *ctx = F->ctx;
  const int nb = F->nbatch;
  const long ps = F->per_design, ts = F->tmp_elems;
  double *T = F->tmp, *U = F->tmp + ts * nb, *E = F->buf + F->E_off;
  int rc = 0;
  HIPCALL(ctx, hipMemsetAsync(F->cr_info, 0, sizeof(int) * (long)F->nlevels * nb, ctx->stream));
  for (int l = 0; l < F->nlevels; l++) {
    const int ml = F->m[l], nn = ml / 6;
    double *S = F->buf + F->S_off[l];
    if (l > 0) {
      const int mp = F->m[l - 1];
      const int *rowp = F->d_cp_rowp + F->noff[l];
      const int ng = (nn + kCoupleNodes - 1) / kCoupleNodes;
      hipLaunchKernelGGL(couple_right_kernel, dim3((mp + 255) / 256, ng, nb), dim3(256), 0, ctx->stream, mp, nn,
                         F->buf + F->S_off[l - 1], mp, ps, T, mp, ts, rowp, F->d_cp_col, E, ps, 1.0, 0.0);
      HIPCALL(ctx, hipGetLastError());
      hipLaunchKernelGGL(transpose_kernel, dim3((mp + 31) / 32, (ml + 31) / 32, nb), dim3(256), 0, ctx->stream, mp,
                         ml, T, mp, ts, U, ml, ts);
      HIPCALL(ctx, hipGetLastError());
      hipLaunchKernelGGL(couple_right_kernel, dim3((ml + 255) / 256, ng, nb), dim3(256), 0, ctx->stream, ml, nn, U,
                         ml, ts, S, ml, ps, rowp, F->d_cp_col, E, ps, -1.0, 1.0);
      HIPCALL(ctx, hipGetLastError());
    }
    if ((rc = spdinv(F, l, 0, ml, 0, F->cr_info + (long)l * nb))) return rc;
  }
  return 0;
}

// x = A^{-1} b by the sequential factorisation for nrhs vectors of every
// design (level-ordered vb -> vx, entries e = r nb + d)
static int apply_schur(shellgpu_factor *F, int nrhs) {
  shellgpu_ctx *ctx = F->ctx;
  const int nb = F->nbatch, nl = F->nlevels;
  const int *act = (nrhs == 1 && F->nact >= 0) ? F->d_act : nullptr;
  const int ne = act ? F->nact : nrhs * nb;
  const long ps = F->per_design, nv = F->nv;
  const double *E = F->buf + F->E_off;
  int rc = 0;
  // forward: t_l = P_l (b_l - E_l t_{l-1}); the last level's t is its x
  for (int l = 0; l < nl; l++) {
    double *zl = F->vb + F->voff[l];
    if (l > 0) {
      const int nn = F->m[l] / 6;
      hipLaunchKernelGGL(couple_fwd_kernel, dim3((6 * nn + 255) / 256, ne), dim3(256), 0, ctx->stream, nn,
                         F->d_cp_rowp + F->noff[l], F->d_cp_col, E, ps, F->vz + F->voff[l - 1], zl, nv, nb, act);
      HIPCALL(ctx, hipGetLastError());
    }
    if ((rc = symv(F, l, zl, 1.0, nullptr, 0.0, (l == nl - 1 ? F->vx : F->vz) + F->voff[l], nrhs))) return rc;
  }
  // backward: x_l = t_l - P_l E_{l+1}^T x_{l+1} (E^T x into vb)
  for (int l = nl - 2; l >= 0; l--) {
    const int nn = F->m[l] / 6;
    double *yl = F->vb + F->voff[l];
    hipLaunchKernelGGL(couple_bwd_kernel, dim3((6 * nn + 255) / 256, ne), dim3(256), 0, ctx->stream, nn,
                       F->d_ct_rowp + F->noff[l], F->d_ct_src, F->d_ct_blk, E, ps, F->vx + F->voff[l + 1], yl, nv, nb,
                       act);
    HIPCALL(ctx, hipGetLastError());
    if ((rc = symv(F, l, yl, -1.0, F->vz + F->voff[l], 1.0, F->vx + F->voff[l], nrhs))) return rc;
  }
  return 0;
}

// First round of cyclic reduction (the odd levels i, neighbours l = i - 1
// and r = i + 1), on the sparse couplings E_i (level i to l) and E_r (r to
// i) as they come from the mesh:
//   X_i = chol(D_i)^{-1},  P_i = X_i^T X_i = D_i^{-1},
//   D_l -= E_i^T P_i E_i,  D_r -= E_r P_i E_r^T,  B_r = -E_r P_i E_i,
// the products with E at the cost of its few 6 x 6 blocks per node (two
// dense-times-coupling products and a transpose per update, as in the
// sequential factorisation), where the dense form spends 7 M^3 flops per
// level on the couplings.  B_r, the coupling of r to l, is dense: the later
// rounds run on the even levels in the dense form.
static int factor_first_round(shellgpu_factor *F, int *info) {
  shellgpu_ctx *ctx = F->ctx;
  const FactorRound &r = F->rounds[0];
  const int nb = F->nbatch, M = F->M, cnt = nb * r.nE, cntR = nb * r.nR;
  const int maxnn = F->maxm / 6, ng = (maxnn + kCoupleNodes - 1) / kCoupleNodes;
  double **P = F->d_ptr + r.off;
  double **pD = P, **pZ = P + 2 * cnt;
  const CoupleJob *J = F->d_jobs;
  int rc = 0;
  if ((rc = cholinv(F, 0, info))) return rc;
  // P_i = X_i^T X_i in the Z_i slot (products over k >= max(row, column))
  if ((rc = mgemm(ctx, true, false, M, M, M, 1.0, pD, M, kGemmUpper, pD, M, kGemmLower, 0.0, pZ, M, 0, cnt)))
    return rc;
  const dim3 gc((M + 255) / 256, ng), gt((M + 31) / 32, (M + 31) / 32);
  auto couple = [&](bool left, long off, int njobs, double alpha, double beta) -> int {
    if (njobs <= 0) return 0;
    const dim3 g(gc.x, gc.y, njobs);
    if (left)
      hipLaunchKernelGGL(couple_job_kernel<true>, g, dim3(256), 0, ctx->stream, J + off, M, F->d_noff, F->d_ct_rowp,
                         F->d_ct_src, F->d_ct_blk, alpha, beta);
    else
      hipLaunchKernelGGL(couple_job_kernel<false>, g, dim3(256), 0, ctx->stream, J + off, M, F->d_noff,
                         F->d_cp_rowp, F->d_cp_col, F->d_cp_col, alpha, beta);
    HIPCALL(ctx, hipGetLastError());
    return 0;
  };
  auto transpose = [&](long off, int njobs) -> int {
    if (njobs <= 0) return 0;
    hipLaunchKernelGGL(transpose_job_kernel, dim3(gt.x, gt.y, njobs), dim3(256), 0, ctx->stream, J + off, M);
    HIPCALL(ctx, hipGetLastError());
    return 0;
  };
  // D_l -= (P_i E_i)^T E_i
  if ((rc = couple(true, F->job_q, cnt, 1.0, 0.0)) || (rc = transpose(F->job_t1, cnt)) ||
      (rc = couple(true, F->job_dl, cnt, -1.0, 1.0)))
    return rc;
  // V_i = P_i E_r^T;  D_r -= V_i^T E_r^T;  B_r = -V_i^T E_i
  if ((rc = couple(false, F->job_v, cntR, 1.0, 0.0)) || (rc = transpose(F->job_t2, cntR)) ||
      (rc = couple(false, F->job_dr, cntR, -1.0, 1.0)) || (rc = couple(true, F->job_br, cntR, -1.0, 0.0)))
    return rc;
  return 0;
}

// Block cyclic reduction of the level-block-tridiagonal matrix in F->buf
// (after the scatter).  The first round runs on the sparse couplings
// (factor_first_round); a later round with stride s eliminates levels
// i = s, 3s, ... with active neighbours l = i - s and r = i + s:
//   C_i C_i^T = D_i,  X_i = C_i^{-1},  W_i = B_r X_i^T,  Y_i = X_i B_i,
//   D_l -= Y_i^T Y_i,  D_r -= W_i W_i^T,  B_r <- -W_i Y_i,
// all entries of a round (levels x designs) in one batched call each, on
// the matrix cores (triangular X_i: half the products; symmetric updates:
// the lower tiles only); Y_i goes to the free slot B_{i-1}.  Level 0 is
// factored last.
static int factor_cr(shellgpu_factor *F) {
  shellgpu_ctx *ctx = F->ctx;
  const int nb = F->nbatch, nl = F->nlevels, M = F->M;
  if (*std::min_element(F->m.begin(), F->m.end()) < M) {
    hipLaunchKernelGGL(pad_identity_kernel, dim3((M + 255) / 256, nl, nb), dim3(256), 0, ctx->stream, M, (long)M * M,
                       F->d_mlev, F->buf, F->per_design);
    HIPCALL(ctx, hipGetLastError());
  }
  const long half = (long)nb * (nl + 1);
  HIPCALL(ctx, hipMemsetAsync(F->cr_info, 0, sizeof(int) * 2 * half, ctx->stream));
  long ioff = 0;
  int rc = 0;
  for (const FactorRound &r : F->rounds) {
    const int cnt = nb * r.nE, cntR = nb * r.nR;
    double **P = F->d_ptr + r.off;
    double **pD = P, **pB = P + cnt, **pZ = P + 2 * cnt, **pDl = P + 3 * cnt, **pDr = P + 4 * cnt,
           **pBr = P + 5 * cnt, **pY = P + 6 * cnt;
    if (r.s == 1) {
      if ((rc = factor_first_round(F, F->cr_info + ioff))) return rc;
      ioff += cnt;
      continue;
    }
    if ((rc = cholinv(F, &r - F->rounds.data(), F->cr_info + ioff)) ||
        // W_i = B_r X_i^T
        (rc = mgemm(ctx, false, true, M, M, M, 1.0, pBr, M, kGemmFull, pD, M, kGemmUpper, 0.0, pZ, M, 0, cntR)) ||
        // Y_i = X_i B_i
        (rc = mgemm(ctx, false, false, M, M, M, 1.0, pD, M, kGemmLower, pB, M, kGemmFull, 0.0, pY, M, 0, cnt)) ||
        // D_l -= Y_i^T Y_i
        (rc = mgemm(ctx, true, false, M, M, M, -1.0, pY, M, kGemmFull, pY, M, kGemmFull, 1.0, pDl, M, 1, cnt)) ||
        // D_r -= W_i W_i^T
        (rc = mgemm(ctx, false, true, M, M, M, -1.0, pZ, M, kGemmFull, pZ, M, kGemmFull, 1.0, pDr, M, 1, cntR)) ||
        // B_r = -W_i Y_i (coupling of r to l)
        (rc = mgemm(ctx, false, false, M, M, M, -1.0, pZ, M, kGemmFull, pY, M, kGemmFull, 0.0, pBr, M, 0, cntR)))
      return rc;
    ioff += cnt;
  }
  if ((rc = cholinv(F, F->rounds.size(), F->cr_info + ioff))) return rc;
  ioff += nb;
  std::vector<int> info(ioff);
  HIPCALL(ctx, hipMemcpyAsync(info.data(), F->cr_info, sizeof(int) * ioff, hipMemcpyDeviceToHost, ctx->stream));
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<int> bad(nb, 0);
  for (long q = 0; q < ioff; q++)
    if (info[q] != 0) bad[q % nb] = 1;
  HIPCALL(ctx, hipMemsetAsync(F->info, 0, sizeof(int) * nb, ctx->stream));
  HIPCALL(ctx, hipMemcpyAsync(F->info, bad.data(), sizeof(int) * nb, hipMemcpyHostToDevice, ctx->stream));
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  for (int b = 0; b < nb; b++)
    if (bad[b])
      return ferr(ctx, 5, "matrix not positive definite in the level Cholesky (design " + std::to_string(b) +
                              "): for K + sigma G the shift is above the first buckling load");
  F->factored = true;
  return 0;
}

// x = A^{-1} b by the cyclic-reduction factors: reduction sweep, root,
// back-substitution sweep (level-ordered vectors vb, vz, vx).  The first
// round's levels i hold P_i: forward t_i = P_i b_i (into x_i), b_k -= E_k
// t_{k-1} + E_{k+1}^T t_{k+1} on the even levels k; backward x_i = t_i -
// P_i (E_i x_{i-1} + E_{i+1}^T x_{i+1}).
static int apply_cr(shellgpu_factor *F) {
  shellgpu_ctx *ctx = F->ctx;
  const int nb = F->nbatch, nl = F->nlevels, M = F->M;
  const long M2 = (long)M * M, PD = F->per_design, NV = F->nv;
  double *D0 = F->buf, *B0 = F->buf + (long)nl * M2, *Z0 = F->buf + 2L * nl * M2;
  const double *E = F->buf + F->E_off;
  const int gx = (F->maxm + 255) / 256;
  auto gather = [&](int k0, const double *src, double *dst, double alpha, double beta) -> int {
    const int nk = (nl - k0 + 1) / 2;
    hipLaunchKernelGGL(cr_gather_kernel, dim3(gx, nk, nb), dim3(256), 0, ctx->stream, k0, nl, M, F->d_noff,
                       F->d_cp_rowp, F->d_cp_col, F->d_ct_rowp, F->d_ct_src, F->d_ct_blk, E, PD, src, dst, NV,
                       alpha, beta);
    HIPCALL(ctx, hipGetLastError());
    return 0;
  };
  int rc = 0;
  // t_i = P_i b_i (symmetric P_i: half of it read)
  if ((rc = symv_cr(F, F->vb + M, 1.0, nullptr, 0.0, F->vx + M))) return rc;
  if ((rc = gather(0, F->vx, F->vb, -1.0, 1.0))) return rc;
  for (const FactorRound &r : F->rounds) {
    const long s = r.s;
    if (s == 1) continue;
    const GemvBatch g{nb, PD, 2 * s * M2, NV, 2 * s * M, NV, 2 * s * M};
    const int cnt = nb * r.nE, cntR = nb * r.nR;
    // c_i = X_i b_i
    if ((rc = gemv_batch(F, false, M, M, D0 + s * M2, M, F->vb + s * M, 1.0, 0.0, F->vz + s * M, 1, g, cnt)))
      return rc;
    // b_l -= Y_i^T c_i (Y_i in the slot B_{i-1})
    if ((rc = gemv_batch(F, true, M, M, B0 + (s - 1) * M2, M, F->vz + s * M, -1.0, 1.0, F->vb, 0, g, cnt)))
      return rc;
    // b_r -= W_i c_i
    if ((rc = gemv_batch(F, false, M, M, Z0 + s * M2, M, F->vz + s * M, -1.0, 1.0, F->vb + 2 * s * M, 0, g, cntR)))
      return rc;
  }
  {
    const GemvBatch g{nb, PD, 0, NV, 0, NV, 0};
    if ((rc = gemv_batch(F, false, M, M, D0, M, F->vb, 1.0, 0.0, F->vz, 1, g, nb))) return rc;
    if ((rc = gemv_batch(F, true, M, M, D0, M, F->vz, 1.0, 0.0, F->vx, 1, g, nb))) return rc;
  }
  for (auto it = F->rounds.rbegin(); it != F->rounds.rend(); ++it) {
    const long s = it->s;
    if (s == 1) continue;
    const GemvBatch g{nb, PD, 2 * s * M2, NV, 2 * s * M, NV, 2 * s * M};
    const int cnt = nb * it->nE, cntR = nb * it->nR;
    // c_i -= Y_i x_l
    if ((rc = gemv_batch(F, false, M, M, B0 + (s - 1) * M2, M, F->vx, -1.0, 1.0, F->vz + s * M, 0, g, cnt)))
      return rc;
    // c_i -= W_i^T x_r
    if ((rc = gemv_batch(F, true, M, M, Z0 + s * M2, M, F->vx + 2 * s * M, -1.0, 1.0, F->vz + s * M, 0, g, cntR)))
      return rc;
    // x_i = X_i^T c_i
    if ((rc = gemv_batch(F, true, M, M, D0 + s * M2, M, F->vz + s * M, 1.0, 0.0, F->vx + s * M, 1, g, cnt)))
      return rc;
  }
  if ((rc = gather(1, F->vx, F->vb, 1.0, 0.0))) return rc;
  return symv_cr(F, F->vb + M, -1.0, F->vx + M, 1.0, F->vx + M);
}

// x = A^{-1} b for R vectors through a cyclic-reduction factor of one
// design, reading the factor once for all of them: the first round's
// symmetric inverses by the multi-vector symmetric product (read in full),
// the dense rounds' products as matrix products with R columns, the
// gathers over the vectors.  The level-ordered vectors lie at a stride of
// whole level pairs (NVR), so vector r of odd level 2k + 1 starts at
// r NVR + (2k + 1) M.  b, x: R vectors of n dofs.
static int apply_cr_multi(shellgpu_factor *F, const double *b, double *x, int R, bool may_alloc) {
  shellgpu_ctx *ctx = F->ctx;
  const int nl = F->nlevels, M = F->M, nbf = (nl + 1) / 2, nE1 = F->rounds[0].nE;
  const long M2 = (long)M * M, NV = F->nv, NVR = 2L * nbf * M, n = F->n;
  const int nT = (M + kSymvTile - 1) / kSymvTile;
  if (R > F->crm_cap) {
    if (!may_alloc) return ferr(ctx, 1, "the factor's vector buffers would have to grow");
    HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
    void **bufs[] = {(void **)&F->crm_b, (void **)&F->crm_z, (void **)&F->crm_x, (void **)&F->crm_part,
                     (void **)&F->d_crm_act};
    for (void **p : bufs) {
      if (*p) HIPCALL(ctx, hipFree(*p));
      *p = nullptr;
    }
    for (double **p : {&F->crm_b, &F->crm_z, &F->crm_x})
      HIPCALL(ctx, hipMalloc((void **)p, sizeof(double) * (size_t)NVR * R));
    HIPCALL(ctx, hipMalloc((void **)&F->crm_part, sizeof(double) * (size_t)R * nbf * nT * nT * kSymvTile));
    std::vector<int> act;
    for (int r = 0; r < R; r++)
      for (int k = 0; k < nE1; k++) act.push_back(r * nbf + k);
    HIPCALL(ctx, hipMalloc((void **)&F->d_crm_act, sizeof(int) * act.size()));
    HIPCALL(ctx, hipMemcpy(F->d_crm_act, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice));
    F->crm_cap = R;
  }
  double *vb = F->crm_b, *vz = F->crm_z, *vx = F->crm_x;
  double *D0 = F->buf, *B0 = F->buf + (long)nl * M2, *Z0 = F->buf + 2L * nl * M2;
  const double *E = F->buf + F->E_off;
  const dim3 pg((unsigned)((NV + 255) / 256), R);
  hipLaunchKernelGGL(permute_rhs_kernel, pg, dim3(256), 0, ctx->stream, NV, n, NVR, F->d_pos2dof, b, vb, 1);
  HIPCALL(ctx, hipGetLastError());
  // y = beta c + alpha P_i x over the odd levels and the R vectors
  auto symv1 = [&](const double *xin, double alpha, const double *c, double beta, double *y) -> int {
    const double *P = Z0 + M2;  // Z slot of level 1, odd levels 2 M2 apart
    switch (R) {
#define SYMV_CRM(RR)                                                                                               \
  case RR:                                                                                                         \
    hipLaunchKernelGGL((symv_multi_pass1_kernel<RR, 1>), dim3(nT * nT, nE1), dim3(256), 0, ctx->stream, M, nT, P,  \
                       M, 2 * M2, xin, 2L * M, nbf, F->crm_part);                                                  \
    break;
      SYMV_CRM(2)
      SYMV_CRM(3)
      SYMV_CRM(4)
      SYMV_CRM(5)
      SYMV_CRM(6)
      SYMV_CRM(7)
      SYMV_CRM(8)
      SYMV_CRM(16)
#undef SYMV_CRM
      default:
        return ferr(ctx, 1, "unsupported number of vectors for the cyclic-reduction solve");
    }
    HIPCALL(ctx, hipGetLastError());
    const SymvBatch sb{nbf, 0, 0, 0, 0, 2L * M, NVR};
    const dim3 g2((M + 63) / 64, R * nE1);
    if (nT >= 64)
      hipLaunchKernelGGL(symv_pass2_kernel<16>, g2, dim3(1024), 0, ctx->stream, M, nT, nT, F->crm_part, alpha, c,
                         beta, y, sb, F->d_crm_act);
    else if (nT >= 32)
      hipLaunchKernelGGL(symv_pass2_kernel<8>, g2, dim3(512), 0, ctx->stream, M, nT, nT, F->crm_part, alpha, c, beta,
                         y, sb, F->d_crm_act);
    else if (nT >= 16)
      hipLaunchKernelGGL(symv_pass2_kernel<4>, g2, dim3(256), 0, ctx->stream, M, nT, nT, F->crm_part, alpha, c, beta,
                         y, sb, F->d_crm_act);
    else if (nT >= 8)
      hipLaunchKernelGGL(symv_pass2_kernel<2>, g2, dim3(128), 0, ctx->stream, M, nT, nT, F->crm_part, alpha, c, beta,
                         y, sb, F->d_crm_act);
    else
      hipLaunchKernelGGL(symv_pass2_kernel<1>, g2, dim3(64), 0, ctx->stream, M, nT, nT, F->crm_part, alpha, c, beta,
                         y, sb, F->d_crm_act);
    HIPCALL(ctx, hipGetLastError());
    return 0;
  };
  const int gx = (F->maxm + 255) / 256;
  auto gather = [&](int k0, const double *src, double *dst, double alpha, double beta) -> int {
    const int nk = (nl - k0 + 1) / 2;
    hipLaunchKernelGGL(cr_gather_kernel, dim3(gx, nk, R), dim3(256), 0, ctx->stream, k0, nl, M, F->d_noff,
                       F->d_cp_rowp, F->d_cp_col, F->d_ct_rowp, F->d_ct_src, F->d_ct_blk, E, 0L, src, dst, NVR,
                       alpha, beta);
    HIPCALL(ctx, hipGetLastError());
    return 0;
  };
  // C = beta C + alpha op(A) X over cnt levels (matrices sA apart, vector
  // segments sv apart), X and C with R columns NVR apart; tri: triangle of
  // op(A)
  auto prod = [&](bool ta, const double *A, long sA, int tri, const double *X, double alpha, double beta, double *C,
                  long sv, int cnt) -> int {
    if (cnt <= 0) return 0;
    GemmArgs g{M, R, M, alpha, beta, A, sA, nullptr, M, X, sv, nullptr, (int)NVR, C, sv, nullptr, (int)NVR};
    g.tri_a = tri;
    // the symmetric products' partial sums are free between them
    g.ws = F->crm_part;
    g.ws_len = (long)R * nbf * nT * nT * kSymvTile;
    HIPCALL(ctx, launch_gemm_f64(ta, false, g, cnt, ctx->stream));
    return 0;
  };
  if (NVR > 0x7fffffffL) return ferr(ctx, 1, "vectors too long for the cyclic-reduction multi-vector solve");
  int rc = 0;
  // t_i = P_i b_i
  if ((rc = symv1(vb + M, 1.0, nullptr, 0.0, vx + M))) return rc;
  if ((rc = gather(0, vx, vb, -1.0, 1.0))) return rc;
  for (const FactorRound &r : F->rounds) {
    const long s = r.s;
    if (s == 1) continue;
    // c_i = X_i b_i; b_l -= Y_i^T c_i (Y_i in the slot B_{i-1}); b_r -= W_i c_i
    if ((rc = prod(false, D0 + s * M2, 2 * s * M2, kGemmLower, vb + s * M, 1.0, 0.0, vz + s * M, 2 * s * M, r.nE)))
      return rc;
    if ((rc = prod(true, B0 + (s - 1) * M2, 2 * s * M2, kGemmFull, vz + s * M, -1.0, 1.0, vb, 2 * s * M, r.nE)))
      return rc;
    if ((rc = prod(false, Z0 + s * M2, 2 * s * M2, kGemmFull, vz + s * M, -1.0, 1.0, vb + 2 * s * M, 2 * s * M,
                   r.nR)))
      return rc;
  }
  if ((rc = prod(false, D0, 0, kGemmLower, vb, 1.0, 0.0, vz, 0, 1))) return rc;
  if ((rc = prod(true, D0, 0, kGemmUpper, vz, 1.0, 0.0, vx, 0, 1))) return rc;
  for (auto it = F->rounds.rbegin(); it != F->rounds.rend(); ++it) {
    const long s = it->s;
    if (s == 1) continue;
    // c_i -= Y_i x_l; c_i -= W_i^T x_r; x_i = X_i^T c_i
    if ((rc = prod(false, B0 + (s - 1) * M2, 2 * s * M2, kGemmFull, vx, -1.0, 1.0, vz + s * M, 2 * s * M, it->nE)))
      return rc;
    if ((rc = prod(true, Z0 + s * M2, 2 * s * M2, kGemmFull, vx + 2 * s * M, -1.0, 1.0, vz + s * M, 2 * s * M,
                   it->nR)))
      return rc;
    if ((rc = prod(true, D0 + s * M2, 2 * s * M2, kGemmUpper, vz + s * M, 1.0, 0.0, vx + s * M, 2 * s * M, it->nE)))
      return rc;
  }
  if ((rc = gather(1, vx, vb, 1.0, 0.0))) return rc;
  if ((rc = symv1(vb + M, -1.0, vx + M, 1.0, vx + M))) return rc;
  hipLaunchKernelGGL(permute_rhs_kernel, pg, dim3(256), 0, ctx->stream, NV, n, NVR, F->d_pos2dof, vx, x, 0);
  HIPCALL(ctx, hipGetLastError());
  return 0;
}

}  // extern "C"

namespace {

// x = A^{-1} b for nrhs vectors per factored design (entry e = r nbatch + d
// of the factor at offset e n of b and x).  The sequential factorisation
// solves them together, reading each level inverse once per solve; cyclic
// reduction solves them one after the other.
// The single-vector solve body (level-ordered vb -> vx) launches kernels
// only, on buffers fixed for the factor's life (until a multi-vector solve
// regrows them), with the factor values read at run time.  With
// SHELLGPU_GRAPH=1 it is captured into a HIP graph at its first use and
// replayed.  It is opt-in: the solves are bound by the GPU, not by their
// launches (no wall-time change measured), and rocprofv3's kernel tracing
// crashes in the capture.
int solve_body(shellgpu_factor *F) { return F->cr ? apply_cr(F) : apply_schur(F, 1); }

int run_solve_body(shellgpu_factor *F) {
  shellgpu_ctx *ctx = F->ctx;
  if (!F->solve_graph && !F->graph_off) {
    const char *env = getenv("SHELLGPU_GRAPH");
    if (!env || strcmp(env, "1") != 0) {
      F->graph_off = true;
    } else {
      hipGraph_t g = nullptr;
      HIPCALL(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      const int rc = solve_body(F);
      const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
      if (rc != 0 || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : ferr(ctx, 2, std::string("solve capture: ") + hipGetErrorString(e));
      }
      const hipError_t ei = hipGraphInstantiate(&F->solve_graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) {
        F->solve_graph = nullptr;
        return ferr(ctx, 2, std::string("solve graph: ") + hipGetErrorString(ei));
      }
    }
  }
  if (F->graph_off) return solve_body(F);
  HIPCALL(ctx, hipGraphLaunch(F->solve_graph, ctx->stream));
  return 0;
}

// may_alloc = false: no capture of the solve graph and no growth of the
// vector buffers (an error instead)
int apply_block(shellgpu_factor *F, const double *b, double *x, int nrhs, bool may_alloc = true) {
  shellgpu_ctx *ctx = F->ctx;
  const int nb = F->nbatch;
  const long nv = F->nv, n = F->n;
  // a Lanczos run may restrict single-vector solves of the sequential form
  // to the designs it has not converged yet (the others keep x unchanged)
  const int *act = (!F->cr && nrhs == 1 && F->nact >= 0) ? F->d_act : nullptr;
  if (act && F->nact == 0) return 0;
  const dim3 pgrid((unsigned)((nv + 255) / 256), act ? F->nact : nb);
  int rc = 0;
  if (F->cr && nb == 1 && nrhs > 1 && (nrhs <= 8 || nrhs == kMaxRhs)) return apply_cr_multi(F, b, x, nrhs, may_alloc);
  if (F->cr || nrhs == 1) {
    for (int r = 0; r < nrhs; r++) {
      const double *br = b + (size_t)r * nb * n;
      double *xr = x + (size_t)r * nb * n;
      hipLaunchKernelGGL(permute_kernel, pgrid, dim3(256), 0, ctx->stream, nv, n, F->d_pos2dof, br, F->vb, 1, act);
      HIPCALL(ctx, hipGetLastError());
      // the captured body covers every design
      const bool direct = act || (!may_alloc && !F->solve_graph);
      if ((rc = direct ? solve_body(F) : run_solve_body(F))) return rc;
      hipLaunchKernelGGL(permute_kernel, pgrid, dim3(256), 0, ctx->stream, nv, n, F->d_pos2dof, F->vx, xr, 0, act);
      HIPCALL(ctx, hipGetLastError());
    }
    return 0;
  }
  if (nrhs > kMaxRhs || (nrhs > 8 && nrhs != kMaxRhs)) return ferr(ctx, 1, "unsupported number of right-hand sides");
  if (nrhs > F->rhs_cap) {
    if (!may_alloc) return ferr(ctx, 1, "the factor's vector buffers would have to grow");
    HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
    if (F->solve_graph) {  // captured on the buffers about to move
      HIPCALL(ctx, hipGraphExecDestroy(F->solve_graph));
      F->solve_graph = nullptr;
    }
    double **bufs[] = {&F->vb, &F->vz, &F->vx};
    for (double **p : bufs) {
      HIPCALL(ctx, hipFree(*p));
      *p = nullptr;
      HIPCALL(ctx, hipMalloc((void **)p, sizeof(double) * (size_t)nv * nb * nrhs));
    }
    const long nT = (F->maxm + kSymvTile - 1) / kSymvTile;
    HIPCALL(ctx, hipFree(F->part));
    F->part = nullptr;
    HIPCALL(ctx,
            hipMalloc((void **)&F->part, sizeof(double) * (size_t)nT * nT * kMaxSplit * kSymvTile * nb * nrhs));
    F->rhs_cap = nrhs;
  }
  const dim3 mgrid((unsigned)((nv + 255) / 256), nrhs * nb);
  hipLaunchKernelGGL(permute_kernel, mgrid, dim3(256), 0, ctx->stream, nv, n, F->d_pos2dof, b, F->vb, 1, nullptr);
  HIPCALL(ctx, hipGetLastError());
  if ((rc = apply_schur(F, nrhs))) return rc;
  hipLaunchKernelGGL(permute_kernel, mgrid, dim3(256), 0, ctx->stream, nv, n, F->d_pos2dof, F->vx, x, 0, nullptr);
  HIPCALL(ctx, hipGetLastError());
  return 0;
}

// block Lanczos: nrhs vectors of every design
int factor_apply_multi(shellgpu_factor *F, const double *b, double *x, int nrhs) {
  if (!F->factored) return ferr(F->ctx, 1, "factor not computed");
  return apply_block(F, b, x, nrhs);
}

// shellgpu_factor_apply without device allocations (see apply_block)
int factor_apply_noalloc(shellgpu_factor *F, const double *b, double *x) {
  shellgpu_ctx *ctx = F->ctx;
  if (!F->factored) return ferr(ctx, 1, "factor not computed");
  if (F->nbatch == ctx->nbatch) return apply_block(F, b, x, 1, false);
  if (F->nbatch != 1) return ferr(ctx, 1, "factor batch size does not match the context's designs");
  const long n = F->n;
  for (int r0 = 0; r0 < ctx->nbatch;) {
    const int nr = rhs_chunk(ctx->nbatch - r0);
    const int rc = apply_block(F, b + (size_t)r0 * n, x + (size_t)r0 * n, nr, false);
    if (rc) return rc;
    r0 += nr;
  }
  return 0;
}

// 0 when factor_apply_noalloc can run on F without allocating
int factor_noalloc_ready(shellgpu_factor *F) {
  shellgpu_ctx *ctx = F->ctx;
  if (!F->factored) return ferr(ctx, 1, "factor not computed");
  if (F->nbatch == ctx->nbatch || F->cr) return 0;
  if (F->nbatch != 1) return ferr(ctx, 1, "factor batch size does not match the context's designs");
  if (rhs_chunk(ctx->nbatch) > F->rhs_cap)
    return ferr(ctx, 1, "a factor shared by " + std::to_string(ctx->nbatch) +
                            " designs needs larger vector buffers: apply it once with shellgpu_factor_apply first");
  return 0;
}

}  // namespace

extern "C" {

// x = A^{-1} b for every design of the context.  A factor of one design
// shared by all (identical designs: load cases of one structure) solves the
// designs' vectors as right-hand sides of that design, 16 (or at most 8)
// at a time.
int shellgpu_factor_apply(shellgpu_factor *F, const double *b, double *x) {
  shellgpu_ctx *ctx = F->ctx;
  if (!F->factored) return ferr(ctx, 1, "factor not computed");
  if (F->nbatch == ctx->nbatch) return apply_block(F, b, x, 1);
  if (F->nbatch != 1) return ferr(ctx, 1, "factor batch size does not match the context's designs");
  const long n = F->n;
  for (int r0 = 0; r0 < ctx->nbatch;) {
    const int nr = rhs_chunk(ctx->nbatch - r0);
    const int rc = apply_block(F, b + (size_t)r0 * n, x + (size_t)r0 * n, nr);
    if (rc) return rc;
    r0 += nr;
  }
  return 0;
}


}  // extern "C"

// ---------------------------------------------------------------------------
// batched vector helpers for GMRES / Lanczos (device vectors, n per design)
// ---------------------------------------------------------------------------
namespace {

// device doubles VecOps needs for nb designs: the dot results, per-design
// coefficients and the partial sums of the two-pass dot
inline long vecops_doubles(int nb) { return 2L * (nb + 1) + (long)kDotChunks * nb; }

struct VecOps {
  shellgpu_ctx *ctx;
  int n, nb;
  double *dwork = nullptr;  // vecops_doubles(nb)
  bool owned = false;
  // ws: caller-provided workspace of vecops_doubles(nb) doubles, or null
  // to allocate one
  int init(shellgpu_ctx *c, int n_, int nb_, double *ws = nullptr) {
    ctx = c;
    n = n_;
    nb = nb_;
    if (ws) {
      dwork = ws;
      return 0;
    }
    HIPCALL(ctx, hipMalloc((void **)&dwork, sizeof(double) * vecops_doubles(nb)));
    owned = true;
    return 0;
  }
  ~VecOps() {
    if (owned && dwork) (void)hipFree(dwork);
  }
  // out[b] = x[b] . y[b]
  int dot(const double *x, const double *y, std::vector<double> &out) {
    const int blocks = std::min(kDotChunks, (n + kDotThreads - 1) / kDotThreads);
    double *part = dwork + 2L * (nb + 1);
    hipLaunchKernelGGL(dot_batched_kernel, dim3(blocks, nb), dim3(kDotThreads), 0, ctx->stream, n, x, (long)n, y,
                       (long)n, part);
    HIPCALL(ctx, hipGetLastError());
    hipLaunchKernelGGL(dot_finish_kernel, dim3(nb), dim3(kDotThreads), 0, ctx->stream, blocks, part, dwork);
    HIPCALL(ctx, hipGetLastError());
    out.resize(nb);
    HIPCALL(ctx, hipMemcpyAsync(out.data(), dwork, sizeof(double) * nb, hipMemcpyDeviceToHost, ctx->stream));
    HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
  }
  int scale(double *x, const std::vector<double> &s) {
    HIPCALL(ctx, hipMemcpyAsync(dwork + nb, s.data(), sizeof(double) * nb, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(scale_batched_kernel, dim3((n + 255) / 256, nb), dim3(256), 0, ctx->stream, n, x, dwork + nb);
    HIPCALL(ctx, hipGetLastError());
    return 0;
  }
  int axpy(const std::vector<double> &a, const double *x, double *y) {
    HIPCALL(ctx, hipMemcpyAsync(dwork + nb, a.data(), sizeof(double) * nb, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(axpy_batched_kernel, dim3((n + 255) / 256, nb), dim3(256), 0, ctx->stream, n, dwork + nb, x,
                       y);
    HIPCALL(ctx, hipGetLastError());
    return 0;
  }
};

}  // namespace

namespace {

// GMRES workspace of nb designs of n dofs: msub + 1 Krylov vectors and one
// work vector (each rounded up to 32 doubles), then the VecOps buffer
inline long gmres_vec_stride(int n, int nb) { return ((long)n * nb + 31) / 32 * 32; }
inline long gmres_doubles(int n, int nb, int msub) {
  return (long)(msub + 2) * gmres_vec_stride(n, nb) + vecops_doubles(nb);
}

// Right-preconditioned restarted GMRES (GMRES::solve, src/bpmat/KSM.cpp:
// 785-955) on the workspace ws (gmres_doubles).  may_alloc = false: the
// preconditioner must not allocate (no capture of its solve graph, no
// growth of a shared factor's vector buffers; the caller checks the
// latter).
int gmres_run(shellgpu_ctx *ctx, shellgpu_mat *A, shellgpu_factor *F, const double *b, double *x, int msub,
              int nrestart, double rtol, double atol, int *iters, double *resnorm, double *ws, bool may_alloc) {
  const int nb = ctx->nbatch, n = 6 * ctx->nnodes;
  const long vs = gmres_vec_stride(n, nb);
  VecOps vo;
  int rc = vo.init(ctx, n, nb, ws + (msub + 2) * vs);
  if (rc) return rc;
  std::vector<double *> W(msub + 1);
  for (int i = 0; i <= msub; i++) W[i] = ws + i * vs;
  double *work = ws + (msub + 1) * vs;
  auto precond = [&](const double *in, double *out) {
    return may_alloc ? shellgpu_factor_apply(F, in, out) : factor_apply_noalloc(F, in, out);
  };
  // per design Hessenberg (column major, (msub+1) x msub), Givens, residuals
  std::vector<std::vector<double>> H(nb, std::vector<double>((size_t)(msub + 1) * msub, 0.0));
  std::vector<std::vector<double>> cs(nb, std::vector<double>(msub)), sn(nb, std::vector<double>(msub)),
      g(nb, std::vector<double>(msub + 1));
  std::vector<double> rhs_norm(nb, 0.0), cur(nb, 0.0), tmp(nb);
  std::vector<int> done(nb, 0);
  int total_iters = 0;
  HIPCALL(ctx, hipMemsetAsync(x, 0, sizeof(double) * (size_t)n * nb, ctx->stream));
  for (int count = 0; count <= nrestart; count++) {
    // W0 = b - A x
    if ((rc = shellgpu_mat_mult(A, x, W[0]))) return rc;
    {
      std::vector<double> m1(nb, -1.0);
      if ((rc = vo.scale(W[0], m1))) return rc;
    }
    {
      std::vector<double> p1(nb, 1.0);
      if ((rc = vo.axpy(p1, b, W[0]))) return rc;
    }
    if ((rc = vo.dot(W[0], W[0], tmp))) return rc;
    for (int d = 0; d < nb; d++) {
      cur[d] = std::sqrt(tmp[d]);
      if (count == 0) rhs_norm[d] = cur[d];
      g[d].assign(msub + 1, 0.0);
      g[d][0] = cur[d];
      if (cur[d] < atol || cur[d] <= rtol * rhs_norm[d]) done[d] = 1;
      tmp[d] = (cur[d] > 0.0) ? 1.0 / cur[d] : 0.0;
    }
    if (std::all_of(done.begin(), done.end(), [](int v) { return v; })) break;
    if ((rc = vo.scale(W[0], tmp))) return rc;
    int niters = 0;
    for (int i = 0; i < msub; i++) {
      if (F) {
        if ((rc = precond(W[i], work))) return rc;
        if ((rc = shellgpu_mat_mult(A, work, W[i + 1]))) return rc;
      } else {
        if ((rc = shellgpu_mat_mult(A, W[i], W[i + 1]))) return rc;
      }
      // modified Gram-Schmidt (GMRES::orthogonalize)
      for (int j = 0; j <= i; j++) {
        if ((rc = vo.dot(W[i + 1], W[j], tmp))) return rc;
        for (int d = 0; d < nb; d++) H[d][(size_t)i * (msub + 1) + j] = tmp[d];
        std::vector<double> m(nb);
        for (int d = 0; d < nb; d++) m[d] = -tmp[d];
        if ((rc = vo.axpy(m, W[j], W[i + 1]))) return rc;
      }
      if ((rc = vo.dot(W[i + 1], W[i + 1], tmp))) return rc;
      std::vector<double> inv(nb);
      for (int d = 0; d < nb; d++) {
        const double hn = std::sqrt(tmp[d]);
        H[d][(size_t)i * (msub + 1) + i + 1] = hn;
        inv[d] = hn > 0.0 ? 1.0 / hn : 0.0;
      }
      if ((rc = vo.scale(W[i + 1], inv))) return rc;
      bool all = true;
      for (int d = 0; d < nb; d++) {
        double *Hc = &H[d][(size_t)i * (msub + 1)];
        for (int k = 0; k < i; k++) {
          const double h1 = Hc[k], h2 = Hc[k + 1];
          Hc[k] = h1 * cs[d][k] + h2 * sn[d][k];
          Hc[k + 1] = -h1 * sn[d][k] + h2 * cs[d][k];
        }
        const double h1 = Hc[i], h2 = Hc[i + 1];
        const double sq = std::sqrt(h1 * h1 + h2 * h2);
        cs[d][i] = sq > 0.0 ? h1 / sq : 1.0;
        sn[d][i] = sq > 0.0 ? h2 / sq : 0.0;
        Hc[i] = h1 * cs[d][i] + h2 * sn[d][i];
        Hc[i + 1] = 0.0;
        const double r0 = g[d][i];
        g[d][i] = r0 * cs[d][i];
        g[d][i + 1] = -r0 * sn[d][i];
        cur[d] = std::fabs(g[d][i + 1]);
        if (!(cur[d] < atol || cur[d] < rtol * rhs_norm[d])) all = false;
      }
      niters++;
      if (all) break;
    }
    total_iters += niters;
    // solve the triangular system, update x (right preconditioning)
    {
      std::vector<std::vector<double>> y(nb, std::vector<double>(niters, 0.0));
      for (int d = 0; d < nb; d++) {
        for (int i = niters - 1; i >= 0; i--) {
          double s = g[d][i];
          for (int j = i + 1; j < niters; j++) s -= H[d][(size_t)j * (msub + 1) + i] * y[d][j];
          y[d][i] = s / H[d][(size_t)i * (msub + 1) + i];
          if (done[d]) y[d][i] = 0.0;
        }
      }
      HIPCALL(ctx, hipMemsetAsync(work, 0, sizeof(double) * (size_t)n * nb, ctx->stream));
      for (int i = 0; i < niters; i++) {
        std::vector<double> c(nb);
        for (int d = 0; d < nb; d++) c[d] = y[d][i];
        if ((rc = vo.axpy(c, W[i], work))) return rc;
      }
      std::vector<double> p1(nb, 1.0);
      if (F) {
        if ((rc = precond(work, W[0]))) return rc;
        if ((rc = vo.axpy(p1, W[0], x))) return rc;
      } else {
        if ((rc = vo.axpy(p1, work, x))) return rc;
      }
    }
    bool all = true;
    for (int d = 0; d < nb; d++)
      if (!(cur[d] < atol || cur[d] < rtol * rhs_norm[d])) all = false;
    if (all) break;
  }
  if (iters) *iters = total_iters;
  if (resnorm) {
    double worst = 0.0;
    for (int d = 0; d < nb; d++) worst = std::max(worst, rhs_norm[d] > 0 ? cur[d] / rhs_norm[d] : 0.0);
    *resnorm = worst;
  }
  return 0;
}

int gmres_check(shellgpu_ctx *ctx, shellgpu_mat *A, shellgpu_factor *F, const double *b, double *x, int msub) {
  if (!ctx) return 1;
  if (!A || !b || !x) return ferr(ctx, 1, "A, b and x are required");
  if (A->nbatch != ctx->nbatch) return ferr(ctx, 1, "matrix batch size");
  if (msub < 1) return ferr(ctx, 1, "msub must be at least 1");
  if (F && F->ctx != ctx) return ferr(ctx, 1, "factor of another context");
  return 0;
}

}  // namespace

extern "C" {

int shellgpu_gmres(shellgpu_ctx *ctx, shellgpu_mat *A, shellgpu_factor *F, const double *b, double *x, int msub,
                   int nrestart, double rtol, double atol, int *iters, double *resnorm) {
  int rc = gmres_check(ctx, A, F, b, x, std::max(1, msub));
  if (rc) return rc;
  msub = std::max(1, msub);
  double *ws = nullptr;
  HIPCALL(ctx, hipMalloc((void **)&ws, sizeof(double) * gmres_doubles(6 * ctx->nnodes, ctx->nbatch, msub)));
  rc = gmres_run(ctx, A, F, b, x, msub, nrestart, rtol, atol, iters, resnorm, ws, true);
  // the stream may still read ws
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(ws);
  return rc;
}

int shellgpu_gmres_noalloc(shellgpu_ctx *ctx, shellgpu_mat *A, shellgpu_factor *F, const double *b, double *x,
                           int msub, int nrestart, double rtol, double atol, int *iters, double *resnorm) {
  int rc = gmres_check(ctx, A, F, b, x, msub);
  if (rc) return rc;
  if (!ctx->scratch) return ferr(ctx, 2, "the device scratch array is not mapped on this device");
  const long need = gmres_doubles(6 * ctx->nnodes, ctx->nbatch, msub);
  if (need > (long)ctx->scratch_doubles)
    return ferr(ctx, 1, "GMRES workspace of " + std::to_string(need * 8) + " bytes exceeds the " +
                            std::to_string(ctx->scratch_doubles * 8) +
                            "-byte device scratch array (fewer designs, a smaller msub or shellgpu_gmres)");
  if (F && (rc = factor_noalloc_ready(F))) return rc;
  return gmres_run(ctx, A, F, b, x, msub, nrestart, rtol, atol, iters, resnorm, ctx->scratch, false);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// linear buckling
// ---------------------------------------------------------------------------
namespace {

// A[b] = K[b] + s[b] G[b] over the batch
__global__ void shifted_kernel(long n, const double *K, const double *G, const double *s, double *A) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long o = (long)blockIdx.y * n + i;
  A[o] = K[o] + s[blockIdx.y] * G[o];
}

struct LanczosOut {
  int bs = 1;     // block size
  int steps = 0;  // basis vectors
  // block size 1: basis vectors per design (a design stops at its own
  // convergence; empty: steps for every design)
  std::vector<int> ns;
  // block size 1: the tridiagonal projection per design
  std::vector<std::vector<double>> alpha, beta;
  // block size > 1: per design the projected matrix (steps x steps,
  // column-major) and the last residual block (bs x bs, column-major)
  std::vector<std::vector<double>> T, R;
};

// Lanczos in the K inner product, designs in lock-step, full CGS2
// re-orthogonalisation.  mode 0: w = A^{-1} K q (shift-invert, A = K +
// sigma G factored in F); mode 1: w = K^{-1} (-G q) (buckling mode, F
// factors K); mode 2: w = A^{-1} K q with the argument K holding the mass
// matrix M and A = K - sigma M (generalised shift-invert, M inner product).
// Ritz value theta relates to the eigenvalue by lambda = -theta sigma /
// (1 - theta) (mode 0), lambda = 1 / theta (mode 1) or lambda = 1 / theta +
// sigma (mode 2).  Mode 3 is mode 1 with the wanted values the nev smallest
// |lambda| of either sign (mode 1 wants the smallest positive loads).  Stops
// when the nev wanted Ritz values have relative residual <= tol.  With a
// factor of the sequential form per design, a design whose wanted values
// have converged keeps its basis size (lo.ns) and its later solves are
// skipped; its later basis vectors and projection entries are not used.
// Otherwise (one factor, or cyclic reduction) every design keeps the whole
// basis.
int lanczos(shellgpu_ctx *ctx, shellgpu_factor *F, shellgpu_mat *K, shellgpu_mat *G, int mode,
            const std::vector<double> &sigma, int nev, int m, double tol, double *Q, double *KQ, double *hv,
            double *work, LanczosOut &lo) {
  const int nb = ctx->nbatch, n = 6 * ctx->nnodes;
  const bool skip = nb > 1 && !F->cr && F->nbatch == nb;
  struct ActiveReset {
    shellgpu_factor *F;
    ~ActiveReset() { F->nact = -1; }
  } active_reset{F};
  std::vector<char> done(nb, 0);
  std::vector<int> act;
  lo.ns.assign(nb, 0);
  rocblas_handle h = (rocblas_handle)ctx->blas;
  const double one = 1.0, mone = -1.0, zero = 0.0;
  VecOps vo;
  int rc = vo.init(ctx, n, nb);
  if (rc) return rc;
  std::vector<double> tmp(nb);
  lo.alpha.assign(nb, {});
  lo.beta.assign(nb, {});
  {
    std::vector<double> q0((size_t)n * nb);
    unsigned long long s = 0x5eedULL;
    for (auto &v : q0) {
      s += 0x9E3779B97F4A7C15ULL;
      unsigned long long z = s;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      z ^= z >> 31;
      v = 2.0 * ((double)(z >> 11) / 9007199254740992.0) - 1.0;
    }
    HIPCALL(ctx, hipMemcpyAsync(Q, q0.data(), sizeof(double) * q0.size(), hipMemcpyHostToDevice, ctx->stream));
    if ((rc = shellgpu_vec_zero_bcs(ctx, Q))) return rc;
    if ((rc = shellgpu_mat_mult(K, Q, KQ))) return rc;
    if ((rc = vo.dot(Q, KQ, tmp))) return rc;
    for (auto &t : tmp) t = 1.0 / std::sqrt(t);
    if ((rc = vo.scale(Q, tmp))) return rc;
    if ((rc = vo.scale(KQ, tmp))) return rc;
  }
  int k;
  for (k = 0; k < m; k++) {
    double *qn = Q + (size_t)(k + 1) * nb * n, *Kqn = KQ + (size_t)(k + 1) * nb * n;
    if (mode == 0 || mode == 2) {
      if ((rc = shellgpu_factor_apply(F, KQ + (size_t)k * nb * n, qn))) return rc;
    } else {
      if ((rc = shellgpu_mat_mult(G, Q + (size_t)k * nb * n, work))) return rc;
      if ((rc = shellgpu_vec_zero_bcs(ctx, work))) return rc;
      if ((rc = shellgpu_factor_apply(F, work, qn))) return rc;
      std::vector<double> m1(nb, -1.0);
      if ((rc = vo.scale(qn, m1))) return rc;
    }
    if ((rc = shellgpu_vec_zero_bcs(ctx, qn))) return rc;
    std::vector<double> hsum((size_t)nb * (k + 1), 0.0), hh((size_t)nb * (k + 1));
    for (int pass = 0; pass < 2; pass++) {
      BLASCALL(ctx, rocblas_dgemv_strided_batched(h, rocblas_operation_transpose, n, k + 1, &one, KQ, nb * n, n, qn,
                                                  1, n, &zero, hv, 1, k + 1, nb));
      BLASCALL(ctx, rocblas_dgemv_strided_batched(h, rocblas_operation_none, n, k + 1, &mone, Q, nb * n, n, hv, 1,
                                                  k + 1, &one, qn, 1, n, nb));
      HIPCALL(ctx, hipMemcpyAsync(hh.data(), hv, sizeof(double) * hh.size(), hipMemcpyDeviceToHost, ctx->stream));
      HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
      for (size_t i = 0; i < hh.size(); i++) hsum[i] += hh[i];
    }
    for (int d = 0; d < nb; d++) lo.alpha[d].push_back(hsum[(size_t)d * (k + 1) + k]);
    if ((rc = shellgpu_mat_mult(K, qn, Kqn))) return rc;
    if ((rc = vo.dot(qn, Kqn, tmp))) return rc;
    for (int d = 0; d < nb; d++) {
      const double bt = std::sqrt(std::max(tmp[d], 0.0));
      lo.beta[d].push_back(bt);
      tmp[d] = bt > 0.0 ? 1.0 / bt : 0.0;
    }
    if ((rc = vo.scale(qn, tmp))) return rc;
    if ((rc = vo.scale(Kqn, tmp))) return rc;
    // convergence checks: every step for one design (its last-row QL is
    // cheap next to a solve), every 5 steps for a batch
    if (k + 1 >= nev && ((k + 1) % (nb == 1 ? 1 : 5) == 0 || k + 1 == m)) {
      bool all = true, newly = false;
      for (int d = 0; d < nb; d++) {
        if (done[d]) continue;
        std::vector<double> w, z;
        const int ns = k + 1;
        std::vector<double> e(lo.beta[d].begin(), lo.beta[d].begin() + (ns - 1));
        std::vector<double> a(lo.alpha[d].begin(), lo.alpha[d].begin() + ns);
        // z: the last component of each Ritz vector
        if (tridiag_ql(ns, a, e, w, z, true) != 0) return ferr(ctx, 6, "tridiagonal eigensolver failed");
        std::vector<int> idx(ns);
        for (int i = 0; i < ns; i++) idx[i] = i;
        auto key = [&](int i) {
          if (mode == 2) return std::fabs(1.0 / w[i] + sigma[d]);
          if (mode == 3) return w[i] != 0.0 ? std::fabs(1.0 / w[i]) : 1e300;
          return mode == 0 ? std::fabs(-w[i] * sigma[d] / (1.0 - w[i])) : (w[i] > 0 ? 1.0 / w[i] : 1e300);
        };
        std::sort(idx.begin(), idx.end(), [&](int a, int b) { return key(a) < key(b); });
        bool conv = true;
        for (int j = 0; j < nev; j++) {
          const int i = idx[j];
          const double err = std::fabs(lo.beta[d][ns - 1] * z[i]);
          if (!(err <= tol * std::fabs(w[i]))) conv = false;
        }
        if (!conv) {
          all = false;
          if (!skip) break;  // every design is checked again at the next check
        } else if (skip) {
          done[d] = 1;  // frozen: its wanted pairs are final
          lo.ns[d] = ns;
          newly = true;
        }
      }
      if (all) {
        k++;
        break;
      }
      if (skip && newly) {
        act.clear();
        for (int d = 0; d < nb; d++)
          if (!done[d]) act.push_back(d);
        HIPCALL(ctx, hipMemcpyAsync(F->d_act, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice,
                                    ctx->stream));
        HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
        F->nact = (int)act.size();
      }
    }
  }
  lo.steps = std::min(k, m);
  for (int d = 0; d < nb; d++)
    if (!done[d]) lo.ns[d] = lo.steps;
  return 0;
}

// Ritz pairs of every design's projection after a Lanczos run: w[d] the
// Ritz values, z[d] the coefficient vectors (column i of an s x s
// column-major matrix) and res[d][i] the residual norm of pair i (in the
// inner product of the basis).  Block size 1: implicit QL of the
// tridiagonal matrix on the host; larger blocks: the dense projection by a
// batched symmetric eigensolver on the device.
int proj_eigs(shellgpu_ctx *ctx, const LanczosOut &lo, int s, std::vector<std::vector<double>> &w,
              std::vector<std::vector<double>> &z, std::vector<std::vector<double>> &res) {
  const int nb = ctx->nbatch, bs = lo.bs;
  w.assign(nb, {});
  z.assign(nb, {});
  res.assign(nb, {});
  if (bs == 1) {
    // a design's own basis size when it stopped before the others
    for (int d = 0; d < nb; d++) {
      const int sd = lo.ns.empty() ? s : std::min(s, lo.ns[d]);
      std::vector<double> e(lo.beta[d].begin(), lo.beta[d].begin() + (sd - 1));
      std::vector<double> a(lo.alpha[d].begin(), lo.alpha[d].begin() + sd);
      if (tridiag_ql(sd, a, e, w[d], z[d]) != 0) return ferr(ctx, 6, "tridiagonal eigensolver failed");
      res[d].resize(sd);
      for (int i = 0; i < sd; i++) res[d][i] = std::fabs(lo.beta[d][sd - 1] * z[d][(size_t)i * sd + sd - 1]);
    }
    return 0;
  }
  const size_t s2 = (size_t)s * s;
  std::vector<double> hT(s2 * nb);
  for (int d = 0; d < nb; d++)
    for (int j = 0; j < s; j++)
      for (int i = 0; i < s; i++) hT[d * s2 + (size_t)j * s + i] = lo.T[d][(size_t)j * lo.steps + i];
  double *dT = nullptr, *dD = nullptr, *dE = nullptr;
  int *dinfo = nullptr;
  auto release = [&]() {
    void *b[] = {dT, dD, dE, dinfo};
    for (void *p : b)
      if (p) (void)hipFree(p);
  };
  if (hipMalloc((void **)&dT, sizeof(double) * s2 * nb) != hipSuccess ||
      hipMalloc((void **)&dD, sizeof(double) * (size_t)s * nb) != hipSuccess ||
      hipMalloc((void **)&dE, sizeof(double) * (size_t)s * nb) != hipSuccess ||
      hipMalloc((void **)&dinfo, sizeof(int) * nb) != hipSuccess) {
    release();
    return ferr(ctx, 2, "projected eigenproblem allocation failed");
  }
  std::vector<double> hD((size_t)s * nb);
  std::vector<int> info(nb);
  rocblas_status st = rocblas_status_success;
  hipError_t he = hipMemcpyAsync(dT, hT.data(), sizeof(double) * hT.size(), hipMemcpyHostToDevice, ctx->stream);
  if (he == hipSuccess)
    st = rocsolver_dsyevd_strided_batched((rocblas_handle)ctx->blas, rocblas_evect_original, rocblas_fill_lower, s, dT,
                                          s, (rocblas_stride)s2, dD, s, dE, s, dinfo, nb);
  if (he == hipSuccess && st == rocblas_status_success)
    he = hipMemcpyAsync(hT.data(), dT, sizeof(double) * hT.size(), hipMemcpyDeviceToHost, ctx->stream);
  if (he == hipSuccess && st == rocblas_status_success)
    he = hipMemcpyAsync(hD.data(), dD, sizeof(double) * hD.size(), hipMemcpyDeviceToHost, ctx->stream);
  if (he == hipSuccess && st == rocblas_status_success)
    he = hipMemcpyAsync(info.data(), dinfo, sizeof(int) * nb, hipMemcpyDeviceToHost, ctx->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
  release();
  if (he != hipSuccess) return ferr(ctx, 2, std::string("projected eigenproblem: ") + hipGetErrorString(he));
  if (st != rocblas_status_success) return ferr(ctx, 4, "projected eigenproblem: rocsolver_dsyevd failed");
  for (int d = 0; d < nb; d++) {
    if (info[d] != 0) return ferr(ctx, 6, "projected eigenproblem did not converge");
    w[d].assign(hD.begin() + (size_t)d * s, hD.begin() + (size_t)(d + 1) * s);
    z[d].assign(hT.begin() + d * s2, hT.begin() + (d + 1) * s2);
    res[d].resize(s);
    for (int i = 0; i < s; i++) {
      const double *zl = z[d].data() + (size_t)i * s + (s - bs);  // last block of coefficient vector i
      double r2 = 0.0;
      for (int a = 0; a < bs; a++) {
        double v = 0.0;
        for (int b = 0; b < bs; b++) v += lo.R[d][(size_t)b * bs + a] * zl[b];
        r2 += v * v;
      }
      res[d][i] = std::sqrt(r2);
    }
  }
  return 0;
}

// Block Lanczos with bs vectors per block (same operators, modes, inner
// product and stopping test as lanczos above): each step applies the
// operator to a whole block (one multi-vector solve: every level inverse of
// the sequential factorisation is read once for the block), removes the
// basis by block classical Gram-Schmidt applied twice and K-orthonormalises
// the block by Cholesky QR applied twice.  The projection is block
// tridiagonal: T_jj = A_j (the block's own Gram-Schmidt coefficients),
// T_{j+1,j} = B_{j+1} (the Cholesky factors).  m bounds the basis size.
int lanczos_block(shellgpu_ctx *ctx, shellgpu_factor *F, shellgpu_mat *K, shellgpu_mat *G, int mode,
                  const std::vector<double> &sigma, int nev, int m, double tol, double *Q, double *KQ, int bs,
                  LanczosOut &lo) {
  const int nb = ctx->nbatch, n = 6 * ctx->nnodes;
  const size_t V = (size_t)nb * n;  // one basis vector of every design
  rocblas_handle h = (rocblas_handle)ctx->blas;
  const double one = 1.0, mone = -1.0, zero = 0.0;
  const int jmax = (m + 1) / bs - 1;  // block steps (basis blocks 0..jmax)
  if (jmax < 1) return ferr(ctx, 1, "Lanczos basis too small for the block size");
  const int smax = jmax * bs;
  lo.bs = bs;
  lo.T.assign(nb, std::vector<double>((size_t)smax * smax, 0.0));
  lo.R.assign(nb, std::vector<double>((size_t)bs * bs, 0.0));
  lo.alpha.clear();
  lo.beta.clear();
  double *dH = nullptr, *dM = nullptr;
  if (hipMalloc((void **)&dH, sizeof(double) * (size_t)nb * (smax + bs) * bs) != hipSuccess ||
      hipMalloc((void **)&dM, sizeof(double) * (size_t)nb * bs * bs) != hipSuccess) {
    if (dH) (void)hipFree(dH);
    return ferr(ctx, 2, "block Lanczos allocation failed");
  }
  struct Free {
    double *a, *b;
    ~Free() {
      (void)hipFree(a);
      (void)hipFree(b);
    }
  } fr{dH, dM};
  int rc = 0;
  // K-orthonormalise the block W (bs vectors, K W in KW) by Cholesky QR
  // twice; Rtot (per design, bs x bs column-major upper) = R2 R1
  std::vector<double> hM((size_t)nb * bs * bs), Rinv((size_t)nb * bs * bs);
  auto cholqr2 = [&](double *W, double *KW, std::vector<double> &Rtot) -> int {
    Rtot.assign((size_t)nb * bs * bs, 0.0);
    for (int d = 0; d < nb; d++)
      for (int a = 0; a < bs; a++) Rtot[(size_t)d * bs * bs + a * bs + a] = 1.0;
    for (int pass = 0; pass < 2; pass++) {
      BLASCALL(ctx, rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, bs, bs, n,
                                                  &one, W, (int)V, n, KW, (int)V, n, &zero, dM, bs, bs * bs, nb));
      HIPCALL(ctx, hipMemcpyAsync(hM.data(), dM, sizeof(double) * hM.size(), hipMemcpyDeviceToHost, ctx->stream));
      HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
      for (int d = 0; d < nb; d++) {
        double *M = hM.data() + (size_t)d * bs * bs, *Ri = Rinv.data() + (size_t)d * bs * bs;
        // upper Cholesky M = R^T R (column-major, R in M's upper triangle)
        for (int j = 0; j < bs; j++) {
          for (int i = 0; i <= j; i++) {
            double v = 0.5 * (M[j * bs + i] + M[i * bs + j]);
            for (int k = 0; k < i; k++) v -= M[i * bs + k] * M[j * bs + k];
            if (i == j) {
              if (!(v > 0.0)) return ferr(ctx, 6, "block Lanczos breakdown (rank-deficient block)");
              M[j * bs + j] = std::sqrt(v);
            } else {
              M[j * bs + i] = v / M[i * bs + i];
            }
          }
        }
        // R stored in M(i, j) = M[j * bs + i] for i <= j; Ri = R^{-1} (upper)
        std::fill(Ri, Ri + bs * bs, 0.0);
        for (int j = 0; j < bs; j++) {
          Ri[j * bs + j] = 1.0 / M[j * bs + j];
          for (int i = j - 1; i >= 0; i--) {
            double v = 0.0;
            for (int k = i + 1; k <= j; k++) v += M[k * bs + i] * Ri[j * bs + k];
            Ri[j * bs + i] = -v / M[i * bs + i];
          }
        }
        // Rtot = R Rtot (both upper)
        double *Rt = Rtot.data() + (size_t)d * bs * bs;
        std::vector<double> P((size_t)bs * bs, 0.0);
        for (int j = 0; j < bs; j++)
          for (int i = 0; i <= j; i++) {
            double v = 0.0;
            for (int k = i; k <= j; k++) v += M[k * bs + i] * Rt[j * bs + k];
            P[j * bs + i] = v;
          }
        std::copy(P.begin(), P.end(), Rt);
      }
      HIPCALL(ctx, hipMemcpyAsync(dM, Rinv.data(), sizeof(double) * Rinv.size(), hipMemcpyHostToDevice, ctx->stream));
      BLASCALL(ctx, rocblas_dtrmm_strided_batched(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                                  rocblas_diagonal_non_unit, n, bs, &one, dM, bs, bs * bs, W, (int)V,
                                                  n, W, (int)V, n, nb));
      BLASCALL(ctx, rocblas_dtrmm_strided_batched(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                                  rocblas_diagonal_non_unit, n, bs, &one, dM, bs, bs * bs, KW, (int)V,
                                                  n, KW, (int)V, n, nb));
      // the host copies of Rinv must outlive the transfers above
      HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
    }
    return 0;
  };
  {
    std::vector<double> q0((size_t)bs * V);
    unsigned long long sd = 0x5eedULL;
    for (auto &v : q0) {
      sd += 0x9E3779B97F4A7C15ULL;
      unsigned long long zz = sd;
      zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ULL;
      zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBULL;
      zz ^= zz >> 31;
      v = 2.0 * ((double)(zz >> 11) / 9007199254740992.0) - 1.0;
    }
    HIPCALL(ctx, hipMemcpyAsync(Q, q0.data(), sizeof(double) * q0.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < bs; c++) {
      if ((rc = shellgpu_vec_zero_bcs(ctx, Q + c * V))) return rc;
      if ((rc = shellgpu_mat_mult(K, Q + c * V, KQ + c * V))) return rc;
    }
    std::vector<double> R0;
    if ((rc = cholqr2(Q, KQ, R0))) return rc;
  }
  std::vector<std::vector<double>> Hsum(nb);
  std::vector<double> hH;
  int j, s = 0;
  for (j = 0; j < jmax; j++) {
    double *Qj = Q + (size_t)j * bs * V, *KQj = KQ + (size_t)j * bs * V;
    double *W = Q + (size_t)(j + 1) * bs * V, *KW = KQ + (size_t)(j + 1) * bs * V;
    if (mode == 0 || mode == 2) {
      if ((rc = factor_apply_multi(F, KQj, W, bs))) return rc;
    } else {
      for (int c = 0; c < bs; c++) {
        if ((rc = shellgpu_mat_mult(G, Qj + c * V, KW + c * V))) return rc;
        if ((rc = shellgpu_vec_zero_bcs(ctx, KW + c * V))) return rc;
      }
      if ((rc = factor_apply_multi(F, KW, W, bs))) return rc;
      if ((rc = shellgpu_vec_scale(ctx, bs * V, -1.0, W))) return rc;
    }
    for (int c = 0; c < bs; c++)
      if ((rc = shellgpu_vec_zero_bcs(ctx, W + c * V))) return rc;
    const int ncol = (j + 1) * bs;
    for (int d = 0; d < nb; d++) Hsum[d].assign((size_t)ncol * bs, 0.0);
    hH.resize((size_t)nb * ncol * bs);
    for (int pass = 0; pass < 2; pass++) {
      BLASCALL(ctx, rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, ncol, bs, n,
                                                  &one, KQ, (int)V, n, W, (int)V, n, &zero, dH, ncol, ncol * bs, nb));
      BLASCALL(ctx, rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, n, bs, ncol,
                                                  &mone, Q, (int)V, n, dH, ncol, ncol * bs, &one, W, (int)V, n, nb));
      HIPCALL(ctx, hipMemcpyAsync(hH.data(), dH, sizeof(double) * hH.size(), hipMemcpyDeviceToHost, ctx->stream));
      HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
      for (int d = 0; d < nb; d++)
        for (size_t q = 0; q < (size_t)ncol * bs; q++) Hsum[d][q] += hH[(size_t)d * ncol * bs + q];
    }
    for (int c = 0; c < bs; c++)
      if ((rc = shellgpu_mat_mult(K, W + c * V, KW + c * V))) return rc;
    std::vector<double> Rb;
    if ((rc = cholqr2(W, KW, Rb))) return rc;
    // projection blocks: T_jj = A_j (symmetrised), T_{j+1,j} = B_{j+1}
    for (int d = 0; d < nb; d++) {
      std::vector<double> &T = lo.T[d];
      const double *Hd = Hsum[d].data();  // ncol x bs, column-major
      for (int b = 0; b < bs; b++)
        for (int a = 0; a < bs; a++) {
          const double v = 0.5 * (Hd[(size_t)b * ncol + j * bs + a] + Hd[(size_t)a * ncol + j * bs + b]);
          T[(size_t)(j * bs + b) * smax + j * bs + a] = v;
        }
      const double *B = Rb.data() + (size_t)d * bs * bs;
      std::copy(B, B + bs * bs, lo.R[d].begin());
      if (j + 1 < jmax)
        for (int b = 0; b < bs; b++)
          for (int a = 0; a <= b; a++) {
            const double v = B[(size_t)b * bs + a];  // B(a, b), upper
            T[(size_t)(j * bs + b) * smax + (j + 1) * bs + a] = v;   // T(j+1 block row a, j block col b)
            T[(size_t)((j + 1) * bs + a) * smax + j * bs + b] = v;   // transpose
          }
    }
    s = (j + 1) * bs;
    if (s >= nev) {
      lo.steps = smax;  // T's leading dimension while iterating
      std::vector<std::vector<double>> w, z, res;
      if ((rc = proj_eigs(ctx, lo, s, w, z, res))) return rc;
      bool all = true;
      for (int d = 0; d < nb && all; d++) {
        std::vector<int> idx(s);
        for (int i = 0; i < s; i++) idx[i] = i;
        auto key = [&](int i) {
          const double th = w[d][i];
          if (mode == 2) return std::fabs(1.0 / th + sigma[d]);
          if (mode == 3) return th != 0.0 ? std::fabs(1.0 / th) : 1e300;
          return mode == 0 ? std::fabs(-th * sigma[d] / (1.0 - th)) : (th > 0 ? 1.0 / th : 1e300);
        };
        std::sort(idx.begin(), idx.end(), [&](int a, int b) { return key(a) < key(b); });
        for (int q = 0; q < nev; q++)
          if (!(res[d][idx[q]] <= tol * std::fabs(w[d][idx[q]]))) all = false;
      }
      if (all) break;
    }
  }
  // compact T to the final basis size s
  for (int d = 0; d < nb; d++) {
    std::vector<double> Tc((size_t)s * s);
    for (int c = 0; c < s; c++)
      for (int r = 0; r < s; r++) Tc[(size_t)c * s + r] = lo.T[d][(size_t)c * smax + r];
    lo.T[d].swap(Tc);
  }
  lo.steps = s;
  return 0;
}

// Vectors per Lanczos block: 1 (the single-vector recurrence) unless
// SHELLGPU_LANCZOS_BLOCK=b (2 ... 8) asks for blocks, which the sequential
// factorisation and a cyclic-reduction factor of one design solve together
// (one pass over the factor per block).  Measured on 64 designs of the
// 100 x 200 cylinder, blocks of 4 read the level inverses 2.5 times less
// per vector but need a basis as large as the single-vector run's, and the
// block Gram-Schmidt products (tall, skinny, transposed GEMMs) cost more
// than the saving.
int lanczos_block_size(const shellgpu_factor *F) {
  if (F->nbatch != F->ctx->nbatch || (F->cr && F->nbatch != 1)) return 1;
  int bs = 1;
  if (const char *env = getenv("SHELLGPU_LANCZOS_BLOCK")) bs = atoi(env);
  return std::max(1, std::min(8, bs));
}

int run_lanczos(shellgpu_ctx *ctx, shellgpu_factor *F, shellgpu_mat *K, shellgpu_mat *G, int mode,
                const std::vector<double> &sigma, int nev, int m, double tol, double *Q, double *KQ, double *hv,
                double *work, LanczosOut &lo, bool single = false) {
  const int bs = single ? 1 : lanczos_block_size(F);
  // the block recurrence hands rocBLAS the basis with a leading dimension of
  // nbatch * n (an int)
  const bool ld_fits = (long)ctx->nbatch * 6 * ctx->nnodes <= 0x7fffffffL;
  if (bs == 1 || (m + 1) / bs < 3 || !ld_fits) {
    lo.bs = 1;
    return lanczos(ctx, F, K, G, mode, sigma, nev, m, tol, Q, KQ, hv, work, lo);
  }
  lo.ns.clear();
  return lanczos_block(ctx, F, K, G, mode, sigma, nev, m, tol, Q, KQ, bs, lo);
}

}  // namespace

struct shellgpu_buckling {
  shellgpu_ctx *ctx = nullptr;
  int model = 0, nev = 0, m = 0;
  double sigma = 0.0, tol = 0.0;
  shellgpu_mat *K = nullptr, *G = nullptr, *A = nullptr;
  shellgpu_factor *F = nullptr;
  // device: path u, work vectors, Lanczos basis Q and K Q, coefficients
  double *u = nullptr, *res = nullptr, *f0 = nullptr, *Q = nullptr, *KQ = nullptr, *hv = nullptr, *dsig = nullptr;
  double *ph = nullptr, *gv = nullptr, *psi = nullptr, *dcoef = nullptr;  // dcoef: 6 per design
  std::vector<double> sig;
  bool solved = false;
  int steps = 0;
  // per design: eigenvalues, Ritz residuals and Ritz coefficient vectors
  // (nev x ns, ns = Lanczos basis size of that design)
  std::vector<std::vector<double>> lam, err, ritz;
  std::vector<int> ns;
  // K, G, A and F belong to the object unless the caller supplied them
  bool own_ops = true;
  // a second factor holding K (zero state) for the static path, the shift
  // estimate and the path adjoint, when device memory allows; otherwise
  // FK == F and K is refactored where needed
  shellgpu_factor *FK = nullptr;
  bool fk_valid = false;
  // FK is the caller's (shellgpu_buckling_set_stiffness_factor): kept, not
  // refactored, not destroyed
  bool fk_ext = false;
  // static-path GMRES settings
  int ksm_m = 10, ksm_restarts = 15;
  double ksm_rtol = 1e-12, ksm_atol = 1e-30;
};

namespace {

int buckling_free(shellgpu_buckling *B) {
  if (B->FK && B->FK != B->F && !B->fk_ext) shellgpu_factor_destroy(B->FK);
  B->FK = nullptr;
  if (B->own_ops) {
    if (B->K) shellgpu_mat_destroy(B->K);
    if (B->G) shellgpu_mat_destroy(B->G);
    if (B->A) shellgpu_mat_destroy(B->A);
    if (B->F) shellgpu_factor_destroy(B->F);
  }
  double *bufs[] = {B->u, B->res, B->f0, B->Q, B->KQ, B->hv, B->dsig, B->ph, B->gv, B->psi, B->dcoef};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  return 0;
}

// Ritz vector idx of every design, Q z, into out (device, nbatch x n); ritz[d]
// holds the wanted Ritz coefficient vectors of design d (ns[d] each)
int ritz_vector(shellgpu_ctx *ctx, const double *Q, const std::vector<int> &ns,
                const std::vector<std::vector<double>> &ritz, int idx, double *hv, double *out) {
  const int nb = ctx->nbatch, n = 6 * ctx->nnodes;
  rocblas_handle h = (rocblas_handle)ctx->blas;
  const double one = 1.0, zero = 0.0;
  for (int d = 0; d < nb; d++) {
    HIPCALL(ctx, hipMemcpyAsync(hv, ritz[d].data() + (size_t)idx * ns[d], sizeof(double) * ns[d],
                                hipMemcpyHostToDevice, ctx->stream));
    BLASCALL(ctx, rocblas_dgemv(h, rocblas_operation_none, n, ns[d], &one, Q + (size_t)d * n, nb * n, hv, 1, &zero,
                                out + (size_t)d * n, 1));
    // hv is reused by the next design
    HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

// phi_n of every design into out (device, nbatch x n)
int buckling_vector(shellgpu_buckling *B, int idx, double *out) {
  return ritz_vector(B->ctx, B->Q, B->ns, B->ritz, idx, B->hv, out);
}

}  // namespace

extern "C" {

int shellgpu_buckling_create(shellgpu_ctx *ctx, int model, double sigma, int nev, int max_lanczos, double tol,
                             shellgpu_buckling **out) {
  if (!ctx || !out) return 1;
  *out = nullptr;
  if (nev <= 0 || max_lanczos < nev || !(tol > 0.0)) return ferr(ctx, 1, "invalid buckling arguments");
  if (!ctx->d_thick) return ferr(ctx, 3, "designs not set (shellgpu_set_designs)");
  shellgpu_buckling *B = new shellgpu_buckling();
  B->ctx = ctx;
  B->model = model == SHELLGPU_NONLINEAR ? SHELLGPU_NONLINEAR : SHELLGPU_LINEAR;
  B->sigma = sigma;
  B->nev = nev;
  B->m = max_lanczos;
  B->tol = tol;
  const int nb = ctx->nbatch, n = 6 * ctx->nnodes, m = max_lanczos;
  int rc = 0;
  if ((rc = ensure_blas(ctx))) goto fail;
  if ((rc = shellgpu_mat_create(ctx, nb, &B->K))) goto fail;
  if ((rc = shellgpu_mat_create(ctx, nb, &B->G))) goto fail;
  if ((rc = shellgpu_mat_create(ctx, nb, &B->A))) goto fail;
  if ((rc = shellgpu_factor_create(ctx, &B->F))) goto fail;
  // the factor of K (shared by identical designs), when memory allows
  if (factor_create(ctx, ctx->same_designs ? 1 : nb, &B->FK) != 0) {
    B->FK = B->F;
    ctx->err.clear();
    (void)hipGetLastError();  // drop the failed allocation's error state
  }
  {
    const size_t v = sizeof(double) * (size_t)n * nb;
    double **vecs[] = {&B->u, &B->res, &B->f0, &B->ph, &B->gv, &B->psi};
    for (double **p : vecs) {
      if (hipMalloc((void **)p, v) != hipSuccess) {
        rc = ferr(ctx, 2, "out of device memory (buckling vectors)");
        goto fail;
      }
    }
    if (hipMalloc((void **)&B->Q, v * (m + 1)) != hipSuccess ||
        hipMalloc((void **)&B->KQ, v * (m + 1)) != hipSuccess ||
        hipMalloc((void **)&B->hv, sizeof(double) * (size_t)nb * (m + 1)) != hipSuccess ||
        hipMalloc((void **)&B->dsig, sizeof(double) * nb) != hipSuccess ||
        hipMalloc((void **)&B->dcoef, sizeof(double) * 6 * nb) != hipSuccess) {
      rc = ferr(ctx, 2, "out of device memory (Lanczos basis)");
      goto fail;
    }
  }
  *out = B;
  return 0;
fail:
  buckling_free(B);
  delete B;
  return rc;
}

void shellgpu_buckling_destroy(shellgpu_buckling *B) {
  if (!B) return;
  buckling_free(B);
  delete B;
}

int shellgpu_buckling_set_stiffness_factor(shellgpu_buckling *B, shellgpu_factor *F) {
  if (!B) return 1;
  shellgpu_ctx *ctx = B->ctx;
  if (!F || F->ctx != ctx) return ferr(ctx, 1, "a factor of this context is required");
  if (F == B->F) return ferr(ctx, 1, "the stiffness factor must differ from the factor of A");
  if (!F->factored) return ferr(ctx, 9, "the stiffness factor holds no factorisation");
  if (F->nbatch != ctx->nbatch) return ferr(ctx, 1, "factor batch size does not match the context's designs");
  if (B->FK && B->FK != B->F && !B->fk_ext) shellgpu_factor_destroy(B->FK);
  B->FK = F;
  B->fk_ext = true;
  B->fk_valid = true;
  return 0;
}

int shellgpu_buckling_set_operators(shellgpu_buckling *B, shellgpu_mat *K, shellgpu_mat *G, shellgpu_mat *A,
                                    shellgpu_factor *F) {
  if (!B) return 1;
  shellgpu_ctx *ctx = B->ctx;
  if (!K || !G || !A || !F) return ferr(ctx, 1, "K, G, A and the factor are required");
  if (K == G || K == A || G == A) return ferr(ctx, 1, "K, G and A must be distinct matrices");
  if (B->fk_ext && F == B->FK) return ferr(ctx, 1, "the factor of A must differ from the stiffness factor");
  const int nb = ctx->nbatch;
  if (K->nbatch != nb || G->nbatch != nb || A->nbatch != nb)
    return ferr(ctx, 1, "matrix batch size does not match the context's designs");
  if (B->FK == B->F) B->FK = nullptr;  // the K factor fell back to the replaced factor
  if (B->own_ops) {
    shellgpu_mat_destroy(B->K);
    shellgpu_mat_destroy(B->G);
    shellgpu_mat_destroy(B->A);
    shellgpu_factor_destroy(B->F);
  }
  B->K = K;
  B->G = G;
  B->A = A;
  B->F = F;
  B->own_ops = false;
  B->solved = false;
  return 0;
}

int shellgpu_buckling_set_ksm(shellgpu_buckling *B, int msub, int nrestart, double rtol, double atol) {
  if (!B) return 1;
  if (msub <= 0 || nrestart < 0 || !(rtol >= 0.0) || !(atol >= 0.0))
    return ferr(B->ctx, 1, "invalid GMRES settings");
  B->ksm_m = msub;
  B->ksm_restarts = nrestart;
  B->ksm_rtol = rtol;
  B->ksm_atol = atol;
  return 0;
}

int shellgpu_buckling_solve(shellgpu_buckling *B, const double *force, const double *u0) {
  if (!B) return 1;
  shellgpu_ctx *ctx = B->ctx;
  const int nb = ctx->nbatch, n = 6 * ctx->nnodes, m = B->m, nev = B->nev;
  int rc = 0;
  VecOps vo;
  LanczosOut lo;
  B->solved = false;
  B->sig.assign(nb, B->sigma);
  if ((rc = vo.init(ctx, n, nb))) return rc;

  // TACSLinearBuckling::solve (src/TACSBuckling.cpp:197-278): K is
  // assembled at zero state, the path is the static solution K u = f (with
  // the BC-driven residual) unless u0 is given
  if ((rc = shellgpu_assemble_mattype(ctx, SHELLGPU_STIFFNESS, nullptr, B->K))) return rc;
  // identical designs (load cases of one structure) share one factor of K;
  // the designs may have changed since the factor was made
  const int fk_nb = ctx->same_designs ? 1 : nb;
  if (!B->fk_ext) {
    if (B->FK && B->FK != B->F && B->FK->nbatch != fk_nb) {
      shellgpu_factor_destroy(B->FK);
      B->FK = nullptr;
    }
    if (!B->FK && factor_create(ctx, fk_nb, &B->FK) != 0) {
      B->FK = B->F;  // no room for a second factor
      ctx->err.clear();
      (void)hipGetLastError();  // drop the failed allocation's error state
    }
    B->fk_valid = false;
  }
  if (!B->fk_valid && (!u0 || B->sigma <= 0.0)) {
    if ((rc = shellgpu_factor_compute(B->FK, B->K))) return rc;
    B->fk_valid = true;
  }
  if (u0) {
    HIPCALL(ctx, hipMemcpyAsync(B->u, u0, sizeof(double) * (size_t)n * nb, hipMemcpyDeviceToDevice, ctx->stream));
  } else {
    if ((rc = shellgpu_assemble_res(ctx, SHELLGPU_LINEAR, nullptr, B->res))) return rc;
    if (force) {
      HIPCALL(ctx,
              hipMemcpyAsync(B->f0, force, sizeof(double) * (size_t)n * nb, hipMemcpyDeviceToDevice, ctx->stream));
      if ((rc = shellgpu_vec_zero_bcs(ctx, B->f0))) return rc;
      std::vector<double> m1(nb, -1.0);
      if ((rc = vo.axpy(m1, B->f0, B->res))) return rc;
    }
    int its = 0;
    double rn = 0.0;
    if ((rc = shellgpu_gmres(ctx, B->K, B->FK, B->res, B->u, B->ksm_m, B->ksm_restarts, B->ksm_rtol, B->ksm_atol,
                             &its, &rn)))
      return rc;
    std::vector<double> m1(nb, -1.0);
    if ((rc = vo.scale(B->u, m1))) return rc;
  }
  if ((rc = shellgpu_vec_set_bcs(ctx, B->u))) return rc;
  if ((rc = shellgpu_assemble_mattype(ctx, SHELLGPU_GEOMETRIC, B->u, B->G))) return rc;

  // Load cases of one structure (identical designs, one shared factor of K)
  // with no shift given first run buckling mode on that factor, w =
  // K^{-1}(-G q): it needs no factorisation per load case, where
  // shift-invert factors K + sigma G for every case.  Its spectrum near the
  // wanted loads is less separated, so it can run out of steps (clustered
  // cylinder loads do); a design with an unconverged wanted pair sends the
  // whole batch down the shift-invert path, with the shift taken from this
  // run's Ritz values.
  bool buck = false, have_shift = false;
  if (B->sigma <= 0.0 && nb > 1 && B->FK != B->F && B->FK->nbatch == 1) {
    if ((rc = run_lanczos(ctx, B->FK, B->K, B->G, 3, B->sig, nev, m, B->tol, B->Q, B->KQ, B->hv, B->res, lo, true)))
      return rc;
    std::vector<std::vector<double>> w, z, res;
    if ((rc = proj_eigs(ctx, lo, lo.steps, w, z, res))) return rc;
    buck = lo.steps >= nev;
    for (int d = 0; d < nb && buck; d++) {
      const int sd = (int)w[d].size();
      std::vector<int> idx(sd);
      for (int i = 0; i < sd; i++) idx[i] = i;
      auto key = [&](int i) { return w[d][i] != 0.0 ? std::fabs(1.0 / w[d][i]) : 1e300; };
      std::sort(idx.begin(), idx.end(), [&](int a, int b) { return key(a) < key(b); });
      for (int j = 0; j < nev; j++)
        if (!(res[d][idx[j]] <= B->tol * std::fabs(w[d][idx[j]]))) buck = false;
    }
    if (!buck) {
      for (int d = 0; d < nb; d++) {
        const int i = (int)(std::max_element(w[d].begin(), w[d].end()) - w[d].begin());
        if (!(w[d][i] > 0.0)) return ferr(ctx, 7, "no positive buckling load under this load case");
        B->sig[d] = 0.98 / (w[d][i] + res[d][i]);
      }
      have_shift = true;
    }
  }

  if (!buck) {
    // shift: sigma > 0 as given; sigma <= 0 places it at 0.98 of a lower
    // bound of the first load from a short buckling-mode Lanczos on
    // K^{-1}(-G): its largest Ritz value theta (residual r) lies within r
    // of the largest eigenvalue 1 / lambda_1, so lambda_1 >= 1 / (theta +
    // r).  Closer shifts converge the main run in fewer steps (45 against
    // 65 at 0.95 on the 200 x 400 cylinder); a shift above lambda_1 is
    // caught by the Cholesky and lowered.
    if (B->sigma <= 0.0 && !have_shift) {
      if ((rc = run_lanczos(ctx, B->FK, B->K, B->G, 1, B->sig, 1, std::min(m, 40), 1e-2, B->Q, B->KQ, B->hv,
                            B->res, lo, true)))
        return rc;
      std::vector<std::vector<double>> w, z, res;
      if ((rc = proj_eigs(ctx, lo, lo.steps, w, z, res))) return rc;
      for (int d = 0; d < nb; d++) {
        const int i = (int)(std::max_element(w[d].begin(), w[d].end()) - w[d].begin());
        const double tmax = w[d][i];
        if (!(tmax > 0.0)) return ferr(ctx, 7, "no positive buckling load under this load case");
        B->sig[d] = 0.98 / (tmax + res[d][i]);
      }
    }
    // A = K + sigma G with BCs; factor, lowering the shift of any design
    // whose shifted matrix is indefinite (shift above its first load)
    for (int attempt = 0;; attempt++) {
      HIPCALL(ctx, hipMemcpyAsync(B->dsig, B->sig.data(), sizeof(double) * nb, hipMemcpyHostToDevice, ctx->stream));
      {
        const long nv = 36L * ctx->nnzb;
        hipLaunchKernelGGL(shifted_kernel, dim3((unsigned)((nv + 255) / 256), nb), dim3(256), 0, ctx->stream, nv,
                           B->K->vals, B->G->vals, B->dsig, B->A->vals);
        B->A->same = false;
        HIPCALL(ctx, hipGetLastError());
      }
      if ((rc = shellgpu_mat_apply_bcs(B->A, 1.0))) return rc;
      if (B->FK == B->F) B->fk_valid = false;
      rc = shellgpu_factor_compute(B->F, B->A);
      if (rc == 0) break;
      if (rc != 5 || attempt >= 8) return rc;
      std::vector<int> info(nb);
      HIPCALL(ctx, hipMemcpy(info.data(), B->F->info, sizeof(int) * nb, hipMemcpyDeviceToHost));
      for (int d = 0; d < nb; d++)
        if (info[d] != 0) B->sig[d] *= (B->sigma <= 0.0 ? 0.8 : 0.5);
    }
    if ((rc = run_lanczos(ctx, B->F, B->K, B->G, 0, B->sig, nev, m, B->tol, B->Q, B->KQ, B->hv, B->res, lo)))
      return rc;
  } else {
    B->sig.assign(nb, 0.0);  // reported shift 0: buckling mode
  }
  B->steps = lo.steps;
  B->lam.assign(nb, {});
  B->err.assign(nb, {});
  B->ritz.assign(nb, {});
  B->ns.assign(nb, 0);
  const int ns = lo.steps;
  if (ns < nev) return ferr(ctx, 8, "Lanczos basis smaller than the number of eigenvalues");
  std::vector<std::vector<double>> wv, zv, rv;
  if ((rc = proj_eigs(ctx, lo, ns, wv, zv, rv))) return rc;
  for (int d = 0; d < nb; d++) {
    const std::vector<double> &w = wv[d], &z = zv[d];
    const int ns = (int)w.size();  // this design's basis size
    if (ns < nev) return ferr(ctx, 8, "Lanczos basis smaller than the number of eigenvalues");
    std::vector<int> idx(ns);
    for (int i = 0; i < ns; i++) idx[i] = i;
    const double sd = B->sig[d];
    // theta = 1 / lambda (buckling mode), lambda / (lambda - sigma)
    // (shift-invert)
    auto lam_of = [&](int i) { return buck ? 1.0 / w[i] : -w[i] * sd / (1.0 - w[i]); };
    // sortEigenvalues, SMALLEST_MAGNITUDE (src/bpmat/GSEP.cpp:693-748)
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return std::fabs(lam_of(a)) < std::fabs(lam_of(b)); });
    B->ns[d] = ns;
    for (int j = 0; j < nev; j++) {
      const int i = idx[j];
      B->lam[d].push_back(lam_of(i));
      B->err[d].push_back(rv[d][i] / std::fabs(w[i]));
      B->ritz[d].insert(B->ritz[d].end(), z.begin() + (size_t)i * ns, z.begin() + (size_t)(i + 1) * ns);
    }
  }
  B->solved = true;
  return 0;
}

int shellgpu_buckling_extract(shellgpu_buckling *B, int idx, double *lambda, double *phi, double *error) {
  if (!B) return 1;
  shellgpu_ctx *ctx = B->ctx;
  if (!B->solved) return ferr(ctx, 9, "buckling problem not solved");
  if (idx < 0 || idx >= B->nev) return ferr(ctx, 1, "eigenvalue index out of range");
  const int nb = ctx->nbatch;
  for (int d = 0; d < nb; d++) {
    if (lambda) lambda[d] = B->lam[d][idx];
    if (error) error[d] = B->err[d][idx];
  }
  if (phi) return buckling_vector(B, idx, phi);
  return 0;
}

int shellgpu_buckling_info(shellgpu_buckling *B, int *nsteps, double *sigma_used, double *path) {
  if (!B) return 1;
  shellgpu_ctx *ctx = B->ctx;
  if (!B->solved) return ferr(ctx, 9, "buckling problem not solved");
  if (nsteps) *nsteps = B->steps;
  if (sigma_used)
    for (int d = 0; d < ctx->nbatch; d++) sigma_used[d] = B->sig[d];
  if (path)
    HIPCALL(ctx, hipMemcpyAsync(path, B->u, sizeof(double) * 6L * ctx->nnodes * ctx->nbatch, hipMemcpyDeviceToDevice,
                                ctx->stream));
  return 0;
}

int shellgpu_buckling_add_eigen_sens(shellgpu_buckling *B, const double *coef, double *dfdt, double *dfdX) {
  if (!B || !coef) return 1;
  shellgpu_ctx *ctx = B->ctx;
  if (!B->solved) return ferr(ctx, 9, "buckling problem not solved");
  if (!dfdt && !dfdX) return 0;
  const int nb = ctx->nbatch, n = 6 * ctx->nnodes, nev = B->nev;
  int rc = 0;
  VecOps vo;
  if ((rc = vo.init(ctx, n, nb))) return rc;
  // coefficient layout: [0, 2 nb) pairs for the thickness kernels, [2 nb,
  // 6 nb) quadruples for the coordinate kernel
  double *c2 = B->dcoef, *c4 = B->dcoef + 2 * nb;
  HIPCALL(ctx, hipMemsetAsync(B->gv, 0, sizeof(double) * (size_t)n * nb, ctx->stream));
  std::vector<double> hc(6 * nb, 0.0);
  for (int j = 0; j < nev; j++) {
    bool any = false;
    for (int d = 0; d < nb; d++) any = any || coef[(size_t)d * nev + j] != 0.0;
    if (!any) continue;
    // phi_j, and the reference's normalisation phi^T G phi
    if ((rc = buckling_vector(B, j, B->ph))) return rc;
    if ((rc = shellgpu_mat_mult(B->G, B->ph, B->res))) return rc;
    std::vector<double> pgp;
    if ((rc = vo.dot(B->ph, B->res, pgp))) return rc;
    for (int d = 0; d < nb; d++) {
      // d lambda = -(phi^T K' phi + lambda phi^T G' phi - lambda psi^T R') / (phi^T G phi)
      const double f = -coef[(size_t)d * nev + j] / pgp[d];
      hc[2 * d] = f;
      hc[2 * d + 1] = f * B->lam[d][j];
      hc[2 * nb + 4 * d] = f;
      hc[2 * nb + 4 * d + 1] = f * B->lam[d][j];
      hc[2 * nb + 4 * d + 2] = 0.0;
    }
    HIPCALL(ctx, hipMemcpyAsync(B->dcoef, hc.data(), sizeof(double) * 6 * nb, hipMemcpyHostToDevice, ctx->stream));
    ElemArgs a = base_args(ctx, B->ph);
    a.u2 = B->u;
    a.u2_stride = n;
    a.coef = c2;
    a.dv = dfdt;
    a.res = B->gv;
    if ((rc = run_colored(ctx, kModeAdjDV, a))) return rc;
    if (dfdX && (rc = run_xpt(ctx, B->ph, B->u, nullptr, c4, 0, kXptK | kXptG, dfdX))) return rc;
    HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  }
  // path adjoint: dR/du psi = sum_j f_j lambda_j d(phi_j^T G phi_j)/du, with
  // the tangent of the residual that defines the path (K for the linear
  // model; the nonlinear tangent at u otherwise)
  if ((rc = shellgpu_vec_zero_bcs(ctx, B->gv))) return rc;
  shellgpu_mat *T = B->K;
  if (B->model == SHELLGPU_NONLINEAR) {
    if ((rc = shellgpu_assemble_jacobian(ctx, SHELLGPU_NONLINEAR, 1.0, 0.0, 0.0, B->u, nullptr, nullptr, nullptr, B->A))) return rc;
    T = B->A;
  }
  {
    int its = 0;
    double rn = 1.0;
    if (B->FK != B->F) {
      // GMRES on the path tangent preconditioned by the factor of K (exact
      // for the linear path; a few iterations for a nonlinear path)
      if (!B->fk_valid) {
        if ((rc = shellgpu_factor_compute(B->FK, B->K))) return rc;
        B->fk_valid = true;
      }
      if ((rc = shellgpu_gmres(ctx, T, B->FK, B->gv, B->psi, 20, 4, 1e-12, 1e-300, &its, &rn))) return rc;
    }
    if (!(rn <= 1e-10)) {
      // factor the tangent itself (the shifted operator's factor is lost
      // when the factors are shared: a later solve refactors)
      if (B->FK == B->F) B->fk_valid = false;
      if ((rc = shellgpu_factor_compute(B->F, T))) return rc;
      if ((rc = shellgpu_gmres(ctx, T, B->F, B->gv, B->psi, 10, 3, 1e-12, 1e-300, &its, &rn))) return rc;
    }
  }
  for (int d = 0; d < nb; d++) {
    hc[d] = -1.0;
    hc[2 * nb + 4 * d] = 0.0;
    hc[2 * nb + 4 * d + 1] = 0.0;
    hc[2 * nb + 4 * d + 2] = -1.0;
  }
  HIPCALL(ctx, hipMemcpyAsync(B->dcoef, hc.data(), sizeof(double) * 6 * nb, hipMemcpyHostToDevice, ctx->stream));
  const int nonlin = B->model == SHELLGPU_NONLINEAR;
  if (dfdt) {
    ElemArgs a = base_args(ctx, B->u);
    a.psi = B->psi;
    a.psi_stride = n;
    a.coef = c2;
    a.dv = dfdt;
    if ((rc = run_colored(ctx, nonlin ? kModeAdjResNl : kModeAdjResLin, a))) return rc;
  }
  if (dfdX && (rc = run_xpt(ctx, nullptr, B->u, B->psi, c4, nonlin, kXptR, dfdX))) return rc;
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int shellgpu_linear_buckling(shellgpu_ctx *ctx, double sigma, const double *force, int nev, int max_lanczos,
                             double tol, double *lambda, double *phi, double *path, int *nsteps, double *resid,
                             double *sigma_used) {
  shellgpu_buckling *B = nullptr;
  int rc = shellgpu_buckling_create(ctx, SHELLGPU_LINEAR, sigma, nev, max_lanczos, tol, &B);
  if (rc) return rc;
  const int nb = ctx->nbatch, n = 6 * ctx->nnodes;
  std::vector<double> lam(nb), err(nb);
  if ((rc = shellgpu_buckling_solve(B, force, nullptr))) goto out;
  if ((rc = shellgpu_buckling_info(B, nsteps, sigma_used, path))) goto out;
  for (int j = 0; j < nev; j++) {
    if ((rc = shellgpu_buckling_extract(B, j, lam.data(), phi ? B->ph : nullptr, err.data()))) goto out;
    for (int d = 0; d < nb; d++) {
      lambda[(size_t)d * nev + j] = lam[d];
      if (resid) resid[(size_t)d * nev + j] = err[d];
      if (phi && hipMemcpyAsync(phi + ((size_t)d * nev + j) * n, B->ph + (size_t)d * n, sizeof(double) * n,
                                hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) {
        rc = ferr(ctx, 2, "eigenvector copy failed");
        goto out;
      }
    }
  }
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = ferr(ctx, 2, "stream synchronisation failed");
out:
  shellgpu_buckling_destroy(B);
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// natural frequencies (TACSFrequencyAnalysis, Lanczos path)
// ---------------------------------------------------------------------------
struct shellgpu_frequency {
  shellgpu_ctx *ctx = nullptr;
  int nev = 0, m = 0;
  double sigma = 0.0, tol = 0.0;
  shellgpu_mat *K = nullptr, *M = nullptr, *A = nullptr;
  shellgpu_factor *F = nullptr;
  // device: Lanczos basis Q and M Q, coefficients, work vectors, shifts
  double *Q = nullptr, *MQ = nullptr, *hv = nullptr, *work = nullptr, *ph = nullptr, *dsig = nullptr;
  std::vector<double> sig;
  bool solved = false;
  int steps = 0;
  std::vector<std::vector<double>> lam, err, ritz;
  std::vector<int> ns;
};

namespace {

void frequency_free(shellgpu_frequency *F) {
  if (F->K) shellgpu_mat_destroy(F->K);
  if (F->M) shellgpu_mat_destroy(F->M);
  if (F->A) shellgpu_mat_destroy(F->A);
  if (F->F) shellgpu_factor_destroy(F->F);
  double *bufs[] = {F->Q, F->MQ, F->hv, F->work, F->ph, F->dsig};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
}

}  // namespace

extern "C" {

int shellgpu_frequency_create(shellgpu_ctx *ctx, double sigma, int nev, int max_lanczos, double tol,
                              shellgpu_frequency **out) {
  if (!ctx || !out) return 1;
  *out = nullptr;
  if (nev <= 0 || max_lanczos < nev || !(tol > 0.0) || !(sigma >= 0.0))
    return ferr(ctx, 1, "invalid frequency-analysis arguments (nev > 0, max_lanczos >= nev, tol > 0, sigma >= 0)");
  if (!ctx->d_thick) return ferr(ctx, 3, "designs not set (shellgpu_set_designs)");
  shellgpu_frequency *F = new shellgpu_frequency();
  F->ctx = ctx;
  F->sigma = sigma;
  F->nev = nev;
  F->m = max_lanczos;
  F->tol = tol;
  const int nb = ctx->nbatch, n = 6 * ctx->nnodes, m = max_lanczos;
  const size_t v = sizeof(double) * (size_t)n * nb;
  int rc = 0;
  if ((rc = ensure_blas(ctx))) goto fail;
  if ((rc = shellgpu_mat_create(ctx, nb, &F->K))) goto fail;
  if ((rc = shellgpu_mat_create(ctx, nb, &F->M))) goto fail;
  if ((rc = shellgpu_mat_create(ctx, nb, &F->A))) goto fail;
  if ((rc = shellgpu_factor_create(ctx, &F->F))) goto fail;
  if (hipMalloc((void **)&F->Q, v * (m + 1)) != hipSuccess || hipMalloc((void **)&F->MQ, v * (m + 1)) != hipSuccess ||
      hipMalloc((void **)&F->hv, sizeof(double) * (size_t)nb * (m + 1)) != hipSuccess ||
      hipMalloc((void **)&F->work, v) != hipSuccess || hipMalloc((void **)&F->ph, v) != hipSuccess ||
      hipMalloc((void **)&F->dsig, sizeof(double) * nb) != hipSuccess) {
    rc = ferr(ctx, 2, "out of device memory (frequency analysis)");
    goto fail;
  }
  *out = F;
  return 0;
fail:
  frequency_free(F);
  delete F;
  return rc;
}

void shellgpu_frequency_destroy(shellgpu_frequency *F) {
  if (!F) return;
  frequency_free(F);
  delete F;
}

int shellgpu_frequency_solve(shellgpu_frequency *F) {
  if (!F) return 1;
  shellgpu_ctx *ctx = F->ctx;
  const int nb = ctx->nbatch, nev = F->nev;
  int rc = 0;
  LanczosOut lo;
  F->solved = false;
  F->sig.assign(nb, F->sigma);
  // TACSFrequencyAnalysis::solve (src/TACSBuckling.cpp:810-893): K and M at
  // zero state, A = K - sigma M with BCs, factor, SEP::solve
  if ((rc = shellgpu_assemble_mattype(ctx, SHELLGPU_STIFFNESS, nullptr, F->K))) return rc;
  if ((rc = shellgpu_assemble_mattype(ctx, SHELLGPU_MASS, nullptr, F->M))) return rc;
  for (int attempt = 0;; attempt++) {
    std::vector<double> ms(nb);
    for (int d = 0; d < nb; d++) ms[d] = -F->sig[d];
    HIPCALL(ctx, hipMemcpyAsync(F->dsig, ms.data(), sizeof(double) * nb, hipMemcpyHostToDevice, ctx->stream));
    {
      const long nv = 36L * ctx->nnzb;
      hipLaunchKernelGGL(shifted_kernel, dim3((unsigned)((nv + 255) / 256), nb), dim3(256), 0, ctx->stream, nv,
                         F->K->vals, F->M->vals, F->dsig, F->A->vals);
      F->A->same = false;
      HIPCALL(ctx, hipGetLastError());
    }
    HIPCALL(ctx, hipStreamSynchronize(ctx->stream));  // ms is released next iteration
    if ((rc = shellgpu_mat_apply_bcs(F->A, 1.0))) return rc;
    rc = shellgpu_factor_compute(F->F, F->A);
    if (rc == 0) break;
    if (rc != 5 || attempt >= 12) return rc;
    // A indefinite: the shift lies above the first eigenvalue of a design
    std::vector<int> info(nb);
    HIPCALL(ctx, hipMemcpy(info.data(), F->F->info, sizeof(int) * nb, hipMemcpyDeviceToHost));
    for (int d = 0; d < nb; d++) {
      if (info[d] == 0) continue;
      if (F->sig[d] == 0.0) return ferr(ctx, 5, "stiffness matrix is not positive definite");
      F->sig[d] = attempt < 10 ? 0.5 * F->sig[d] : 0.0;
    }
    ctx->err.clear();
  }
  if ((rc = run_lanczos(ctx, F->F, F->M, nullptr, 2, F->sig, nev, F->m, F->tol, F->Q, F->MQ, F->hv, F->work, lo)))
    return rc;
  F->steps = lo.steps;
  F->lam.assign(nb, {});
  F->err.assign(nb, {});
  F->ritz.assign(nb, {});
  F->ns.assign(nb, 0);
  const int ns = lo.steps;
  if (ns < nev) return ferr(ctx, 8, "Lanczos basis smaller than the number of eigenvalues");
  std::vector<std::vector<double>> wv, zv, rv;
  if ((rc = proj_eigs(ctx, lo, ns, wv, zv, rv))) return rc;
  for (int d = 0; d < nb; d++) {
    const std::vector<double> &w = wv[d], &z = zv[d];
    const int ns = (int)w.size();  // this design's basis size
    if (ns < nev) return ferr(ctx, 8, "Lanczos basis smaller than the number of eigenvalues");
    std::vector<int> idx(ns);
    for (int i = 0; i < ns; i++) idx[i] = i;
    const double sd = F->sig[d];
    // EPGeneralizedShiftInvert::convertEigenvalue (src/bpmat/GSEP.cpp:184-186)
    auto lam_of = [&](int i) { return 1.0 / w[i] + sd; };
    // sortEigenvalues, SMALLEST_MAGNITUDE (src/bpmat/GSEP.cpp:693-748)
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return std::fabs(lam_of(a)) < std::fabs(lam_of(b)); });
    F->ns[d] = ns;
    for (int j = 0; j < nev; j++) {
      const int i = idx[j];
      F->lam[d].push_back(lam_of(i));
      F->err[d].push_back(rv[d][i] / std::fabs(w[i]));
      F->ritz[d].insert(F->ritz[d].end(), z.begin() + (size_t)i * ns, z.begin() + (size_t)(i + 1) * ns);
    }
  }
  F->solved = true;
  return 0;
}

int shellgpu_frequency_extract(shellgpu_frequency *F, int idx, double *lambda, double *phi, double *error) {
  if (!F) return 1;
  shellgpu_ctx *ctx = F->ctx;
  if (!F->solved) return ferr(ctx, 9, "frequency problem not solved");
  if (idx < 0 || idx >= F->nev) return ferr(ctx, 1, "eigenvalue index out of range");
  for (int d = 0; d < ctx->nbatch; d++) {
    if (lambda) lambda[d] = F->lam[d][idx];
    if (error) error[d] = F->err[d][idx];
  }
  if (phi) return ritz_vector(ctx, F->Q, F->ns, F->ritz, idx, F->hv, phi);
  return 0;
}

int shellgpu_frequency_info(shellgpu_frequency *F, int *nsteps, double *sigma_used) {
  if (!F) return 1;
  shellgpu_ctx *ctx = F->ctx;
  if (!F->solved) return ferr(ctx, 9, "frequency problem not solved");
  if (nsteps) *nsteps = F->steps;
  if (sigma_used)
    for (int d = 0; d < ctx->nbatch; d++) sigma_used[d] = F->sig[d];
  return 0;
}

int shellgpu_frequency_add_eigen_sens(shellgpu_frequency *F, const double *coef, double *dfdt, double *dfdX) {
  if (!F || !coef) return 1;
  shellgpu_ctx *ctx = F->ctx;
  if (!F->solved) return ferr(ctx, 9, "frequency problem not solved");
  if (!dfdt && !dfdX) return 0;
  const int nb = ctx->nbatch, n = 6 * ctx->nnodes, nev = F->nev;
  int rc = 0;
  VecOps vo;
  if ((rc = vo.init(ctx, n, nb))) return rc;
  std::vector<double> hc(4 * nb, 0.0);
  for (int j = 0; j < nev; j++) {
    bool any = false;
    for (int d = 0; d < nb; d++) any = any || coef[(size_t)d * nev + j] != 0.0;
    if (!any) continue;
    // d lambda = phi^T (K' - lambda M') phi / (phi^T M phi)
    if ((rc = ritz_vector(ctx, F->Q, F->ns, F->ritz, j, F->hv, F->ph))) return rc;
    if ((rc = shellgpu_mat_mult(F->M, F->ph, F->work))) return rc;
    std::vector<double> pmp;
    if ((rc = vo.dot(F->ph, F->work, pmp))) return rc;
    if (dfdX) {
      // coordinates: (c, 0, 0, -c lambda) per design, K and M in one pass
      for (int d = 0; d < nb; d++) {
        const double f = coef[(size_t)d * nev + j] / pmp[d];
        hc[4 * d] = f;
        hc[4 * d + 1] = hc[4 * d + 2] = 0.0;
        hc[4 * d + 3] = -f * F->lam[d][j];
      }
      HIPCALL(ctx, hipMemcpyAsync(ctx->d_coef, hc.data(), sizeof(double) * 4 * nb, hipMemcpyHostToDevice,
                                  ctx->stream));
      if ((rc = run_xpt(ctx, F->ph, nullptr, nullptr, ctx->d_coef, 0, kXptK | kXptM, dfdX))) return rc;
      HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (!dfdt) continue;
    // stiffness term: (c, 0) pairs of the thickness-adjoint kernel
    for (int d = 0; d < nb; d++) {
      hc[2 * d] = coef[(size_t)d * nev + j] / pmp[d];
      hc[2 * d + 1] = 0.0;
    }
    HIPCALL(ctx, hipMemcpyAsync(ctx->d_coef, hc.data(), sizeof(double) * 2 * nb, hipMemcpyHostToDevice, ctx->stream));
    {
      ElemArgs a = base_args(ctx, F->ph);
      a.u2 = nullptr;
      a.res = nullptr;
      a.coef = ctx->d_coef;
      a.dv = dfdt;
      if ((rc = run_colored(ctx, kModeAdjDV, a))) return rc;
    }
    HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
    // mass term: one coefficient per design
    for (int d = 0; d < nb; d++) hc[d] = -coef[(size_t)d * nev + j] * F->lam[d][j] / pmp[d];
    HIPCALL(ctx, hipMemcpyAsync(ctx->d_coef, hc.data(), sizeof(double) * nb, hipMemcpyHostToDevice, ctx->stream));
    {
      ElemArgs a = base_args(ctx, F->ph);
      a.coef = ctx->d_coef;
      a.dv = dfdt;
      if ((rc = run_colored(ctx, kModeMassDV, a))) return rc;
    }
    HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

int shellgpu_ks_aggregate(int n, double rho, const double *f, double *ks, double *dks) {
  if (n <= 0 || !f || !ks || !(rho > 0.0)) return 1;
  double fmax = f[0];
  for (int i = 1; i < n; i++) fmax = std::max(fmax, f[i]);
  double sum = 0.0;
  for (int i = 0; i < n; i++) {
    const double v = std::exp(rho * (f[i] - fmax));
    sum += v;
    if (dks) dks[i] = v;
  }
  if (dks)
    for (int i = 0; i < n; i++) dks[i] /= sum;
  *ks = fmax + std::log(sum) / rho;
  return 0;
}

static int newton_body(shellgpu_ctx *ctx, shellgpu_factor *Fext, const double *force, int nsteps, int max_iters,
                       double rtol, double *u, int *iters, double *relres);

int shellgpu_newton_solve(shellgpu_ctx *ctx, const double *force, int nsteps, int max_iters, double rtol, double *u,
                          int *iters, double *relres) {
  return newton_body(ctx, nullptr, force, nsteps, max_iters, rtol, u, iters, relres);
}

int shellgpu_newton_solve_factor(shellgpu_ctx *ctx, shellgpu_factor *F, const double *force, int nsteps,
                                 int max_iters, double rtol, double *u, int *iters, double *relres) {
  if (!F || F->ctx != ctx) return ferr(ctx, 1, "a factor of this context is required");
  if (F->nbatch != ctx->nbatch) return ferr(ctx, 1, "factor batch size does not match the context's designs");
  return newton_body(ctx, F, force, nsteps, max_iters, rtol, u, iters, relres);
}

static int newton_body(shellgpu_ctx *ctx, shellgpu_factor *Fext, const double *force, int nsteps, int max_iters,
                       double rtol, double *u, int *iters, double *relres) {
  if (!ctx || !force || !u || nsteps <= 0 || max_iters <= 0 || !(rtol > 0.0))
    return ferr(ctx, 1, "invalid Newton arguments");
  if (!ctx->d_thick) return ferr(ctx, 3, "designs not set (shellgpu_set_designs)");
  const int nb = ctx->nbatch, n = 6 * ctx->nnodes;
  int rc = 0, total = 0;
  shellgpu_mat *A = nullptr;
  shellgpu_factor *F = nullptr;
  double *R = nullptr, *f0 = nullptr, *du = nullptr;
  VecOps vo;
  std::vector<double> fn, rn(nb, 0.0);
  bool factored = false;
  if ((rc = vo.init(ctx, n, nb))) return rc;
  if ((rc = ensure_blas(ctx))) return rc;
  if ((rc = shellgpu_mat_create(ctx, nb, &A))) goto out;
  if (Fext)
    F = Fext;
  else if ((rc = shellgpu_factor_create(ctx, &F)))
    goto out;
  if (hipMalloc((void **)&R, sizeof(double) * (size_t)n * nb) != hipSuccess ||
      hipMalloc((void **)&f0, sizeof(double) * (size_t)n * nb) != hipSuccess ||
      hipMalloc((void **)&du, sizeof(double) * (size_t)n * nb) != hipSuccess) {
    rc = ferr(ctx, 2, "out of device memory (Newton vectors)");
    goto out;
  }
  if (hipMemcpyAsync(f0, force, sizeof(double) * (size_t)n * nb, hipMemcpyDeviceToDevice, ctx->stream) !=
      hipSuccess) {
    rc = ferr(ctx, 2, "force copy failed");
    goto out;
  }
  if ((rc = shellgpu_vec_zero_bcs(ctx, f0))) goto out;
  if ((rc = vo.dot(f0, f0, fn))) goto out;
  for (double &v : fn) v = v > 0.0 ? std::sqrt(v) : 1.0;
  for (int k = 1; k <= nsteps; k++) {
    const double s = (double)k / nsteps;
    for (int it = 0;; it++) {
      // R(u) - s f with BC rows u - ubar; tangent of the nonlinear model
      if ((rc = shellgpu_assemble_jacobian(ctx, SHELLGPU_NONLINEAR, 1.0, 0.0, 0.0, u, nullptr, nullptr, R, A))) goto out;
      {
        std::vector<double> ms(nb, -s);
        if ((rc = vo.axpy(ms, f0, R))) goto out;
      }
      std::vector<double> r2;
      if ((rc = vo.dot(R, R, r2))) goto out;
      bool conv = true;
      for (int d = 0; d < nb; d++) {
        rn[d] = std::sqrt(r2[d]) / fn[d];
        conv = conv && rn[d] <= rtol;
      }
      if (conv) break;
      if (it >= max_iters) {
        rc = ferr(ctx, 11, "Newton did not converge within max_iters");
        goto out;
      }
      // Newton step K_t du = R by GMRES, preconditioned by the last factored
      // tangent; refactor at the current state when that stalls
      int its = 0;
      double res = 0.0;
      if (!factored) {
        if ((rc = shellgpu_factor_compute(F, A))) goto out;
        factored = true;
      }
      if ((rc = shellgpu_gmres(ctx, A, F, R, du, 20, 2, 1e-12, 1e-300, &its, &res))) goto out;
      if (!(res <= 1e-10)) {
        if ((rc = shellgpu_factor_compute(F, A))) goto out;
        if ((rc = shellgpu_gmres(ctx, A, F, R, du, 20, 5, 1e-12, 1e-300, &its, &res))) goto out;
      }
      {
        std::vector<double> m1(nb, -1.0);
        if ((rc = vo.axpy(m1, du, u))) goto out;
      }
      total++;
    }
  }
  if (iters) *iters = total;
  if (relres)
    for (int d = 0; d < nb; d++) relres[d] = rn[d];
out:
  if (A) shellgpu_mat_destroy(A);
  if (F && F != Fext) shellgpu_factor_destroy(F);
  double *bufs[] = {R, f0, du};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  return rc;
}

int shellgpu_vec_axpby(shellgpu_ctx *ctx, size_t n, double alpha, double beta, const double *x, double *y) {
  if (!ctx || !x || !y || n > 0x7fffffff) return ferr(ctx, 1, "invalid vector arguments");
  int rc = ensure_blas(ctx);
  if (rc) return rc;
  rocblas_handle h = (rocblas_handle)ctx->blas;
  if (beta != 1.0) BLASCALL(ctx, rocblas_dscal(h, (rocblas_int)n, &beta, y, 1));
  BLASCALL(ctx, rocblas_daxpy(h, (rocblas_int)n, &alpha, x, 1, y, 1));
  return 0;
}

int shellgpu_vec_scale(shellgpu_ctx *ctx, size_t n, double alpha, double *x) {
  if (!ctx || !x || n > 0x7fffffff) return ferr(ctx, 1, "invalid vector arguments");
  int rc = ensure_blas(ctx);
  if (rc) return rc;
  BLASCALL(ctx, rocblas_dscal((rocblas_handle)ctx->blas, (rocblas_int)n, &alpha, x, 1));
  return 0;
}

int shellgpu_vec_dot(shellgpu_ctx *ctx, size_t n, const double *x, const double *y, double *out) {
  if (!ctx || !x || !y || !out || n > 0x7fffffff) return ferr(ctx, 1, "invalid vector arguments");
  int rc = ensure_blas(ctx);
  if (rc) return rc;
  // host pointer mode: the call returns when the result is in *out
  BLASCALL(ctx, rocblas_ddot((rocblas_handle)ctx->blas, (rocblas_int)n, x, 1, y, 1, out));
  return 0;
}

}  // extern "C"
