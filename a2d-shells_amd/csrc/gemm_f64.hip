// gemm_f64.hip -- batched FP64 GEMM on the CDNA4 matrix cores for the level
// factorisations (the dense updates of the Schur-complement inverses and of
// block cyclic reduction):
//
//   C_b = alpha op(A_b) op(B_b) + beta C_b,   b = 0 .. batch-1,
//
// column-major with leading dimensions, op = identity or transpose, the
// matrices of entry b given by pointer arrays or by a base and a stride;
// op(A) or op(B) may be triangular (the K loop of a tile runs over the
// range where both are nonzero) and C may be computed on and below the
// diagonal only (the symmetric updates of a factorisation).
//
// Work split: a workgroup of 4 wavefronts computes a 64 x 64 tile of C, each
// wavefront a 32 x 32 quarter as 2 x 2 blocks of v_mfma_f64_16x16x4f64
// (16 x 16 x 4 per instruction).  K advances 16 at a time through LDS:
// every thread stages 4 doubles of the A tile (64 x 16) and 4 of the B tile
// (16 x 64), double-buffered so the next tile's global loads are in flight
// during the current tile's 16 matrix instructions per wavefront.  Tiles are
// stored k-row by k-row with 80 doubles per row: the four k-rows read by one
// matrix instruction then fall in distinct LDS banks.
//
// Operand layout of v_mfma_f64_16x16x4f64 (one 64-lane wavefront), as
// measured on gfx950: lane l supplies X[l % 16][l / 16] and Y[l / 16][l % 16]
// of the 16 x 4 and 4 x 16 operands of D = X Y, and holds D[4 i + l / 16][l % 16],
// i = 0..3.  The kernel computes the transposed block, X from B and Y from A
// (D = (A B)^T), so lanes 0-15 of a quarter hold 16 consecutive rows of one
// column of C and the stores coalesce.
#include "shellgpu_internal.h"

namespace shellgpu {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kTM = 64, kTN = 64, kTK = 16;
constexpr int kRow = 80;  // doubles per staged k-row

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_f64_kernel(GemmArgs g) {
  __shared__ double As[2][kTK][kRow];
  __shared__ double Bs[2][kTK][kRow];
  const long bz = blockIdx.z;
  const double *A = g.Ap ? g.Ap[bz] : g.A + bz * g.sA;
  const double *B = g.Bp ? g.Bp[bz] : g.B + bz * g.sB;
  double *C = g.Cp ? g.Cp[bz] : g.C + bz * g.sC;
  const int M = g.M, N = g.N, K = g.K;
  const int m0 = blockIdx.x * kTM, n0 = blockIdx.y * kTN;
  if (g.lower_c && n0 >= m0 + kTM) return;  // tile strictly above the diagonal
  // K range holding the tile's nonzero products of triangular operands
  int kbeg = 0, kend = K;
  if (g.tri_a == kGemmLower) kend = min(kend, m0 + kTM);
  if (g.tri_a == kGemmUpper) kbeg = max(kbeg, m0);
  if (g.tri_b == kGemmLower) kbeg = max(kbeg, n0);
  if (g.tri_b == kGemmUpper) kend = min(kend, n0 + kTN);
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;

  // staging assignment: A tile (m 0..63, k 0..15), B tile (k 0..15, n 0..63)
  // N-op A / T-op B: 4 consecutive m (n) at one k; T-op A / N-op B: 4
  // consecutive k at one m (n) -- the contiguous direction in memory
  const int a_r = TA ? (t >> 2) : ((t & 15) << 2), a_k = TA ? ((t & 3) << 2) : (t >> 4);
  const int b_c = TB ? ((t & 15) << 2) : (t >> 2), b_k = TB ? (t >> 4) : ((t & 3) << 2);
  double ra[4], rb[4];

  const int tra = g.tri_a, trb = g.tri_b;
  auto load = [&](int k0) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      // A element (m, k) of op(A)
      const int m = m0 + a_r + (TA ? 0 : j), k = k0 + a_k + (TA ? j : 0);
      const bool ina = m < M && k < K && !(tra == kGemmLower && k > m) && !(tra == kGemmUpper && k < m);
      ra[j] = ina ? (TA ? A[k + (long)m * g.lda] : A[m + (long)k * g.lda]) : 0.0;
      // B element (k, n) of op(B)
      const int n = n0 + b_c + (TB ? j : 0), kb = k0 + b_k + (TB ? 0 : j);
      const bool inb = n < N && kb < K && !(trb == kGemmLower && kb < n) && !(trb == kGemmUpper && kb > n);
      rb[j] = inb ? (TB ? B[n + (long)kb * g.ldb] : B[kb + (long)n * g.ldb]) : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (TA)
        As[buf][a_k + j][a_r] = ra[j];
      else
        As[buf][a_k][a_r + j] = ra[j];
      if (TB)
        Bs[buf][b_k][b_c + j] = rb[j];
      else
        Bs[buf][b_k + j][b_c] = rb[j];
    }
  };

  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  if (kbeg < kend) {
    load(kbeg);
    store(0);
  }
  __syncthreads();
  int cur = 0;
  const int lr = lane & 15, lk = lane >> 4;
  for (int k0 = kbeg; k0 < kend; k0 += kTK) {
    const bool more = k0 + kTK < kend;
    if (more) load(k0 + kTK);
#pragma unroll
    for (int kk = 0; kk < kTK; kk += 4) {
      double a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; i++) {
        a[i] = As[cur][kk + lk][wm + 16 * i + lr];
        b[i] = Bs[cur][kk + lk][wn + 16 * i + lr];
      }
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
    }
    if (more) store(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }

  // C = alpha acc + beta C
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const int m = m0 + wm + 16 * i + lr;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int n = n0 + wn + 16 * j + 4 * e + lk;
        if (n >= N) continue;
        double *c = C + m + (long)n * g.ldc;
        const double v = g.alpha * acc[i][j][e];
        *c = g.beta == 0.0 ? v : v + g.beta * *c;
      }
  }
}

}  // namespace

hipError_t launch_gemm_f64(bool ta, bool tb, const GemmArgs &g, int batch, hipStream_t stream) {
  if (g.M <= 0 || g.N <= 0 || batch <= 0) return hipSuccess;
  const dim3 grid((g.M + kTM - 1) / kTM, (g.N + kTN - 1) / kTN, batch);
  if (!ta && !tb)
    hipLaunchKernelGGL((gemm_f64_kernel<false, false>), grid, dim3(256), 0, stream, g);
  else if (!ta && tb)
    hipLaunchKernelGGL((gemm_f64_kernel<false, true>), grid, dim3(256), 0, stream, g);
  else if (ta && !tb)
    hipLaunchKernelGGL((gemm_f64_kernel<true, false>), grid, dim3(256), 0, stream, g);
  else
    hipLaunchKernelGGL((gemm_f64_kernel<true, true>), grid, dim3(256), 0, stream, g);
  return hipGetLastError();
}

}  // namespace shellgpu
