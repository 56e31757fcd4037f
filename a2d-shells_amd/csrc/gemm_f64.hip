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
// Work split: a workgroup of 4 wavefronts computes a T x T tile of C (T =
// 128 for large problems, 64 where 128-tiles would leave the GPU short of
// workgroups), each wavefront a T/2 x T/2 quarter as (T/32)^2 blocks of
// v_mfma_f64_16x16x4f64 (16 x 16 x 4 per instruction, 64 cycles).  K
// advances 16 at a time through LDS: every thread stages T/16 doubles of the
// A tile (T x 16) and of the B tile (16 x T), double-buffered so the next
// tile's global loads are in flight during the current tile's (T/32)^2 x 4
// matrix instructions per wavefront.  Tiles are stored k-row by k-row with an
// odd row length (T + 1 doubles): the staging stores (along a row or down a
// column) and the operand reads are free of LDS bank conflicts.
//
// Operand layout of v_mfma_f64_16x16x4f64 (one 64-lane wavefront), as
// measured on gfx950: lane l supplies X[l % 16][l / 16] and Y[l / 16][l % 16]
// of the 16 x 4 and 4 x 16 operands of D = X Y, and holds D[4 i + l / 16][l % 16],
// i = 0..3.  The kernel computes the transposed block, X from B and Y from A
// (D = (A B)^T), so lanes 0-15 of a quarter hold 16 consecutive rows of one
// column of C and the stores coalesce.
#include <algorithm>

#include "shellgpu_internal.h"

namespace shellgpu {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
// global-memory views of the operands: the compiler then issues global
// loads (a generic pointer read from a struct or a pointer array gives flat
// loads, which also count against the LDS wait counter and so serialise
// the prefetch of the next tile with the matrix instructions)
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) double gdouble;

constexpr int kTK = 16;
constexpr int kGroup = 8;  // tile rows per group of the work order

// WT: 16 x 16 blocks per wavefront and direction (tile T = 32 WT)
template <bool TA, bool TB, int WT>
__global__ __launch_bounds__(256) void gemm_f64_kernel(GemmArgs g) {
  constexpr int kT = 32 * WT;      // workgroup tile
  constexpr int kRow = kT + 1;     // doubles per staged k-row: odd, so a column walk hits every bank
  constexpr int kLd = kT / 16;     // staged doubles per thread and operand
  constexpr int kStep = 256 / kT;  // k-rows per staging pass (lane along m or n)
  __shared__ double As[2][kTK][kRow];
  __shared__ double Bs[2][kTK][kRow];
  // Work order.  The dispatcher hands consecutive workgroups to the 8 XCDs
  // in turn, each XCD with its own L2: give XCD x the x-th contiguous eighth
  // of the work order, which runs batch entry by batch entry and within one
  // in groups of kGroup tile rows, column by column, so the workgroups
  // resident on one XCD share a few panels of A and B in its L2.
  const int tM = (g.M + kT - 1) / kT, tN = (g.N + kT - 1) / kT;
  const unsigned G = gridDim.x, G8 = G & ~7u, gid = blockIdx.x;
  const unsigned w = gid < G8 ? (gid & 7u) * (G8 >> 3) + (gid >> 3) : gid;
  const long bz = w / ((unsigned)tM * tN);
  const int wt = (int)(w - bz * tM * tN);
  const int gm0 = wt / (kGroup * tN) * kGroup, gsz = min(tM - gm0, kGroup), tg = wt - gm0 * tN;
  const int tile_m = gm0 + tg % gsz, tile_n = tg / gsz;
  gcdouble *A = (gcdouble *)(g.Ap ? g.Ap[bz] : g.A + bz * g.sA);
  gcdouble *B = (gcdouble *)(g.Bp ? g.Bp[bz] : g.B + bz * g.sB);
  gdouble *C = (gdouble *)(g.Cp ? g.Cp[bz] : g.C + bz * g.sC);
  const int M = g.M, N = g.N, K = g.K;
  const int m0 = tile_m * kT, n0 = tile_n * kT;
  if (g.lower_c && n0 >= m0 + kT) return;  // tile strictly above the diagonal
  // K range holding the tile's nonzero products of triangular operands
  int kbeg = 0, kend = K;
  if (g.tri_a == kGemmLower) kend = min(kend, m0 + kT);
  if (g.tri_a == kGemmUpper) kbeg = max(kbeg, m0);
  if (g.tri_b == kGemmLower) kbeg = max(kbeg, n0);
  if (g.tri_b == kGemmUpper) kend = min(kend, n0 + kT);
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int wm = (wave & 1) * (kT / 2), wn = (wave >> 1) * (kT / 2);

  // staging assignment: A tile (m 0..T-1, k 0..15), B tile (k 0..15, n 0..T-1),
  // consecutive lanes along the operand's contiguous direction in memory
  // (coalesced loads, and with the odd LDS row length conflict-free stores):
  // N-op A / T-op B: lane -> m (n) = t % T, k = t / T + kStep j;
  // T-op A / N-op B: lane -> k = t % 16, m (n) = t / 16 + 16 j
  const int a_r = TA ? (t >> 4) : (t % kT), a_k = TA ? (t & 15) : (t / kT);
  const int b_c = TB ? (t % kT) : (t >> 4), b_k = TB ? (t / kT) : (t & 15);
  double ra[kLd], rb[kLd];

  const int tra = g.tri_a, trb = g.tri_b;
  // every load reads an element of the operand (indices clamped into the
  // matrix; K >= 1 wherever the K loop runs): no branches.  Out-of-range
  // and out-of-triangle elements are zeroed when the tile is stored to LDS,
  // so the loaded values are first used after the current tile's matrix
  // instructions and the loads overlap them.
  unsigned keep = 0;
  auto load = [&](int k0) {
    keep = 0;
#pragma unroll
    for (int j = 0; j < kLd; j++) {
      // A element (m, k) of op(A)
      const int m = m0 + a_r + (TA ? 16 * j : 0), k = k0 + a_k + (TA ? 0 : kStep * j);
      const bool ina = m < M && k < K && !(tra == kGemmLower && k > m) && !(tra == kGemmUpper && k < m);
      const int mc = min(m, M - 1), kc = min(k, K - 1);
      ra[j] = TA ? A[kc + (long)mc * g.lda] : A[mc + (long)kc * g.lda];
      // B element (k, n) of op(B)
      const int n = n0 + b_c + (TB ? 0 : 16 * j), kb = k0 + b_k + (TB ? kStep * j : 0);
      const bool inb = n < N && kb < K && !(trb == kGemmLower && kb < n) && !(trb == kGemmUpper && kb > n);
      const int nc = min(n, N - 1), kbc = min(kb, K - 1);
      rb[j] = TB ? B[nc + (long)kbc * g.ldb] : B[kbc + (long)nc * g.ldb];
      keep |= (ina ? 1u : 0u) << j | (inb ? 1u : 0u) << (16 + j);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < kLd; j++) {
      const double va = (keep >> j) & 1u ? ra[j] : 0.0, vb = (keep >> (16 + j)) & 1u ? rb[j] : 0.0;
      if (TA)
        As[buf][a_k][a_r + 16 * j] = va;
      else
        As[buf][a_k + kStep * j][a_r] = va;
      if (TB)
        Bs[buf][b_k + kStep * j][b_c] = vb;
      else
        Bs[buf][b_k][b_c + 16 * j] = vb;
    }
  };

  v4d acc[WT][WT];
#pragma unroll
  for (int i = 0; i < WT; i++)
#pragma unroll
    for (int j = 0; j < WT; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  // counted loop with an unconditional prefetch of the next tile (clamped
  // and zero-masked past the end): no branches in the loop body
  const int nk = kbeg < kend ? (kend - kbeg + kTK - 1) / kTK : 0;
  if (nk > 0) {
    load(kbeg);
    store(0);
  }
  __syncthreads();
  const int lr = lane & 15, lk = lane >> 4;
  for (int it = 0; it < nk; it++) {
    const int cur = it & 1;
    load(kbeg + (it + 1) * kTK);
#pragma unroll
    for (int kk = 0; kk < kTK; kk += 4) {
      double a[WT], b[WT];
#pragma unroll
      for (int i = 0; i < WT; i++) {
        a[i] = As[cur][kk + lk][wm + 16 * i + lr];
        b[i] = Bs[cur][kk + lk][wn + 16 * i + lr];
      }
#pragma unroll
      for (int i = 0; i < WT; i++)
#pragma unroll
        for (int j = 0; j < WT; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
    }
    store(cur ^ 1);
    __syncthreads();
  }

  // C = alpha acc + beta C
#pragma unroll
  for (int i = 0; i < WT; i++) {
    const int m = m0 + wm + 16 * i + lr;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < WT; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int n = n0 + wn + 16 * j + 4 * e + lk;
        if (n >= N) continue;
        gdouble *c = C + m + (long)n * g.ldc;
        const double v = g.alpha * acc[i][j][e];
        *c = g.beta == 0.0 ? v : v + g.beta * *c;
      }
  }
}

// Products with at most 16 columns, op(B) = B (a block of vectors): C =
// alpha op(A) B + beta C in 64 x 16 workgroup tiles with the four
// wavefronts stacked along M (16 rows each, one v_mfma_f64_16x16x4f64
// block per k-step), where the square tiles above would spend three
// quarters of their matrix instructions on padding columns.  Such a product
// is bound by the reads of A: the A tile streams through LDS 16 k-rows at a
// time, double-buffered, with the global loads of the next two k-steps in
// flight, and triangular operands skip their zero K range.
// With ksplit > 1 (a launch of few tiles) each tile's K range is cut into
// ksplit chunks of whole k-steps, one workgroup each, whose unscaled sums go
// to the workspace and are added in chunk order by gemm_narrow_reduce_kernel.
template <bool TA>
__global__ __launch_bounds__(256) void gemm_narrow_kernel(GemmArgs g, int ksplit) {
  constexpr int kTM = 64, kRowA = kTM + 1, kRowB = 17;
  __shared__ double As[2][kTK][kRowA];
  __shared__ double Bs[2][kTK][kRowB];
  const int tM = (g.M + kTM - 1) / kTM;
  const long tix = blockIdx.x / ksplit;
  const int ks = (int)(blockIdx.x - tix * ksplit);
  const long bz = tix / tM;
  const int tile_m = (int)(tix - bz * tM);
  gcdouble *A = (gcdouble *)(g.Ap ? g.Ap[bz] : g.A + bz * g.sA);
  gcdouble *B = (gcdouble *)(g.Bp ? g.Bp[bz] : g.B + bz * g.sB);
  gdouble *C = (gdouble *)(g.Cp ? g.Cp[bz] : g.C + bz * g.sC);
  const int M = g.M, N = g.N, K = g.K, tra = g.tri_a;
  const int m0 = tile_m * kTM;
  int kbeg = 0, kend = K;
  if (tra == kGemmLower) kend = min(kend, m0 + kTM);
  if (tra == kGemmUpper) kbeg = max(kbeg, m0);
  if (ksplit > 1 && kbeg < kend) {
    const int steps = (kend - kbeg + kTK - 1) / kTK, per = (steps + ksplit - 1) / ksplit;
    const int b0 = kbeg + ks * per * kTK;
    kend = min(kend, b0 + per * kTK);
    kbeg = b0;
  }
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  // staging, lanes along each operand's contiguous direction: A tile
  // (64 m x 16 k, 4 per thread): N-op m = t % 64, k = t / 64 + 4 j; T-op
  // k = t % 16, m = t / 16 + 16 j; B tile (16 k x 16 n): k = t % 16, n = t / 16.
  // Addresses are clamped into the operands, so the loads past the K range
  // (the prefetch beyond the last step) read valid memory and are dropped.
  const int a_r = TA ? (t >> 4) : (t & 63), a_k = TA ? (t & 15) : (t >> 6);
  const int b_k = t & 15, b_n = t >> 4;
  struct Stage {
    double a[4], b;
    unsigned keep;
  };
  auto load = [&](Stage &r, int k0) {
    r.keep = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int m = m0 + a_r + (TA ? 16 * j : 0), k = k0 + a_k + (TA ? 0 : 4 * j);
      const bool ina = m < M && k < K && !(tra == kGemmLower && k > m) && !(tra == kGemmUpper && k < m);
      const int mc = min(m, M - 1), kc = min(k, K - 1);
      r.a[j] = TA ? A[kc + (long)mc * g.lda] : A[mc + (long)kc * g.lda];
      r.keep |= (ina ? 1u : 0u) << j;
    }
    const int kb = k0 + b_k;
    r.b = B[min(kb, K - 1) + (long)min(b_n, N - 1) * g.ldb];
    r.keep |= (b_n < N && kb < K ? 1u : 0u) << 4;
  };
  auto store = [&](const Stage &r, int buf) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const double va = (r.keep >> j) & 1u ? r.a[j] : 0.0;
      if (TA)
        As[buf][a_k][a_r + 16 * j] = va;
      else
        As[buf][a_k + 4 * j][a_r] = va;
    }
    Bs[buf][b_k][b_n] = (r.keep >> 4) & 1u ? r.b : 0.0;
  };
  const int lr = lane & 15, lk = lane >> 4;
  v4d acc = v4d{0.0, 0.0, 0.0, 0.0};
  auto step = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < kTK; kk += 4)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Bs[buf][kk + lk][lr], As[buf][kk + lk][16 * wave + lr], acc, 0, 0,
                                                 0);
  };
  const int nk = kbeg < kend ? (kend - kbeg + kTK - 1) / kTK : 0;
  // two register stages: the loads of the next two k-steps are in flight
  // during a step's products (the loop is unrolled by two so each stage has
  // fixed registers), twice the bytes in flight per workgroup of a single
  // stage at the same LDS footprint: 5.3 TB/s against 5.0 on 102 levels of
  // 2,256 x 2,256 times 16 vectors
  Stage r0, r1;
  if (nk > 0) {
    load(r0, kbeg);
    load(r1, kbeg + kTK);
    store(r0, 0);
  }
  __syncthreads();
  for (int it = 0; it < nk; it += 2) {
    // buffer 0 holds step it, r1 step it + 1
    load(r0, kbeg + (it + 2) * kTK);
    step(0);
    store(r1, 1);
    __syncthreads();
    if (it + 1 >= nk) break;
    load(r1, kbeg + (it + 3) * kTK);
    step(1);
    store(r0, 0);
    __syncthreads();
  }
  // the transposed block: lane l holds C[16 wave + l % 16][4 e + l / 16]
  const int m = m0 + 16 * wave + lr;
  if (m >= M) return;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int n = 4 * e + lk;
    if (n >= N) continue;
    if (ksplit > 1) {
      g.ws[((bz * ksplit + ks) * N + n) * (long)M + m] = acc[e];
      continue;
    }
    gdouble *c = C + m + (long)n * g.ldc;
    const double v = g.alpha * acc[e];
    *c = g.beta == 0.0 ? v : v + g.beta * *c;
  }
}

// C = beta C + alpha (sum of the ksplit K chunks, in chunk order)
__global__ void gemm_narrow_reduce_kernel(GemmArgs g, int ksplit) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long MN = (long)g.M * g.N;
  if (i >= MN) return;
  const long bz = blockIdx.y;
  const int n = (int)(i / g.M), m = (int)(i - (long)n * g.M);
  double s = 0.0;
  for (int ks = 0; ks < ksplit; ks++) s += g.ws[(bz * ksplit + ks) * MN + i];
  double *C = g.Cp ? g.Cp[bz] : g.C + bz * g.sC;
  double *c = C + m + (long)n * g.ldc;
  const double v = g.alpha * s;
  *c = g.beta == 0.0 ? v : v + g.beta * *c;
}

template <int WT>
hipError_t launch_tiles(bool ta, bool tb, const GemmArgs &g, long tiles, hipStream_t stream) {
  const dim3 grid((unsigned)tiles);
  if (!ta && !tb)
    hipLaunchKernelGGL((gemm_f64_kernel<false, false, WT>), grid, dim3(256), 0, stream, g);
  else if (!ta && tb)
    hipLaunchKernelGGL((gemm_f64_kernel<false, true, WT>), grid, dim3(256), 0, stream, g);
  else if (ta && !tb)
    hipLaunchKernelGGL((gemm_f64_kernel<true, false, WT>), grid, dim3(256), 0, stream, g);
  else
    hipLaunchKernelGGL((gemm_f64_kernel<true, true, WT>), grid, dim3(256), 0, stream, g);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_gemm_f64(bool ta, bool tb, const GemmArgs &g, int batch, hipStream_t stream) {
  if (g.M <= 0 || g.N <= 0 || batch <= 0) return hipSuccess;
  if (g.N <= 16 && !tb && g.tri_b == kGemmFull && !g.lower_c) {
    const long tiles = (long)((g.M + 63) / 64) * batch;
    // split K when the tiles would not fill the GPU with workgroups (each
    // then runs the whole K chain): about 2,048 workgroups, chunks of at
    // least 4 k-steps, within the workspace
    int ksplit = 1;
    if (g.ws && tiles < 1024) {
      const long steps = (g.K + kTK - 1) / kTK;
      ksplit = (int)std::min<long>((2048 + tiles - 1) / tiles, std::max<long>(1, steps / 4));
      const long per = (long)g.M * g.N * batch;
      if (per > 0) ksplit = (int)std::min<long>(ksplit, g.ws_len / per);
      if (ksplit < 2) ksplit = 1;
    }
    if (tiles * ksplit > 0x7fffffffL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(tiles * ksplit));
    if (ta)
      hipLaunchKernelGGL(gemm_narrow_kernel<true>, grid, dim3(256), 0, stream, g, ksplit);
    else
      hipLaunchKernelGGL(gemm_narrow_kernel<false>, grid, dim3(256), 0, stream, g, ksplit);
    if (ksplit > 1) {
      const long MN = (long)g.M * g.N;
      hipLaunchKernelGGL(gemm_narrow_reduce_kernel, dim3((unsigned)((MN + 255) / 256), batch), dim3(256), 0, stream,
                         g, ksplit);
    }
    return hipGetLastError();
  }
  // 128-tiles while they give at least two workgroups per CU and both
  // dimensions exceed 64 (a 128-tile on a narrow product, such as a
  // product with 16 vectors, would compute mostly padding), else 64-tiles
  const long t128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128) * batch;
  const long t64 = (long)((g.M + 63) / 64) * ((g.N + 63) / 64) * batch;
  if (t64 > 0x7fffffffL) return hipErrorInvalidValue;
  if (t128 >= 512 && g.M > 64 && g.N > 64) return launch_tiles<4>(ta, tb, g, t128, stream);
  return launch_tiles<2>(ta, tb, g, t64, stream);
}

}  // namespace shellgpu
