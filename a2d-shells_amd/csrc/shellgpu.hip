// shellgpu.hip -- C-ABI layer: context, mesh / pattern / colouring set-up,
// matrix objects, assembly launches, boundary conditions and the 6x6-block
// SpMV.  Declarations and reference citations are in include/shellgpu.h.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>

#include "shellgpu.h"
#include "shellgpu_ctx.h"

using namespace shellgpu;

// Workspace of the entry points that allocate nothing
// (shellgpu_gmres_noalloc): a fixed 64 MiB array in device memory.  Every
// GPU the library runs on holds its own copy, at its own address, until the
// process ends.  Host code must not use the name `scratch` as a device
// address: shellgpu_create maps it once per device with
// hipGetSymbolAddress and every context keeps that pointer.  All calls on
// one device share the array, so calls that overlap on one GPU race.
constexpr size_t kScratchDoubles = size_t(8) << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 256;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// the current device's address of `scratch` (device must be current)
hipError_t map_scratch(int device, double **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

int set_err(shellgpu_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

#define HIPCALL(ctx, call)                                                        \
  do {                                                                            \
    hipError_t _e = (call);                                                       \
    if (_e != hipSuccess)                                                         \
      return set_err(ctx, 2, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

template <typename T>
int upload(shellgpu_ctx *ctx, T **dptr, const T *h, size_t n) {
  if (*dptr) {
    (void)hipFree(*dptr);
    *dptr = nullptr;
  }
  if (n == 0) return 0;
  HIPCALL(ctx, hipMalloc((void **)dptr, n * sizeof(T)));
  HIPCALL(ctx, hipMemcpyAsync(*dptr, h, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// BCSRMat::zeroRow for every BC node (BCSRMat.cpp:2005-2034)
__global__ void bsr_bc_kernel(const int *rowp, const int *cols, double *vals, long vstride,
                              const int *bc_nodes, const int *bc_mask, double diag) {
  const int b = blockIdx.x;
  vals += (long)blockIdx.y * vstride;
  const int row = bc_nodes[b], mask = bc_mask[b];
  const int p0 = rowp[row], p1 = rowp[row + 1];
  for (int idx = threadIdx.x; idx < 36 * (p1 - p0); idx += blockDim.x) {
    const int p = p0 + idx / 36, e = idx % 36, r = e / 6, c = e % 6;
    if (mask & (1 << r)) vals[36L * p + e] = (cols[p] == row && r == c) ? diag : 0.0;
  }
}

// TACSBVec::applyBCs / setBCs (TACSBVec.cpp:546-625)
// mode 0: x[bc] = 0; 1: x[bc] = ubar; 2: x[bc] = u[bc] - ubar (u may be null)
__global__ void vec_bc_kernel(double *x, const double *u, long stride, int nbc,
                              const int *bc_nodes, const int *bc_mask, const double *bc_vals,
                              int mode) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 6 * nbc) return;
  const int b = i / 6, k = i % 6;
  if (!(bc_mask[b] & (1 << k))) return;
  const long off = (long)blockIdx.y * stride + 6L * bc_nodes[b] + k;
  const double ub = bc_vals ? bc_vals[6 * b + k] : 0.0;
  if (mode == 0) {
    x[off] = 0.0;
  } else if (mode == 1) {
    x[off] = ub;
  } else {
    x[off] = (u ? u[off] : 0.0) - ub;
  }
}

// y = A x, 6x6 blocks; one lane per scalar row (BCSRMatVecMult6,
// BCSRMatMult6.cpp:82-125)
template <bool kVec>
__global__ __launch_bounds__(192) void bsr_spmv6_kernel(int nrows, const int *__restrict__ rowp,
                                                        const int *__restrict__ cols,
                                                        const double *__restrict__ vals, long vstride,
                                                        const double *__restrict__ x,
                                                        double *__restrict__ y, long xstride) {
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(tid / 6), r = (int)(tid % 6);
  if (row >= nrows) return;
  vals += (long)blockIdx.y * vstride;
  x += (long)blockIdx.y * xstride;
  y += (long)blockIdx.y * xstride;
  double acc = 0.0;
  int p = rowp[row];
  const int p1 = rowp[row + 1];
  if constexpr (kVec) {
    // the lane's 48-byte rows of up to kAhead blocks are all requested
    // before the first product (16-byte loads: blocks are 288 B and node
    // segments 48 B, so both are 16-byte aligned when x is), which keeps
    // enough bytes in flight for HBM: 68 % of its peak on 64 matrices of the
    // 100 x 200 cylinder against 52 % loading block by block.  The sum runs
    // in block order either way.
    constexpr int kAhead = 9;  // the row length of a structured quad mesh
    double2 a[kAhead][3];
    int c[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; u++) {
      if (p + u < p1) {
        const double2 *ap = reinterpret_cast<const double2 *>(vals + 36L * (p + u) + 6 * r);
        a[u][0] = ap[0];
        a[u][1] = ap[1];
        a[u][2] = ap[2];
        c[u] = cols[p + u];
      }
    }
#pragma unroll
    for (int u = 0; u < kAhead; u++) {
      if (p + u < p1) {
        const double2 *xp = reinterpret_cast<const double2 *>(x + 6L * c[u]);
        const double2 x0 = xp[0], x1 = xp[1], x2 = xp[2];
        acc += a[u][0].x * x0.x + a[u][0].y * x0.y + a[u][1].x * x1.x + a[u][1].y * x1.y + a[u][2].x * x2.x +
               a[u][2].y * x2.y;
      }
    }
    p += kAhead;
  }
  for (; p < p1; p++) {
    const double *a = vals + 36L * p + 6 * r;
    const double *xx = x + 6L * cols[p];
    acc += a[0] * xx[0] + a[1] * xx[1] + a[2] * xx[2] + a[3] * xx[3] + a[4] * xx[4] + a[5] * xx[5];
  }
  y[6L * row + r] = acc;
}

// y_v = A x_v for the vectors of every design when all designs' copies of A
// hold the same values (load cases of one structure): one lane per scalar
// row reads the copy of design 0 once for kV vectors (grid.y counts groups
// of kV), against once per vector above.  x and y are 16-byte aligned, the
// sums run in block order and each has the expression of the kernel above,
// so y agrees with it bit for bit.
template <int kV>
__global__ __launch_bounds__(192) void bsr_spmv6_same_kernel(int nrows, const int *__restrict__ rowp,
                                                             const int *__restrict__ cols,
                                                             const double *__restrict__ vals, int nv,
                                                             const double *__restrict__ x,
                                                             double *__restrict__ y, long xstride) {
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(tid / 6), r = (int)(tid % 6);
  if (row >= nrows) return;
  const int v0 = blockIdx.y * kV;
  const int nvl = min(kV, nv - v0);
  x += (long)v0 * xstride;
  y += (long)v0 * xstride;
  double acc[kV];
#pragma unroll
  for (int v = 0; v < kV; v++) acc[v] = 0.0;
  const int p1 = rowp[row + 1];
  for (int p = rowp[row]; p < p1; p++) {
    const double2 *ap = reinterpret_cast<const double2 *>(vals + 36L * p + 6 * r);
    const double2 a0 = ap[0], a1 = ap[1], a2 = ap[2];
    const long c = 6L * cols[p];
#pragma unroll
    for (int v = 0; v < kV; v++) {
      if (v < nvl) {
        const double2 *xp = reinterpret_cast<const double2 *>(x + v * xstride + c);
        const double2 x0 = xp[0], x1 = xp[1], x2 = xp[2];
        acc[v] += a0.x * x0.x + a0.y * x0.y + a1.x * x1.x + a1.y * x1.y + a2.x * x2.x + a2.y * x2.y;
      }
    }
  }
#pragma unroll
  for (int v = 0; v < kV; v++)
    if (v < nvl) y[v * xstride + 6L * row + r] = acc[v];
}

// y = A^T x, 6x6 blocks: row j of A^T gathers block (i, j) of A, transposed,
// for every block (j, i) of row j (the pattern is structurally symmetric)
__global__ __launch_bounds__(192) void bsr_spmv6_t_kernel(int nrows, const int *__restrict__ rowp,
                                                          const int *__restrict__ cols,
                                                          const int *__restrict__ tpos,
                                                          const double *__restrict__ vals, long vstride,
                                                          const double *__restrict__ x, double *__restrict__ y,
                                                          long xstride) {
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(tid / 6), r = (int)(tid % 6);
  if (row >= nrows) return;
  vals += (long)blockIdx.y * vstride;
  x += (long)blockIdx.y * xstride;
  y += (long)blockIdx.y * xstride;
  double acc = 0.0;
  for (int p = rowp[row]; p < rowp[row + 1]; p++) {
    const double *a = vals + 36L * tpos[p] + r;  // column r of block (cols[p], row)
    const double *xx = x + 6L * cols[p];
    acc += a[0] * xx[0] + a[6] * xx[1] + a[12] * xx[2] + a[18] * xx[3] + a[24] * xx[4] + a[30] * xx[5];
  }
  y[6L * row + r] = acc;
}

__global__ void scale_kernel(long n, double alpha, double *x) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] *= alpha;
}

// the 6 diagonal entries of every diagonal block += alpha (BCSRMat::addDiag)
__global__ void add_diag_kernel(int nnodes, const int *diag, double *vals, long vstride, double alpha) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 6 * nnodes) return;
  vals[(long)blockIdx.y * vstride + 36L * diag[i / 6] + 7 * (i % 6)] += alpha;
}

// dst block blk[s] += the sum, in order, of the 36-value blocks
// vals[ptr[s] .. ptr[s + 1]) (one thread per block entry)
__global__ void add_blocks_kernel(int nseg, const int *ptr, const int *blk, const double *vals, double *dst) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 36L * nseg) return;
  const int sgm = (int)(t / 36), e = (int)(t % 36);
  double acc = 0.0;
  for (int k = ptr[sgm]; k < ptr[sgm + 1]; k++) acc += vals[36L * k + e];
  dst[36L * blk[sgm] + e] += acc;
}

// w+ = a + b, w- = a - b
__global__ void sum_diff_kernel(long n, const double *a, const double *b, double *wp, double *wm) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    wp[i] = a[i] + b[i];
    wm[i] = a[i] - b[i];
  }
}

__global__ void axpy_kernel(long n, double alpha, const double *x, double *y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] += alpha * x[i];
}

int launch_grid(long n) { return (int)std::min<long>((n + 255) / 256, 65535L * 4); }

// greedy first-fit colouring of the element dual graph (the scheme of
// TacsComputeSerialMultiColor, src/utils/TacsUtilities.cpp:430-487, applied
// to elements that share a node): strategy 0 visits the elements in their
// order, strategy 1 those with the most neighbouring elements first (ties in
// element order), which can need fewer colours on meshes with junctions
void color_elements(int nnodes, int nelems, const std::vector<int> &conn, std::vector<int> &color,
                    int &ncolors, int strategy = 0) {
  std::vector<int> nep(nnodes + 1, 0);
  for (int e = 0; e < nelems; e++)
    for (int k = 0; k < 4; k++) nep[conn[4 * e + k] + 1]++;
  for (int i = 0; i < nnodes; i++) nep[i + 1] += nep[i];
  std::vector<int> n2e(nep[nnodes]), fill(nep.begin(), nep.end() - 1);
  for (int e = 0; e < nelems; e++)
    for (int k = 0; k < 4; k++) n2e[fill[conn[4 * e + k]]++] = e;
  std::vector<int> order(nelems);
  for (int e = 0; e < nelems; e++) order[e] = e;
  if (strategy == 1) {
    // neighbour counts (with repeats over shared nodes: a proxy of the degree)
    std::vector<int> deg(nelems, 0);
    for (int e = 0; e < nelems; e++)
      for (int k = 0; k < 4; k++) {
        const int n = conn[4 * e + k];
        deg[e] += nep[n + 1] - nep[n] - 1;
      }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return deg[x] > deg[y]; });
  }
  color.assign(nelems, -1);
  ncolors = 0;
  std::vector<int> mark;
  for (int e : order) {
    mark.assign(ncolors + 1, 0);
    for (int k = 0; k < 4; k++) {
      const int n = conn[4 * e + k];
      for (int p = nep[n]; p < nep[n + 1]; p++) {
        const int c = color[n2e[p]];
        if (c >= 0) mark[c] = 1;
      }
    }
    int c = 0;
    while (mark[c]) c++;
    color[e] = c;
    ncolors = std::max(ncolors, c + 1);
  }
}

int check_batch(shellgpu_ctx *ctx, const shellgpu_mat *A) {
  if (!ctx->d_thick) return set_err(ctx, 3, "designs not set (shellgpu_set_designs)");
  if (A && A->nbatch != ctx->nbatch)
    return set_err(ctx, 3, "matrix batch size does not match the context's designs");
  return 0;
}

int record_begin(shellgpu_ctx *ctx, std::vector<std::pair<hipEvent_t, hipEvent_t>> &ev) {
  if (!ctx->timing) return 0;
  hipEvent_t a, b;
  HIPCALL(ctx, hipEventCreate(&a));
  HIPCALL(ctx, hipEventCreate(&b));
  ev.push_back({a, b});
  HIPCALL(ctx, hipEventRecord(a, ctx->stream));
  return 0;
}

int record_end(shellgpu_ctx *ctx, std::vector<std::pair<hipEvent_t, hipEvent_t>> &ev) {
  if (!ctx->timing) return 0;
  HIPCALL(ctx, hipEventRecord(ev.back().second, ctx->stream));
  return 0;
}

void drop_events(std::vector<std::pair<hipEvent_t, hipEvent_t>> &ev) {
  for (auto &p : ev) {
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  ev.clear();
}

// colour the elements with the given strategy (color_elements) and upload
// the colour lists and the first-touch bits of the colour launch order
int apply_coloring(shellgpu_ctx *ctx, int strategy) {
  const int nelems = ctx->nelems;
  const std::vector<int> &conn = ctx->h_conn;
  color_elements(ctx->nnodes, nelems, conn, ctx->h_color, ctx->ncolors, strategy);
  ctx->coloring = strategy;
  ctx->h_color_ptr.assign(ctx->ncolors + 1, 0);
  for (int e = 0; e < nelems; e++) ctx->h_color_ptr[ctx->h_color[e] + 1]++;
  for (int c = 0; c < ctx->ncolors; c++) ctx->h_color_ptr[c + 1] += ctx->h_color_ptr[c];
  ctx->h_color_elems.assign(nelems, 0);
  {
    std::vector<int> pos(ctx->h_color_ptr.begin(), ctx->h_color_ptr.end() - 1);
    for (int e = 0; e < nelems; e++) ctx->h_color_elems[pos[ctx->h_color[e]]++] = e;
  }

  // first-touch bits in colour launch order (elements of one colour share no
  // node, so no block is reached twice within a colour)
  std::vector<int> first(nelems, 0);
  bool distinct = true;
  for (int e = 0; e < nelems && distinct; e++)
    for (int i = 0; i < 4; i++)
      for (int j = i + 1; j < 4; j++) distinct = distinct && conn[4 * e + i] != conn[4 * e + j];
  if (distinct) {
    std::vector<char> seen(ctx->nnzb, 0);
    for (int e : ctx->h_color_elems)
      for (int k = 0; k < 16; k++) {
        const int b = ctx->h_blk[16L * e + k];
        if (!seen[b]) {
          seen[b] = 1;
          first[e] |= 1 << k;
        }
      }
  }
  int rc = 0;
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->d_first) {
    (void)hipFree(ctx->d_first);
    ctx->d_first = nullptr;
  }
  if (distinct && (rc = upload(ctx, &ctx->d_first, first.data(), first.size()))) return rc;
  if ((rc = upload(ctx, &ctx->d_color_elems, ctx->h_color_elems.data(), ctx->h_color_elems.size())))
    return rc;
  return 0;
}

}  // namespace

namespace shellgpu {

ElemArgs base_args(shellgpu_ctx *ctx, const double *u) {
  ElemArgs a{};
  a.conn = ctx->d_conn;
  a.blk = ctx->d_blk;
  a.Xpts = ctx->d_X;
  a.state = u;
  a.state_stride = 6L * ctx->nnodes;
  a.thick = ctx->d_thick;
  a.temp = ctx->d_temp;
  a.nelems = ctx->nelems;
  a.mat = ctx->mat;
  a.alpha = 1.0;
  a.vals_stride = 36L * ctx->nnzb;
  a.res_stride = 6L * ctx->nnodes;
  return a;
}

// one launch per colour, back to back on the stream; with timing enabled one
// event pair brackets the colour sequence (events between the launches
// would add about 3 us each to the pass)
int run_colored(shellgpu_ctx *ctx, int mode, ElemArgs a) {
  int rc = 0;
  if ((rc = record_begin(ctx, ctx->elem_ev))) return rc;
  for (int c = 0; c < ctx->ncolors; c++) {
    a.elems = ctx->d_color_elems + ctx->h_color_ptr[c];
    const int nb = ctx->h_color_ptr[c + 1] - ctx->h_color_ptr[c];
    HIPCALL(ctx, launch_elem_kernel(mode, ctx->elem_variant[mode & 31], a, nb, ctx->nbatch, ctx->stream));
  }
  if (ctx->timing) ctx->elem_launches += ctx->ncolors;
  return record_end(ctx, ctx->elem_ev);
}

int run_xpt(shellgpu_ctx *ctx, const double *phi, const double *u, const double *psi, const double *coef,
            int nonlin, int terms, double *dfdX) {
  XptArgs a{};
  a.conn = ctx->d_conn;
  a.Xpts = ctx->d_X;
  a.thick = ctx->d_thick;
  a.temp = ctx->d_temp;
  a.nelems = ctx->nelems;
  a.mat = ctx->mat;
  a.phi = phi;
  a.u = u;
  a.psi = psi;
  a.stride = 6L * ctx->nnodes;
  a.coef = coef;
  a.nonlin = nonlin;
  a.dfdX = dfdX;
  a.dX_stride = 3L * ctx->nnodes;
  a.terms = terms;
  for (int c = 0; c < ctx->ncolors; c++) {
    a.elems = ctx->d_color_elems + ctx->h_color_ptr[c];
    const int nb = ctx->h_color_ptr[c + 1] - ctx->h_color_ptr[c];
    HIPCALL(ctx, launch_xpt_kernel(a, nb, ctx->nbatch, ctx->stream));
  }
  return 0;
}

}  // namespace shellgpu

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int shellgpu_create(shellgpu_ctx **out, int device) {
  if (!out) return 1;
  *out = nullptr;
  shellgpu_ctx *ctx = new shellgpu_ctx();
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreate(&ctx->stream);
  if (e == hipSuccess) {
    e = map_scratch(device, &ctx->scratch);
    ctx->scratch_doubles = kScratchDoubles;
  }
  if (e != hipSuccess) {
    fprintf(stderr, "shellgpu_create: %s\n", hipGetErrorString(e));
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 2;
  }
  *out = ctx;
  return 0;
}

void shellgpu_destroy(shellgpu_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  void *bufs[] = {ctx->d_conn, ctx->d_rowp,     ctx->d_cols,    ctx->d_blk, ctx->d_diag, ctx->d_tpos,
                  ctx->d_color_elems, ctx->d_X, ctx->d_bc_nodes, ctx->d_bc_mask,
                  ctx->d_bc_vals, ctx->d_thick, ctx->d_temp, ctx->d_coef, ctx->d_first};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  drop_events(ctx->elem_ev);
  drop_events(ctx->spmv_ev);
  if (ctx->t0) (void)hipEventDestroy(ctx->t0);
  if (ctx->t1) (void)hipEventDestroy(ctx->t1);
  if (ctx->blas) (void)rocblas_destroy_handle((rocblas_handle)ctx->blas);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *shellgpu_last_error(const shellgpu_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int shellgpu_synchronize(shellgpu_ctx *ctx) {
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

void *shellgpu_get_stream(shellgpu_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int shellgpu_malloc(shellgpu_ctx *ctx, void **dptr, size_t bytes) {
  HIPCALL(ctx, hipMalloc(dptr, bytes));
  return 0;
}
int shellgpu_free(shellgpu_ctx *ctx, void *dptr) {
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  HIPCALL(ctx, hipFree(dptr));
  return 0;
}
int shellgpu_memcpy_h2d(shellgpu_ctx *ctx, void *dst, const void *src, size_t bytes) {
  HIPCALL(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}
int shellgpu_memcpy_d2h(shellgpu_ctx *ctx, void *dst, const void *src, size_t bytes) {
  HIPCALL(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}
int shellgpu_memcpy_d2d(shellgpu_ctx *ctx, void *dst, const void *src, size_t bytes) {
  HIPCALL(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return 0;
}
int shellgpu_memset(shellgpu_ctx *ctx, void *dst, int value, size_t bytes) {
  HIPCALL(ctx, hipMemsetAsync(dst, value, bytes, ctx->stream));
  return 0;
}

int shellgpu_set_mesh(shellgpu_ctx *ctx, int nnodes, int nelems, const int *conn, const double *Xpts) {
  if (!ctx || nnodes <= 0 || nelems <= 0 || !conn || !Xpts) return set_err(ctx, 1, "invalid mesh arguments");
  for (long i = 0; i < 4L * nelems; i++)
    if (conn[i] < 0 || conn[i] >= nnodes) return set_err(ctx, 1, "connectivity entry out of range");
  ctx->nnodes = nnodes;
  ctx->nelems = nelems;
  ctx->h_conn.assign(conn, conn + 4L * nelems);

  // node-to-node pattern: sorted, unique, diagonal kept
  // (TACSAssembler::computeLocalNodeToNodeCSR + TacsSortAndUniquifyCSR)
  std::vector<int> nep(nnodes + 1, 0);
  for (int e = 0; e < nelems; e++)
    for (int k = 0; k < 4; k++) nep[conn[4 * e + k] + 1]++;
  for (int i = 0; i < nnodes; i++) nep[i + 1] += nep[i];
  std::vector<int> n2e(nep[nnodes]), fill(nep.begin(), nep.end() - 1);
  for (int e = 0; e < nelems; e++)
    for (int k = 0; k < 4; k++) n2e[fill[conn[4 * e + k]]++] = e;
  ctx->h_rowp.assign(nnodes + 1, 0);
  ctx->h_cols.clear();
  std::vector<int> row;
  for (int i = 0; i < nnodes; i++) {
    row.clear();
    for (int p = nep[i]; p < nep[i + 1]; p++)
      for (int k = 0; k < 4; k++) row.push_back(conn[4 * n2e[p] + k]);
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    ctx->h_cols.insert(ctx->h_cols.end(), row.begin(), row.end());
    ctx->h_rowp[i + 1] = (int)ctx->h_cols.size();
  }
  ctx->nnzb = (int)ctx->h_cols.size();

  // element -> BSR block table
  ctx->h_blk.assign(16L * nelems, 0);
  for (int e = 0; e < nelems; e++)
    for (int i = 0; i < 4; i++) {
      const int r = conn[4 * e + i];
      const int *c0 = &ctx->h_cols[ctx->h_rowp[r]], *c1 = &ctx->h_cols[ctx->h_rowp[r + 1]];
      for (int j = 0; j < 4; j++) {
        const int *it = std::lower_bound(c0, c1, conn[4 * e + j]);
        ctx->h_blk[16L * e + 4 * i + j] = (int)(it - ctx->h_cols.data());
      }
    }

  // diagonal block of every node; transposed position of every block
  std::vector<int> diag(nnodes), tpos(ctx->nnzb);
  for (int i = 0; i < nnodes; i++) {
    const int *c0 = &ctx->h_cols[ctx->h_rowp[i]], *c1 = &ctx->h_cols[ctx->h_rowp[i + 1]];
    diag[i] = (int)(std::lower_bound(c0, c1, i) - ctx->h_cols.data());
    for (int p = ctx->h_rowp[i]; p < ctx->h_rowp[i + 1]; p++) {
      const int j = ctx->h_cols[p];
      const int *d0 = &ctx->h_cols[ctx->h_rowp[j]], *d1 = &ctx->h_cols[ctx->h_rowp[j + 1]];
      tpos[p] = (int)(std::lower_bound(d0, d1, i) - ctx->h_cols.data());
    }
  }

  int rc = apply_coloring(ctx, ctx->coloring);
  if (rc) return rc;
  if ((rc = upload(ctx, &ctx->d_conn, conn, 4L * nelems))) return rc;
  if ((rc = upload(ctx, &ctx->d_X, Xpts, 3L * nnodes))) return rc;
  if ((rc = upload(ctx, &ctx->d_rowp, ctx->h_rowp.data(), ctx->h_rowp.size()))) return rc;
  if ((rc = upload(ctx, &ctx->d_cols, ctx->h_cols.data(), ctx->h_cols.size()))) return rc;
  if ((rc = upload(ctx, &ctx->d_blk, ctx->h_blk.data(), ctx->h_blk.size()))) return rc;
  if ((rc = upload(ctx, &ctx->d_diag, diag.data(), diag.size()))) return rc;
  if ((rc = upload(ctx, &ctx->d_tpos, tpos.data(), tpos.size()))) return rc;
  return 0;
}

int shellgpu_set_nodes(shellgpu_ctx *ctx, const double *Xpts) {
  if (!ctx || !Xpts) return set_err(ctx, 1, "invalid node arguments");
  if (ctx->nnodes <= 0 || !ctx->d_X) return set_err(ctx, 1, "set the mesh before the nodes");
  HIPCALL(ctx, hipMemcpyAsync(ctx->d_X, Xpts, sizeof(double) * 3 * (size_t)ctx->nnodes, hipMemcpyDeviceToDevice,
                              ctx->stream));
  return 0;
}

int shellgpu_get_nodes(shellgpu_ctx *ctx, double *Xpts) {
  if (!ctx || !Xpts) return set_err(ctx, 1, "invalid node arguments");
  if (ctx->nnodes <= 0 || !ctx->d_X) return set_err(ctx, 1, "no mesh");
  HIPCALL(ctx, hipMemcpyAsync(Xpts, ctx->d_X, sizeof(double) * 3 * (size_t)ctx->nnodes, hipMemcpyDeviceToDevice,
                              ctx->stream));
  return 0;
}

int shellgpu_set_material(shellgpu_ctx *ctx, double E, double nu, double rho, double cte, double k_drill) {
  if (!ctx) return 1;
  if (!(E > 0.0) || !(nu > -1.0 && nu < 0.5)) return set_err(ctx, 1, "invalid material");
  ctx->mat = make_material(E, nu, rho, cte, k_drill);
  return 0;
}

int shellgpu_set_bcs(shellgpu_ctx *ctx, int nbc, const int *nodes, const int *mask, const double *vals) {
  if (!ctx || nbc < 0) return 1;
  for (int i = 0; i < nbc; i++)
    if (nodes[i] < 0 || nodes[i] >= ctx->nnodes) return set_err(ctx, 1, "BC node out of range");
  ctx->nbc = nbc;
  ctx->h_bc_nodes.assign(nodes, nodes + nbc);
  ctx->h_bc_mask.assign(mask, mask + nbc);
  int rc = 0;
  if ((rc = upload(ctx, &ctx->d_bc_nodes, nodes, (size_t)nbc))) return rc;
  if ((rc = upload(ctx, &ctx->d_bc_mask, mask, (size_t)nbc))) return rc;
  std::vector<double> v(6L * nbc, 0.0);
  if (vals) std::copy(vals, vals + 6L * nbc, v.begin());
  if ((rc = upload(ctx, &ctx->d_bc_vals, v.data(), v.size()))) return rc;
  return 0;
}

int shellgpu_set_designs(shellgpu_ctx *ctx, int nbatch, const double *thick, const double *temp) {
  if (!ctx || nbatch <= 0 || !thick) return set_err(ctx, 1, "invalid design arguments");
  if (ctx->nelems <= 0) return set_err(ctx, 1, "set the mesh before the designs");
  ctx->nbatch = nbatch;
  {
    const size_t ne = ctx->nelems;
    bool same = true;
    for (int d = 1; d < nbatch && same; d++)
      same = std::equal(thick, thick + ne, thick + d * ne) && (!temp || std::equal(temp, temp + ne, temp + d * ne));
    ctx->same_designs = same;
  }
  int rc = 0;
  if ((rc = upload(ctx, &ctx->d_thick, thick, (size_t)nbatch * ctx->nelems))) return rc;
  if (temp) {
    if ((rc = upload(ctx, &ctx->d_temp, temp, (size_t)nbatch * ctx->nelems))) return rc;
  } else if (ctx->d_temp) {
    (void)hipFree(ctx->d_temp);
    ctx->d_temp = nullptr;
  }
  if (ctx->d_coef) (void)hipFree(ctx->d_coef);
  ctx->d_coef = nullptr;
  HIPCALL(ctx, hipMalloc((void **)&ctx->d_coef, sizeof(double) * 4 * nbatch));
  return 0;
}

int shellgpu_get_sizes(shellgpu_ctx *ctx, int *nnodes, int *nelems, int *nnzb, int *ncolors, int *nbatch) {
  if (!ctx) return 1;
  if (nnodes) *nnodes = ctx->nnodes;
  if (nelems) *nelems = ctx->nelems;
  if (nnzb) *nnzb = ctx->nnzb;
  if (ncolors) *ncolors = ctx->ncolors;
  if (nbatch) *nbatch = ctx->nbatch;
  return 0;
}

int shellgpu_get_pattern(shellgpu_ctx *ctx, int *rowp, int *cols) {
  if (!ctx) return 1;
  if (rowp) std::copy(ctx->h_rowp.begin(), ctx->h_rowp.end(), rowp);
  if (cols) std::copy(ctx->h_cols.begin(), ctx->h_cols.end(), cols);
  return 0;
}

int shellgpu_get_colors(shellgpu_ctx *ctx, int *elem_color) {
  if (!ctx) return 1;
  std::copy(ctx->h_color.begin(), ctx->h_color.end(), elem_color);
  return 0;
}

int shellgpu_mat_create(shellgpu_ctx *ctx, int nbatch, shellgpu_mat **A) {
  if (!ctx || !A || nbatch <= 0) return set_err(ctx, 1, "invalid matrix arguments");
  if (ctx->nnzb <= 0) return set_err(ctx, 1, "set the mesh before creating matrices");
  shellgpu_mat *m = new shellgpu_mat();
  m->ctx = ctx;
  m->nbatch = nbatch;
  hipError_t e = hipMalloc((void **)&m->vals, sizeof(double) * 36L * ctx->nnzb * nbatch);
  if (e != hipSuccess) {
    delete m;
    return set_err(ctx, 2, std::string("hipMalloc matrix: ") + hipGetErrorString(e));
  }
  e = hipMemsetAsync(m->vals, 0, sizeof(double) * 36L * ctx->nnzb * nbatch, ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(m->vals);
    delete m;
    return set_err(ctx, 2, "hipMemset matrix");
  }
  m->same = true;
  *A = m;
  return 0;
}

void shellgpu_mat_destroy(shellgpu_mat *A) {
  if (!A) return;
  (void)hipStreamSynchronize(A->ctx->stream);
  (void)hipFree(A->vals);
  delete A;
}

double *shellgpu_mat_values(shellgpu_mat *A) {
  if (!A) return nullptr;
  A->same = false;  // the caller may write design by design
  return A->vals;
}

int shellgpu_mat_get_values(shellgpu_mat *A, double *h) {
  shellgpu_ctx *ctx = A->ctx;
  HIPCALL(ctx, hipMemcpyAsync(h, A->vals, sizeof(double) * 36L * ctx->nnzb * A->nbatch, hipMemcpyDeviceToHost,
                              ctx->stream));
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int shellgpu_mat_set_values(shellgpu_mat *A, const double *h) {
  shellgpu_ctx *ctx = A->ctx;
  A->same = false;
  HIPCALL(ctx, hipMemcpyAsync(A->vals, h, sizeof(double) * 36L * ctx->nnzb * A->nbatch, hipMemcpyHostToDevice,
                              ctx->stream));
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int shellgpu_mat_zero(shellgpu_mat *A) {
  shellgpu_ctx *ctx = A->ctx;
  HIPCALL(ctx, hipMemsetAsync(A->vals, 0, sizeof(double) * 36L * ctx->nnzb * A->nbatch, ctx->stream));
  A->same = true;
  return 0;
}

int shellgpu_mat_copy(shellgpu_mat *dst, const shellgpu_mat *src) {
  shellgpu_ctx *ctx = dst->ctx;
  if (dst->nbatch != src->nbatch || dst->ctx != src->ctx) return set_err(ctx, 1, "matrix shapes differ");
  HIPCALL(ctx, hipMemcpyAsync(dst->vals, src->vals, sizeof(double) * 36L * ctx->nnzb * src->nbatch,
                              hipMemcpyDeviceToDevice, ctx->stream));
  dst->same = src->same;
  return 0;
}

int shellgpu_mat_axpy(shellgpu_mat *Y, double alpha, const shellgpu_mat *X) {
  shellgpu_ctx *ctx = Y->ctx;
  if (Y->nbatch != X->nbatch || Y->ctx != X->ctx) return set_err(ctx, 1, "matrix shapes differ");
  const long n = 36L * ctx->nnzb * Y->nbatch;
  hipLaunchKernelGGL(axpy_kernel, dim3(launch_grid(n)), dim3(256), 0, ctx->stream, n, alpha, X->vals, Y->vals);
  HIPCALL(ctx, hipGetLastError());
  Y->same = Y->same && X->same;
  return 0;
}

int shellgpu_mat_scale(shellgpu_mat *A, double alpha) {
  shellgpu_ctx *ctx = A->ctx;
  const long n = 36L * ctx->nnzb * A->nbatch;
  hipLaunchKernelGGL(scale_kernel, dim3(launch_grid(n)), dim3(256), 0, ctx->stream, n, alpha, A->vals);
  HIPCALL(ctx, hipGetLastError());
  return 0;
}

int shellgpu_mat_add_diag(shellgpu_mat *A, double alpha) {
  shellgpu_ctx *ctx = A->ctx;
  hipLaunchKernelGGL(add_diag_kernel, dim3((6 * ctx->nnodes + 255) / 256, A->nbatch), dim3(256), 0, ctx->stream,
                     ctx->nnodes, ctx->d_diag, A->vals, 36L * ctx->nnzb, alpha);
  HIPCALL(ctx, hipGetLastError());
  return 0;
}

int shellgpu_mat_add_blocks(shellgpu_mat *A, int design, int nblk, const int *brow, const int *bcol,
                            const double *vals) {
  shellgpu_ctx *ctx = A ? A->ctx : nullptr;
  if (!A) return 1;
  if (design < 0 || design >= A->nbatch) return set_err(ctx, 1, "design out of range");
  if (nblk < 0 || (nblk > 0 && (!brow || !bcol || !vals))) return set_err(ctx, 1, "invalid block arguments");
  // target block of every input block (-1: not in the pattern, skipped)
  std::vector<int> tgt(nblk, -1);
  long missing = 0;
  int first_r = -1, first_c = -1;
  for (int k = 0; k < nblk; k++) {
    const int r = brow[k], c = bcol[k];
    if (r < 0 || c < 0) continue;  // negative indices are skipped, as in BCSRMat::addRowValues
    if (r < ctx->nnodes && c < ctx->nnodes) {
      const int *c0 = &ctx->h_cols[ctx->h_rowp[r]], *c1 = &ctx->h_cols[ctx->h_rowp[r + 1]];
      const int *it = std::lower_bound(c0, c1, c);
      if (it != c1 && *it == c) tgt[k] = (int)(it - ctx->h_cols.data());
    }
    if (tgt[k] < 0 && missing++ == 0) {
      first_r = r;
      first_c = c;
    }
  }
  // stable order by target block: every block's additions are summed in
  // input order, so the result does not depend on the launch
  std::vector<int> order;
  for (int k = 0; k < nblk; k++)
    if (tgt[k] >= 0) order.push_back(k);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return tgt[x] < tgt[y]; });
  std::vector<double> sv(36 * order.size());
  std::vector<int> ptr{0}, blk;
  for (size_t q = 0; q < order.size(); q++) {
    std::copy(vals + 36L * order[q], vals + 36L * order[q] + 36, sv.begin() + 36 * q);
    if (q > 0 && tgt[order[q]] != tgt[order[q - 1]]) ptr.push_back((int)q);
    if (q == 0 || tgt[order[q]] != tgt[order[q - 1]]) blk.push_back(tgt[order[q]]);
  }
  ptr.push_back((int)order.size());
  const int nseg = (int)blk.size();
  if (nseg > 0) {
    double *dv = nullptr;
    int *dp = nullptr, *db = nullptr;
    hipError_t e = hipMalloc((void **)&dv, sizeof(double) * sv.size());
    if (e == hipSuccess) e = hipMalloc((void **)&dp, sizeof(int) * ptr.size());
    if (e == hipSuccess) e = hipMalloc((void **)&db, sizeof(int) * blk.size());
    if (e == hipSuccess) e = hipMemcpyAsync(dv, sv.data(), sizeof(double) * sv.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dp, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(db, blk.data(), sizeof(int) * blk.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
      A->same = false;
      hipLaunchKernelGGL(add_blocks_kernel, dim3((unsigned)((36L * nseg + 255) / 256)), dim3(256), 0, ctx->stream,
                         nseg, dp, db, dv, A->vals + (long)design * 36L * ctx->nnzb);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dv);
    (void)hipFree(dp);
    (void)hipFree(db);
    if (e != hipSuccess) return set_err(ctx, 2, std::string("add blocks: ") + hipGetErrorString(e));
  }
  if (missing > 0)
    return set_err(ctx, 7, std::to_string(missing) + " block(s) not in the pattern (first: row " +
                               std::to_string(first_r) + ", column " + std::to_string(first_c) +
                               "); the others were added");
  return 0;
}

int shellgpu_mat_mult_transpose(shellgpu_mat *A, const double *x, double *y) {
  shellgpu_ctx *ctx = A->ctx;
  const long nthreads = 6L * ctx->nnodes;
  const int blocks = (int)((nthreads + 191) / 192);
  hipLaunchKernelGGL(bsr_spmv6_t_kernel, dim3(blocks, A->nbatch), dim3(192), 0, ctx->stream, ctx->nnodes,
                     ctx->d_rowp, ctx->d_cols, ctx->d_tpos, A->vals, 36L * ctx->nnzb, x, y, 6L * ctx->nnodes);
  HIPCALL(ctx, hipGetLastError());
  return 0;
}

int shellgpu_mat_apply_bcs(shellgpu_mat *A, double diag) {
  shellgpu_ctx *ctx = A->ctx;
  if (ctx->nbc == 0) return 0;
  hipLaunchKernelGGL(bsr_bc_kernel, dim3(ctx->nbc, A->nbatch), dim3(64), 0, ctx->stream, ctx->d_rowp, ctx->d_cols,
                     A->vals, 36L * ctx->nnzb, ctx->d_bc_nodes, ctx->d_bc_mask, diag);
  HIPCALL(ctx, hipGetLastError());
  return 0;
}

int shellgpu_mat_mult(shellgpu_mat *A, const double *x, double *y) {
  shellgpu_ctx *ctx = A->ctx;
  const long nthreads = 6L * ctx->nnodes;
  const int blocks = (int)((nthreads + 191) / 192);
  int rc = record_begin(ctx, ctx->spmv_ev);
  if (rc) return rc;
  // 16-byte loads need x on a 16-byte boundary (the matrix values are)
  const bool aligned = (reinterpret_cast<uintptr_t>(x) & 15) == 0;
  if (aligned && A->same && A->nbatch > 1) {
    constexpr int kV = 8;
    hipLaunchKernelGGL(bsr_spmv6_same_kernel<kV>, dim3(blocks, (A->nbatch + kV - 1) / kV), dim3(192), 0, ctx->stream,
                       ctx->nnodes, ctx->d_rowp, ctx->d_cols, A->vals, A->nbatch, x, y, 6L * ctx->nnodes);
  } else if (aligned)
    hipLaunchKernelGGL(bsr_spmv6_kernel<true>, dim3(blocks, A->nbatch), dim3(192), 0, ctx->stream, ctx->nnodes,
                       ctx->d_rowp, ctx->d_cols, A->vals, 36L * ctx->nnzb, x, y, 6L * ctx->nnodes);
  else
    hipLaunchKernelGGL(bsr_spmv6_kernel<false>, dim3(blocks, A->nbatch), dim3(192), 0, ctx->stream, ctx->nnodes,
                       ctx->d_rowp, ctx->d_cols, A->vals, 36L * ctx->nnzb, x, y, 6L * ctx->nnodes);
  HIPCALL(ctx, hipGetLastError());
  return record_end(ctx, ctx->spmv_ev);
}

// one matrix type written (accumulate = false: the previous values are
// overwritten) or added (accumulate = true) into A with factor alpha; no BCs
static int assemble_type(shellgpu_ctx *ctx, int mattype, const double *u, shellgpu_mat *A, double alpha,
                         bool accumulate) {
  if (mattype != SHELLGPU_STIFFNESS && mattype != SHELLGPU_GEOMETRIC && mattype != SHELLGPU_GEOMETRIC_FD &&
      mattype != SHELLGPU_TANGENT_NONLINEAR && mattype != SHELLGPU_MASS)
    return set_err(ctx, 1, "unknown matrix type");
  int rc = 0;
  ElemArgs a = base_args(ctx, u);
  a.vals = A->vals;
  a.alpha = alpha;
  // the stiffness-type modes write each block at its first touch, which
  // replaces zeroing the matrix; the mass kernel adds
  a.first = (accumulate || mattype == SHELLGPU_MASS) ? nullptr : ctx->d_first;
  if (!accumulate && !a.first)
    HIPCALL(ctx, hipMemsetAsync(A->vals, 0, sizeof(double) * 36L * ctx->nnzb * A->nbatch, ctx->stream));
  // identical designs give identical copies unless a per-design state enters
  const bool state_free = !u || mattype == SHELLGPU_STIFFNESS || mattype == SHELLGPU_MASS;
  A->same = (accumulate ? A->same : true) && ctx->same_designs && state_free;
  if (mattype == SHELLGPU_STIFFNESS) {
    if ((rc = run_colored(ctx, kModeKLin, a))) return rc;
  } else if (mattype == SHELLGPU_GEOMETRIC) {
    if ((rc = run_colored(ctx, kModeG, a))) return rc;
  } else if (mattype == SHELLGPU_GEOMETRIC_FD) {
    a.fd_sign = 1;
    if ((rc = run_colored(ctx, kModeKNl, a))) return rc;
    a.fd_sign = -1;
    a.first = nullptr;  // the second difference adds to the first
    if ((rc = run_colored(ctx, kModeKNl, a))) return rc;
  } else if (mattype == SHELLGPU_TANGENT_NONLINEAR) {
    if ((rc = run_colored(ctx, kModeKNl, a))) return rc;
  } else {
    if ((rc = run_colored(ctx, kModeMass, a))) return rc;
  }
  return 0;
}

int shellgpu_assemble_mattype(shellgpu_ctx *ctx, int mattype, const double *u, shellgpu_mat *A) {
  int rc = check_batch(ctx, A);
  if (rc) return rc;
  if ((rc = assemble_type(ctx, mattype, u, A, 1.0, false))) return rc;
  return shellgpu_mat_apply_bcs(A, 1.0);
}

int shellgpu_assemble_mat_combo(shellgpu_ctx *ctx, int nmats, const int *mattypes, const double *scale,
                                const double *u, shellgpu_mat *A) {
  int rc = check_batch(ctx, A);
  if (rc) return rc;
  if (nmats < 0 || (nmats > 0 && (!mattypes || !scale))) return set_err(ctx, 1, "invalid matrix combination");
  HIPCALL(ctx, hipMemsetAsync(A->vals, 0, sizeof(double) * 36L * ctx->nnzb * A->nbatch, ctx->stream));
  A->same = true;
  for (int k = 0; k < nmats; k++)
    if ((rc = assemble_type(ctx, mattypes[k], u, A, scale[k], true))) return rc;
  return shellgpu_mat_apply_bcs(A, 1.0);
}

int shellgpu_add_jacobian_vec_product(shellgpu_ctx *ctx, int model, double scale, double alpha, double beta,
                                      double gamma, const double *u, const double *dudt, const double *d2udt2,
                                      const double *x, double *y, int transpose) {
  int rc = check_batch(ctx, nullptr);
  if (rc) return rc;
  if (!x || !y) return set_err(ctx, 1, "x and y are required");
  // no velocity terms (TACSDirector.h:368-432); u'' enters only the residual
  (void)beta;
  (void)dudt;
  (void)d2udt2;
  ElemArgs a = base_args(ctx, u);
  a.alpha = alpha;
  a.gamma = gamma;
  a.mvx = x;
  a.mvy = y;
  a.mv_stride = 6L * ctx->nnodes;
  a.mv_scale = scale;
  a.mv_trans = transpose ? 1 : 0;
  const bool nonlin = model == SHELLGPU_NONLINEAR;
  if ((rc = run_colored(ctx, nonlin ? kModeJvpNl : kModeJvpLin, a))) return rc;
  return shellgpu_vec_zero_bcs(ctx, y);
}

int shellgpu_assemble_res(shellgpu_ctx *ctx, int model, const double *u, double *R) {
  int rc = check_batch(ctx, nullptr);
  if (rc) return rc;
  HIPCALL(ctx, hipMemsetAsync(R, 0, sizeof(double) * 6L * ctx->nnodes * ctx->nbatch, ctx->stream));
  ElemArgs a = base_args(ctx, u);
  a.res = R;
  if ((rc = run_colored(ctx, model == SHELLGPU_NONLINEAR ? kModeRNl : kModeRLin, a))) return rc;
  if (ctx->nbc > 0) {
    hipLaunchKernelGGL(vec_bc_kernel, dim3((6 * ctx->nbc + 255) / 256, ctx->nbatch), dim3(256), 0, ctx->stream, R,
                       u, 6L * ctx->nnodes, ctx->nbc, ctx->d_bc_nodes, ctx->d_bc_mask, ctx->d_bc_vals, 2);
    HIPCALL(ctx, hipGetLastError());
  }
  return 0;
}

int shellgpu_element_output(shellgpu_ctx *ctx, int model, const double *u, double ys, double *out) {
  int rc = check_batch(ctx, nullptr);
  if (rc) return rc;
  if (!out) return set_err(ctx, 1, "out is required");
  if (!(ys > 0.0)) return set_err(ctx, 1, "the yield stress must be positive");
  ElemArgs a = base_args(ctx, u);
  a.out = out;
  a.ys = ys;
  return run_colored(ctx, model == SHELLGPU_NONLINEAR ? kModeOutNl : kModeOutLin, a);
}

int shellgpu_assemble_jacobian(shellgpu_ctx *ctx, int model, double alpha, double beta, double gamma,
                               const double *u, const double *dudt, const double *d2udt2, double *R,
                               shellgpu_mat *A) {
  int rc = check_batch(ctx, A);
  if (rc) return rc;
  if (!A) return set_err(ctx, 1, "A is required");
  // the linearized-rotation director has no velocity-dependent terms
  // (TACSLinearizedRotation::addDirectorJacobian, TACSDirector.h:368-432):
  // beta and du/dt do not enter
  (void)beta;
  (void)dudt;
  ElemArgs a = base_args(ctx, u);
  a.vals = A->vals;
  a.alpha = alpha;
  a.gamma = gamma;
  a.ddu = d2udt2;
  a.ddu_stride = 6L * ctx->nnodes;
  a.first = ctx->d_first;
  if (!a.first)
    HIPCALL(ctx, hipMemsetAsync(A->vals, 0, sizeof(double) * 36L * ctx->nnzb * A->nbatch, ctx->stream));
  const bool nonlin = model == SHELLGPU_NONLINEAR;
  // u enters the matrix only through the nonlinear tangent; u'' only R
  A->same = ctx->same_designs && (!nonlin || !u);
  if (!R && gamma == 0.0) {
    // alpha K (or alpha K_t) alone
    if ((rc = run_colored(ctx, nonlin ? kModeKNl : kModeKLin, a))) return rc;
  } else {
    // the residual (with M u'') and alpha K + gamma M in one pass
    if (R) {
      HIPCALL(ctx, hipMemsetAsync(R, 0, sizeof(double) * 6L * ctx->nnodes * ctx->nbatch, ctx->stream));
      a.res = R;
    }
    if ((rc = run_colored(ctx, nonlin ? kModeJacNl : kModeJacLin, a))) return rc;
    if (R && ctx->nbc > 0) {
      hipLaunchKernelGGL(vec_bc_kernel, dim3((6 * ctx->nbc + 255) / 256, ctx->nbatch), dim3(256), 0, ctx->stream,
                         R, u, 6L * ctx->nnodes, ctx->nbc, ctx->d_bc_nodes, ctx->d_bc_mask, ctx->d_bc_vals, 2);
      HIPCALL(ctx, hipGetLastError());
    }
  }
  return shellgpu_mat_apply_bcs(A, 1.0);
}

// per-design coefficients (c0, c1) into ctx->d_coef
static int set_coef(shellgpu_ctx *ctx, double c0, double c1) {
  std::vector<double> h(2 * ctx->nbatch);
  for (int d = 0; d < ctx->nbatch; d++) {
    h[2 * d] = c0;
    h[2 * d + 1] = c1;
  }
  HIPCALL(ctx, hipMemcpyAsync(ctx->d_coef, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, ctx->stream));
  // h is released on return: the copy must have consumed it
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int shellgpu_add_adj_res_product(shellgpu_ctx *ctx, int model, double scale, const double *psi, const double *u,
                                 double *dfdt) {
  int rc = check_batch(ctx, nullptr);
  if (rc) return rc;
  if (!psi || !dfdt) return set_err(ctx, 1, "psi and dfdt are required");
  std::vector<double> h(ctx->nbatch, scale);
  HIPCALL(ctx, hipMemcpyAsync(ctx->d_coef, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, ctx->stream));
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  ElemArgs a = base_args(ctx, u);
  a.psi = psi;
  a.psi_stride = 6L * ctx->nnodes;
  a.coef = ctx->d_coef;
  a.dv = dfdt;
  return run_colored(ctx, model == SHELLGPU_NONLINEAR ? kModeAdjResNl : kModeAdjResLin, a);
}

// psi^T A phi for a symmetric A (K, G(u), M and their derivatives) by
// polarization, psi^T A phi = (w+^T A w+ - w-^T A w-) / 4 with
// w+- = psi +- phi, so that products with psi != phi reuse the symmetric
// kernels: runs fn(w, c) for (phi, 1) when psi is null or phi, else for
// (w+, 1/4) and (w-, -1/4).
}  // extern "C"
template <class Fn>
static int polarized(shellgpu_ctx *ctx, const double *psi, const double *phi, Fn fn) {
  if (!psi || psi == phi) return fn(phi, 1.0);
  const long n = 6L * ctx->nnodes * ctx->nbatch;
  double *w = nullptr;
  HIPCALL(ctx, hipMalloc((void **)&w, sizeof(double) * 2 * n));
  hipLaunchKernelGGL(sum_diff_kernel, dim3(launch_grid(n)), dim3(256), 0, ctx->stream, n, psi, phi, w, w + n);
  int rc = hipGetLastError() == hipSuccess ? 0 : set_err(ctx, 2, "sum/difference launch failed");
  if (!rc) rc = fn(w, 0.25);
  if (!rc) rc = fn(w + n, -0.25);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(w);
  return rc;
}
extern "C" {

// dfdt += scale phi^T (dA/dt) phi
static int mat_dv_sym(shellgpu_ctx *ctx, int mattype, double scale, const double *phi, const double *u,
                      double *dfdt) {
  int rc = 0;
  if (mattype == SHELLGPU_MASS) {
    // one coefficient per design
    std::vector<double> h(ctx->nbatch, scale);
    HIPCALL(ctx, hipMemcpyAsync(ctx->d_coef, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                                ctx->stream));
    HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
    ElemArgs a = base_args(ctx, phi);
    a.coef = ctx->d_coef;
    a.dv = dfdt;
    return run_colored(ctx, kModeMassDV, a);
  }
  const bool geo = mattype == SHELLGPU_GEOMETRIC;
  if ((rc = set_coef(ctx, geo ? 0.0 : scale, geo ? scale : 0.0))) return rc;
  ElemArgs a = base_args(ctx, phi);
  a.u2 = u;
  a.u2_stride = 6L * ctx->nnodes;
  a.coef = ctx->d_coef;
  a.dv = dfdt;
  return run_colored(ctx, kModeAdjDV, a);
}

int shellgpu_add_mat_dv_sens_inner_product(shellgpu_ctx *ctx, int mattype, double scale, const double *psi,
                                           const double *phi, const double *u, double *dfdt) {
  int rc = check_batch(ctx, nullptr);
  if (rc) return rc;
  if (!phi || !dfdt) return set_err(ctx, 1, "phi and dfdt are required");
  if (mattype != SHELLGPU_STIFFNESS && mattype != SHELLGPU_GEOMETRIC && mattype != SHELLGPU_MASS)
    return set_err(ctx, 1, "matrix type must be SHELLGPU_STIFFNESS, SHELLGPU_GEOMETRIC or SHELLGPU_MASS");
  return polarized(ctx, psi, phi,
                   [&](const double *w, double c) { return mat_dv_sym(ctx, mattype, c * scale, w, u, dfdt); });
}

int shellgpu_eval_mat_sv_sens_inner_product(shellgpu_ctx *ctx, const double *psi, const double *phi,
                                            double *dfdu) {
  int rc = check_batch(ctx, nullptr);
  if (rc) return rc;
  if (!phi || !dfdu) return set_err(ctx, 1, "phi and dfdu are required");
  HIPCALL(ctx, hipMemsetAsync(dfdu, 0, sizeof(double) * 6L * ctx->nnodes * ctx->nbatch, ctx->stream));
  return polarized(ctx, psi, phi, [&](const double *w, double c) {
    int r = set_coef(ctx, 0.0, c);
    if (r) return r;
    ElemArgs a = base_args(ctx, w);
    a.coef = ctx->d_coef;
    a.res = dfdu;
    return run_colored(ctx, kModeAdjDV, a);
  });
}

// per-design coefficient quadruples (c0, c1, c2, 0) into ctx->d_coef
static int set_coef4(shellgpu_ctx *ctx, double c0, double c1, double c2, double c3 = 0.0) {
  std::vector<double> h(4 * ctx->nbatch, 0.0);
  for (int d = 0; d < ctx->nbatch; d++) {
    h[4 * d] = c0;
    h[4 * d + 1] = c1;
    h[4 * d + 2] = c2;
    h[4 * d + 3] = c3;
  }
  HIPCALL(ctx, hipMemcpyAsync(ctx->d_coef, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, ctx->stream));
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int shellgpu_add_mat_xpt_sens_inner_product(shellgpu_ctx *ctx, int mattype, double scale, const double *psi,
                                            const double *phi, const double *u, double *dfdX) {
  int rc = check_batch(ctx, nullptr);
  if (rc) return rc;
  if (!phi || !dfdX) return set_err(ctx, 1, "phi and dfdX are required");
  if (mattype != SHELLGPU_STIFFNESS && mattype != SHELLGPU_GEOMETRIC && mattype != SHELLGPU_MASS)
    return set_err(ctx, 1, "matrix type must be SHELLGPU_STIFFNESS, SHELLGPU_GEOMETRIC or SHELLGPU_MASS");
  const bool geo = mattype == SHELLGPU_GEOMETRIC;
  const int slot = mattype == SHELLGPU_STIFFNESS ? 0 : (geo ? 1 : 3);
  return polarized(ctx, psi, phi, [&](const double *w, double c) {
    double cc[4] = {0.0, 0.0, 0.0, 0.0};
    cc[slot] = c * scale;
    int r = set_coef4(ctx, cc[0], cc[1], cc[2], cc[3]);
    if (r) return r;
    return run_xpt(ctx, w, geo ? u : nullptr, nullptr, ctx->d_coef, 0, slot == 0 ? kXptK : (geo ? kXptG : kXptM),
                   dfdX);
  });
}

int shellgpu_add_adj_res_xpt_product(shellgpu_ctx *ctx, int model, double scale, const double *psi,
                                     const double *u, double *dfdX) {
  int rc = check_batch(ctx, nullptr);
  if (rc) return rc;
  if (!psi || !dfdX) return set_err(ctx, 1, "psi and dfdX are required");
  if ((rc = set_coef4(ctx, 0.0, 0.0, scale))) return rc;
  return run_xpt(ctx, nullptr, u, psi, ctx->d_coef, model == SHELLGPU_NONLINEAR, kXptR, dfdX);
}

int shellgpu_vec_zero_bcs(shellgpu_ctx *ctx, double *x) {
  if (ctx->nbc == 0) return 0;
  hipLaunchKernelGGL(vec_bc_kernel, dim3((6 * ctx->nbc + 255) / 256, std::max(ctx->nbatch, 1)), dim3(256), 0,
                     ctx->stream, x, (const double *)nullptr, 6L * ctx->nnodes, ctx->nbc, ctx->d_bc_nodes,
                     ctx->d_bc_mask, ctx->d_bc_vals, 0);
  HIPCALL(ctx, hipGetLastError());
  return 0;
}

int shellgpu_vec_set_bcs(shellgpu_ctx *ctx, double *x) {
  if (ctx->nbc == 0) return 0;
  hipLaunchKernelGGL(vec_bc_kernel, dim3((6 * ctx->nbc + 255) / 256, std::max(ctx->nbatch, 1)), dim3(256), 0,
                     ctx->stream, x, (const double *)nullptr, 6L * ctx->nnodes, ctx->nbc, ctx->d_bc_nodes,
                     ctx->d_bc_mask, ctx->d_bc_vals, 1);
  HIPCALL(ctx, hipGetLastError());
  return 0;
}

int shellgpu_device_info(shellgpu_ctx *ctx, char *name, int name_len, char *arch, int arch_len, char *pci_bus_id,
                         int pci_len, int *compute_units, size_t *hbm_bytes) {
  if (!ctx) return 1;
  hipDeviceProp_t prop;
  HIPCALL(ctx, hipGetDeviceProperties(&prop, ctx->device));
  if (name && name_len > 0) {
    snprintf(name, name_len, "%s", prop.name);
    // some drivers report no marketing name in the properties
    if (!name[0] && hipDeviceGetName(name, name_len, ctx->device) != hipSuccess) name[0] = 0;
  }
  if (arch && arch_len > 0) snprintf(arch, arch_len, "%s", prop.gcnArchName);
  char pci[64] = {0};
  HIPCALL(ctx, hipDeviceGetPCIBusId(pci, sizeof(pci), ctx->device));
  if (pci_bus_id && pci_len > 0) snprintf(pci_bus_id, pci_len, "%s", pci);
  if (name && name_len > 0 && !name[0]) {
    // without the libdrm marketing-name table the runtime reports none: the
    // kernel driver's product name of this PCI device
    std::string path = "/sys/bus/pci/devices/";
    for (const char *c = pci; *c; c++) path += (char)tolower((unsigned char)*c);
    path += "/product_name";
    if (FILE *f = fopen(path.c_str(), "r")) {
      if (fgets(name, name_len, f)) {
        size_t k = strlen(name);
        while (k > 0 && (name[k - 1] == '\n' || name[k - 1] == ' ')) name[--k] = 0;
      } else {
        name[0] = 0;
      }
      fclose(f);
    }
  }
  if (compute_units) *compute_units = prop.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
  return 0;
}

// the element-kernel classes with compiled schedules (shellgpu_tune)
static const int kTunedModes[] = {kModeKLin, kModeKNl, kModeG, kModeJacLin, kModeJacNl};
static const char *kTunedNames[] = {"stiffness", "tangent", "geometric", "jacobian_linear", "jacobian_nonlinear"};
constexpr int kTunedCount = 5;

int shellgpu_set_elem_schedule(shellgpu_ctx *ctx, int kernel_class, int variant) {
  if (!ctx) return 1;
  if (kernel_class < 0 || kernel_class >= kTunedCount || variant < 0 || variant >= kElemVariants)
    return set_err(ctx, 1, "kernel class or schedule out of range");
  ctx->elem_variant[kTunedModes[kernel_class]] = variant;
  return 0;
}

int shellgpu_tune(shellgpu_ctx *ctx, int flags) {
  int rc = check_batch(ctx, nullptr);
  if (rc) return rc;
  const size_t nv = 6L * ctx->nnodes * ctx->nbatch;
  shellgpu_mat *A = nullptr;
  double *u = nullptr, *R = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::string rep = "{";
  const bool timing = ctx->timing;
  ctx->timing = false;
  // a smooth state of small displacements and rotations, for the state-
  // dependent modes (the timing does not depend on its values)
  std::vector<double> hu(nv);
  for (size_t i = 0; i < nv; i++) hu[i] = 1e-6 * std::sin(0.001 * (double)i);
  auto trial = [&](int mode, float &best_ms) -> int {
    ElemArgs a = base_args(ctx, u);
    a.vals = A->vals;
    if (mode == kModeJacLin || mode == kModeJacNl) a.res = R;
    int r = run_colored(ctx, mode, a);  // warm-up
    if (r) return r;
    HIPCALL(ctx, hipEventRecord(e0, ctx->stream));
    for (int k = 0; k < 3; k++)
      if ((r = run_colored(ctx, mode, a))) return r;
    HIPCALL(ctx, hipEventRecord(e1, ctx->stream));
    HIPCALL(ctx, hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCALL(ctx, hipEventElapsedTime(&ms, e0, e1));
    best_ms = ms / 3.0f;
    return 0;
  };
  if ((rc = shellgpu_mat_create(ctx, ctx->nbatch, &A))) goto out;
  if (hipMalloc((void **)&u, sizeof(double) * nv) != hipSuccess ||
      hipMalloc((void **)&R, sizeof(double) * nv) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = set_err(ctx, 2, "tuner allocation failed");
    goto out;
  }
  if (hipMemcpy(u, hu.data(), sizeof(double) * nv, hipMemcpyHostToDevice) != hipSuccess) {
    rc = set_err(ctx, 2, "tuner state upload failed");
    goto out;
  }
  // colouring (flags bit 0): the stiffness assembly under each strategy
  if (flags & 1) {
    char buf[160];
    float t[2] = {0.0f, 0.0f};
    int nc[2] = {0, 0};
    for (int st = 0; st < 2; st++) {
      if ((rc = apply_coloring(ctx, st))) goto out;
      nc[st] = ctx->ncolors;
      if ((rc = trial(kModeKLin, t[st]))) goto out;
    }
    const int best = t[1] < t[0] ? 1 : 0;
    if ((rc = apply_coloring(ctx, best))) goto out;
    snprintf(buf, sizeof(buf),
             "\"coloring\": {\"chosen\": \"%s\", \"colors\": %d, \"us\": {\"element_order\": %.1f, "
             "\"most_neighbours_first\": %.1f}}, ",
             best ? "most_neighbours_first" : "element_order", nc[best], 1e3 * t[0], 1e3 * t[1]);
    rep += buf;
  }
  // compiled schedule of every matrix mode
  rep += "\"elem\": {";
  for (int c = 0; c < kTunedCount; c++) {
    const int mode = kTunedModes[c];
    float t[kElemVariants];
    int best = 0;
    for (int v = 0; v < kElemVariants; v++) {
      ctx->elem_variant[mode] = v;
      if ((rc = trial(mode, t[v]))) goto out;
      if (t[v] < t[best]) best = v;
    }
    ctx->elem_variant[mode] = best;
    char buf[512];
    int len = snprintf(buf, sizeof(buf), "%s\"%s\": {\"schedule\": %d, \"us\": [", c ? ", " : "", kTunedNames[c],
                       best);
    for (int v = 0; v < kElemVariants; v++)
      len += snprintf(buf + len, sizeof(buf) - len, "%s%.1f", v ? ", " : "", 1e3 * t[v]);
    snprintf(buf + len, sizeof(buf) - len, "]}");
    rep += buf;
  }
  rep += "}, \"schedules\": [\"block reads before the quadrature loop\", \"block reads at the scatter\", "
         "\"reads before the loop, 2 wavefronts/SIMD register budget\", "
         "\"reads at the scatter, 2 wavefronts/SIMD register budget\", "
         "\"wide workgroups, reads before the loop\", \"wide workgroups, reads at the scatter\", "
         "\"wide workgroups, reads before the loop, 2 wavefronts/SIMD\", "
         "\"wide workgroups, reads at the scatter, 2 wavefronts/SIMD\"]}";
  ctx->schedule = rep;
out:
  ctx->timing = timing;
  (void)hipStreamSynchronize(ctx->stream);
  if (A) shellgpu_mat_destroy(A);
  if (u) (void)hipFree(u);
  if (R) (void)hipFree(R);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}

int shellgpu_get_schedule(shellgpu_ctx *ctx, char *buf, int len) {
  if (!ctx || !buf || len <= 0) return 1;
  snprintf(buf, len, "%s", ctx->schedule.empty() ? "{}" : ctx->schedule.c_str());
  return (int)ctx->schedule.size() >= len ? set_err(ctx, 1, "buffer too small for the schedule") : 0;
}

int shellgpu_set_timing(shellgpu_ctx *ctx, int enable) {
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  drop_events(ctx->elem_ev);
  drop_events(ctx->spmv_ev);
  ctx->elem_launches = 0;
  ctx->timing = enable != 0;
  return 0;
}

int shellgpu_get_timing(shellgpu_ctx *ctx, double *elem_ms, int *elem_launches, double *spmv_ms,
                        int *spmv_launches) {
  HIPCALL(ctx, hipStreamSynchronize(ctx->stream));
  double te = 0.0, ts = 0.0;
  for (auto &p : ctx->elem_ev) {
    float ms = 0.0f;
    HIPCALL(ctx, hipEventElapsedTime(&ms, p.first, p.second));
    te += ms;
  }
  for (auto &p : ctx->spmv_ev) {
    float ms = 0.0f;
    HIPCALL(ctx, hipEventElapsedTime(&ms, p.first, p.second));
    ts += ms;
  }
  if (elem_ms) *elem_ms = te;
  if (elem_launches) *elem_launches = ctx->elem_launches;
  if (spmv_ms) *spmv_ms = ts;
  if (spmv_launches) *spmv_launches = (int)ctx->spmv_ev.size();
  drop_events(ctx->elem_ev);
  drop_events(ctx->spmv_ev);
  ctx->elem_launches = 0;
  return 0;
}

int shellgpu_timer_start(shellgpu_ctx *ctx) {
  if (!ctx->t0) HIPCALL(ctx, hipEventCreate(&ctx->t0));
  if (!ctx->t1) HIPCALL(ctx, hipEventCreate(&ctx->t1));
  HIPCALL(ctx, hipEventRecord(ctx->t0, ctx->stream));
  return 0;
}

int shellgpu_timer_stop(shellgpu_ctx *ctx, double *ms) {
  if (!ctx->t0) return set_err(ctx, 1, "timer not started");
  HIPCALL(ctx, hipEventRecord(ctx->t1, ctx->stream));
  HIPCALL(ctx, hipEventSynchronize(ctx->t1));
  float f = 0.0f;
  HIPCALL(ctx, hipEventElapsedTime(&f, ctx->t0, ctx->t1));
  *ms = f;
  return 0;
}

}  // extern "C"
