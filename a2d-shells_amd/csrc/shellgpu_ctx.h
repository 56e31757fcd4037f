// shellgpu_ctx.h -- the context and matrix objects behind the C ABI
// (host-side state shared by the translation units of the library).
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "shellgpu_internal.h"

struct shellgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  int nnodes = 0, nelems = 0, nnzb = 0, ncolors = 0, nbatch = 0;
  std::vector<int> h_conn, h_rowp, h_cols, h_color, h_color_ptr, h_color_elems, h_blk;
  int *d_conn = nullptr, *d_rowp = nullptr, *d_cols = nullptr, *d_blk = nullptr;
  // per node the block index of its diagonal block; per block the index of
  // the transposed block (the pattern is structurally symmetric)
  int *d_diag = nullptr, *d_tpos = nullptr;
  int *d_color_elems = nullptr;
  // per element, bit 4 i + j: this element is the first, in colour launch
  // order, to reach BSR block (node i, node j), so a matrix assembly can
  // write that block instead of adding to a zeroed one; null when an element
  // repeats a node (two of its blocks coincide)
  int *d_first = nullptr;
  double *d_X = nullptr;

  // element colouring: 0 greedy first fit in element order, 1 greedy first
  // fit with the elements of most neighbours first (shellgpu_tune chooses)
  int coloring = 0;
  // compiled schedule of each element-kernel mode (launch_elem_kernel's
  // variant; shellgpu_tune chooses) and the tuner's report
  int elem_variant[32] = {0};
  std::string schedule;

  shellgpu::Material mat = shellgpu::make_material(70e9, 0.3, 2700.0, 0.0, 10.0);

  int nbc = 0;
  std::vector<int> h_bc_nodes, h_bc_mask;
  int *d_bc_nodes = nullptr, *d_bc_mask = nullptr;
  double *d_bc_vals = nullptr;

  double *d_thick = nullptr, *d_temp = nullptr;
  bool same_designs = false;  // every design has the same thickness and temperature
  double *d_coef = nullptr;  // 4 * nbatch per-design kernel coefficients
  // timing
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> elem_ev, spmv_ev;
  int elem_launches = 0;  // element-kernel launches inside the elem_ev pairs
  hipEvent_t t0 = nullptr, t1 = nullptr;

  // the device's copy of the library's static scratch array (mapped at
  // shellgpu_create; shared by every context on the device) and its size
  double *scratch = nullptr;
  size_t scratch_doubles = 0;

  // solver state
  std::vector<int> order_seed;  // BFS seed nodes of the level ordering
  void *blas = nullptr;         // rocblas_handle bound to `stream`
};

struct shellgpu_mat {
  shellgpu_ctx *ctx = nullptr;
  int nbatch = 0;
  double *vals = nullptr;
  // every design's values are equal (zeroed, or assembled from identical
  // designs without a per-design state, and changed only uniformly since):
  // a product then reads one copy for all designs' vectors
  bool same = false;
};

namespace shellgpu {
// element-kernel arguments for the context's mesh, material and designs
// with state u (NULL = zero)
ElemArgs base_args(shellgpu_ctx *ctx, const double *u);
// launch one element-kernel mode colour by colour (no two elements of a
// launch share a node, so scatters need no atomics)
int run_colored(shellgpu_ctx *ctx, int mode, ElemArgs a);
// coordinate-derivative kernel over all colours; coef (device, 4 per
// design), vectors with stride 6*nnodes, dfdX with stride 3*nnodes
int run_xpt(shellgpu_ctx *ctx, const double *phi, const double *u, const double *psi, const double *coef,
            int nonlin, int terms, double *dfdX);
}  // namespace shellgpu
