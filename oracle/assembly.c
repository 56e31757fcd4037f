/*
 * assembly.c -- oracle restatement of the reference's assembly path:
 * node-to-node BSR pattern, element loop + block scatter, BC application and
 * the 6x6 block SpMV.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Pinned by: brute-force node adjacency and dense scatter in
 * tests/test_oracle_assembly.py (the reference holds no assembled fixtures;
 * its assembler needs METIS and is not built here).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* TACSAssembler::computeNodeToElementCSR + computeLocalNodeToNodeCSR
 * (TA:1731-1992, no dependent nodes) and TacsSortAndUniquifyCSR
 * (src/utils/TacsUtilities.cpp:280-311) with the diagonal kept. */
int or_node_csr(int nnodes, int nelems, const int *conn, int **rowp_out, int **cols_out) {
  int *nep = (int *)calloc((size_t)nnodes + 1, sizeof(int));
  for (int e = 0; e < nelems; e++)
    for (int k = 0; k < 4; k++) nep[conn[4 * e + k] + 1]++;
  for (int i = 0; i < nnodes; i++) nep[i + 1] += nep[i];
  int *n2e = (int *)malloc(sizeof(int) * (size_t)(nep[nnodes] > 0 ? nep[nnodes] : 1));
  int *fill = (int *)malloc(sizeof(int) * (size_t)(nnodes > 0 ? nnodes : 1));
  memcpy(fill, nep, sizeof(int) * (size_t)nnodes);
  for (int e = 0; e < nelems; e++)
    for (int k = 0; k < 4; k++) n2e[fill[conn[4 * e + k]]++] = e;

  int *rowp = (int *)calloc((size_t)nnodes + 1, sizeof(int));
  int cap = 0;
  for (int i = 0; i < nnodes; i++) cap += 4 * (nep[i + 1] - nep[i]);
  int *cols = (int *)malloc(sizeof(int) * (size_t)(cap > 0 ? cap : 1));
  int nnz = 0;
  for (int i = 0; i < nnodes; i++) {
    int start = nnz;
    for (int p = nep[i]; p < nep[i + 1]; p++) {
      int e = n2e[p];
      for (int k = 0; k < 4; k++) cols[nnz++] = conn[4 * e + k];
    }
    qsort(&cols[start], (size_t)(nnz - start), sizeof(int), cmp_int);
    int w = start;
    for (int r = start; r < nnz; r++)
      if (r == start || cols[r] != cols[w - 1]) cols[w++] = cols[r];
    nnz = w;
    rowp[i + 1] = nnz;
  }
  free(nep);
  free(n2e);
  free(fill);
  *rowp_out = rowp;
  *cols_out = cols;
  return nnz;
}

/* TACSCreator node renumbering (src/TACSCreator.cpp:1141-1202): nodes are
 * numbered in order of first appearance in the element connectivity. */
void or_first_appearance(int nnodes, int nelems, const int *conn, int *new_of_old) {
  for (int i = 0; i < nnodes; i++) new_of_old[i] = -1;
  int next = 0;
  for (int e = 0; e < nelems; e++)
    for (int k = 0; k < 4; k++) {
      int n = conn[4 * e + k];
      if (new_of_old[n] < 0) new_of_old[n] = next++;
    }
}

/* BCSRMat::addRowValues for one element (BCSRMat.cpp:1778-1855): block
 * (i,j) entry (r,c) += Ke[(6i+r)*24 + 6j+c]; column located by binary
 * search in the sorted row (TacsSearchArray). */
static void scatter_elem(const int *nodes, const double *Ke, const int *rowp, const int *cols, double *vals) {
  for (int i = 0; i < 4; i++) {
    int row = nodes[i];
    const int *c0 = &cols[rowp[row]];
    int len = rowp[row + 1] - rowp[row];
    for (int j = 0; j < 4; j++) {
      int col = nodes[j];
      int lo = 0, hi = len - 1, pos = -1;
      while (lo <= hi) {
        int mid = (lo + hi) / 2;
        if (c0[mid] == col) {
          pos = mid;
          break;
        } else if (c0[mid] < col) {
          lo = mid + 1;
        } else {
          hi = mid - 1;
        }
      }
      if (pos < 0) continue; /* not in the pattern: the reference prints and skips */
      double *a = &vals[36 * (size_t)(rowp[row] + pos)];
      for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) a[6 * r + c] += Ke[(6 * i + r) * 24 + 6 * j + c];
    }
  }
}

static void gather_elem(const int *nodes, const double *Xpts, const double *u, double X[12], double q[24]) {
  for (int k = 0; k < 4; k++) {
    for (int d = 0; d < 3; d++) X[3 * k + d] = Xpts[3 * nodes[k] + d];
    for (int d = 0; d < 6; d++) q[6 * k + d] = u ? u[6 * nodes[k] + d] : 0.0;
  }
}

/* TACSAssembler::assembleMatType / assembleJacobian element loop, serial
 * (TA:4084-4250) */
void or_assemble_mat(int kind, int model, const or_material *m, int nnodes, int nelems, const int *conn,
                     const double *Xpts, const double *thick, const double *temp, const double *u,
                     const int *rowp, const int *cols, double *vals) {
  memset(vals, 0, sizeof(double) * 36 * (size_t)rowp[nnodes]);
  for (int e = 0; e < nelems; e++) {
    double X[12], q[24], Ke[576], res[24];
    const int *nodes = &conn[4 * e];
    gather_elem(nodes, Xpts, u, X, q);
    double T = temp ? temp[e] : 0.0;
    if (kind == 0) {
      or_elem_stiffness(model, m, thick[e], T, X, q, Ke);
    } else if (kind == 1) {
      or_elem_geometric_fd(m, thick[e], T, X, q, Ke);
    } else if (kind == 3) {
      or_elem_geometric_step(m, thick[e], T, X, q, 1.0, Ke);
    } else if (kind == 4) {
      or_elem_geometric_step(m, thick[e], T, X, q, 1e5, Ke);
    } else if (kind == 5) {
      or_elem_mass(m, thick[e], X, Ke);
    } else if (kind == 6) {
      or_elem_geometric_step_ld(m, thick[e], T, X, q, 1.0, Ke);
    } else {
      memset(Ke, 0, sizeof(Ke));
      memset(res, 0, sizeof(res));
      or_elem_jacobian(OR_NONLINEAR, m, thick[e], T, 1.0, X, q, res, Ke);
    }
    scatter_elem(nodes, Ke, rowp, cols, vals);
  }
}

/* TACSAssembler::assembleRes element loop (TA:4000-4070), no BCs */
void or_assemble_res(int model, const or_material *m, int nnodes, int nelems, const int *conn,
                     const double *Xpts, const double *thick, const double *temp, const double *u, double *R) {
  memset(R, 0, sizeof(double) * 6 * (size_t)nnodes);
  for (int e = 0; e < nelems; e++) {
    double X[12], q[24], re[24];
    const int *nodes = &conn[4 * e];
    gather_elem(nodes, Xpts, u, X, q);
    memset(re, 0, sizeof(re));
    or_elem_residual(model, m, thick[e], temp ? temp[e] : 0.0, X, q, re);
    for (int k = 0; k < 4; k++)
      for (int d = 0; d < 6; d++) R[6 * nodes[k] + d] += re[6 * k + d];
  }
}

/* BCSRMat::zeroRow (BCSRMat.cpp:2005-2034) over the BC map */
void or_bsr_apply_bcs(int nnodes, const int *rowp, const int *cols, double *vals, int nbc, const int *bc_nodes,
                      const int *bc_mask, double diag) {
  (void)nnodes;
  for (int b = 0; b < nbc; b++) {
    int row = bc_nodes[b], mask = bc_mask[b];
    for (int p = rowp[row]; p < rowp[row + 1]; p++) {
      double *a = &vals[36 * (size_t)p];
      for (int r = 0; r < 6; r++)
        if (mask & (1 << r))
          for (int c = 0; c < 6; c++) a[6 * r + c] = 0.0;
      if (cols[p] == row)
        for (int r = 0; r < 6; r++)
          if (mask & (1 << r)) a[7 * r] = diag;
    }
  }
}

/* TACSBVec::applyBCs without values (TACSBVec.cpp:546-575) */
void or_vec_zero_bcs(double *x, int nbc, const int *bc_nodes, const int *bc_mask) {
  for (int b = 0; b < nbc; b++)
    for (int r = 0; r < 6; r++)
      if (bc_mask[b] & (1 << r)) x[6 * bc_nodes[b] + r] = 0.0;
}

/* BCSRMatVecMult6 (BCSRMatMult6.cpp:82-125) */
void or_bsr_matvec(int nnodes, const int *rowp, const int *cols, const double *vals, const double *x, double *y) {
  for (int i = 0; i < nnodes; i++) {
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int p = rowp[i]; p < rowp[i + 1]; p++) {
      const double *a = &vals[36 * (size_t)p];
      const double *xx = &x[6 * cols[p]];
      for (int r = 0; r < 6; r++)
        acc[r] += a[6 * r] * xx[0] + a[6 * r + 1] * xx[1] + a[6 * r + 2] * xx[2] + a[6 * r + 3] * xx[3] +
                  a[6 * r + 4] * xx[4] + a[6 * r + 5] * xx[5];
    }
    for (int r = 0; r < 6; r++) y[6 * i + r] = acc[r];
  }
}
