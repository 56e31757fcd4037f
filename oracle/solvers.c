/*
 * solvers.c -- oracle restatement of the reference's solve path: static
 * solve, geometric stiffness at the path, shift-invert Lanczos with full
 * modified Gram-Schmidt in the K inner product, tridiagonal eigenproblem.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference factors with an AMD-ordered complete block LU (TACSSchurPc,
 * levFill = 1e6, src/bpmat/TACSSchurMat.cpp:944-1652) and calls LAPACK
 * dstevr for the Lanczos tridiagonal (src/bpmat/GSEP.cpp:259-370); neither
 * is reproduced operation by operation.  The oracle uses an RCM-ordered band
 * LU (exact direct solve, round-off level agreement) and textbook implicit
 * QL.  Pinned by: dense scipy solves / generalized eigensolves on small
 * meshes and the closed-form spectrum of the Toeplitz tridiagonal
 * (tests/test_oracle_solvers.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---------------- RCM ordering of the node graph ---------------- */
static int bfs_levels(int n, const int *rowp, const int *cols, int start, int *level, int *queue) {
  for (int i = 0; i < n; i++) level[i] = -1;
  int head = 0, tail = 0, maxl = 0;
  level[start] = 0;
  queue[tail++] = start;
  while (head < tail) {
    int v = queue[head++];
    for (int p = rowp[v]; p < rowp[v + 1]; p++) {
      int w = cols[p];
      if (level[w] < 0) {
        level[w] = level[v] + 1;
        if (level[w] > maxl) maxl = level[w];
        queue[tail++] = w;
      }
    }
  }
  return maxl;
}

void or_rcm(int n, const int *rowp, const int *cols, int *perm) {
  int *level = (int *)malloc(sizeof(int) * (size_t)n);
  int *queue = (int *)malloc(sizeof(int) * (size_t)n);
  int *order = (int *)malloc(sizeof(int) * (size_t)n);
  char *seen = (char *)calloc((size_t)n, 1);
  int cnt = 0;
  for (int comp = 0; comp < n; comp++) {
    if (seen[comp]) continue;
    /* pseudo-peripheral node (George-Liu): repeat BFS from the min-degree
     * node of the last level while the eccentricity grows */
    int start = comp, ecc = bfs_levels(n, rowp, cols, start, level, queue);
    for (int it = 0; it < 8; it++) {
      int best = -1, bdeg = 1 << 30;
      for (int i = 0; i < n; i++)
        if (level[i] == ecc) {
          int d = rowp[i + 1] - rowp[i];
          if (d < bdeg) {
            bdeg = d;
            best = i;
          }
        }
      int e2 = bfs_levels(n, rowp, cols, best, level, queue);
      if (e2 <= ecc) break;
      ecc = e2;
      start = best;
    }
    /* Cuthill-McKee from start, neighbours by increasing degree */
    int head = cnt;
    order[cnt++] = start;
    seen[start] = 1;
    while (head < cnt) {
      int v = order[head++];
      int first = cnt;
      for (int p = rowp[v]; p < rowp[v + 1]; p++) {
        int w = cols[p];
        if (!seen[w]) {
          seen[w] = 1;
          order[cnt++] = w;
        }
      }
      for (int a = first + 1; a < cnt; a++) { /* insertion sort by degree */
        int w = order[a], dw = rowp[w + 1] - rowp[w], b = a - 1;
        while (b >= first && rowp[order[b] + 1] - rowp[order[b]] > dw) {
          order[b + 1] = order[b];
          b--;
        }
        order[b + 1] = w;
      }
    }
  }
  for (int i = 0; i < n; i++) perm[i] = order[n - 1 - i]; /* reverse */
  free(level);
  free(queue);
  free(order);
  free(seen);
}

/* ---------------- band LU (no pivoting) in RCM order ---------------- */
struct or_band {
  int n;        /* scalar dofs */
  int bw;       /* half bandwidth */
  double *a;    /* row i stored at a[i*(2bw+1) + (j - i + bw)] */
  int *pos_of;  /* new position of old dof */
};

or_band *or_band_factor(int nnodes, const int *rowp, const int *cols, const double *vals, const int *perm) {
  or_band *f = (or_band *)calloc(1, sizeof(or_band));
  const int n = 6 * nnodes;
  f->n = n;
  int *inv = (int *)malloc(sizeof(int) * (size_t)nnodes);
  for (int i = 0; i < nnodes; i++) inv[perm[i]] = i; /* node perm[i] goes to slot i */
  f->pos_of = (int *)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < nnodes; i++)
    for (int d = 0; d < 6; d++) f->pos_of[6 * i + d] = 6 * inv[i] + d;
  int bwn = 0;
  for (int i = 0; i < nnodes; i++)
    for (int p = rowp[i]; p < rowp[i + 1]; p++) {
      int d = abs(inv[i] - inv[cols[p]]);
      if (d > bwn) bwn = d;
    }
  f->bw = 6 * bwn + 5;
  const int w = 2 * f->bw + 1;
  f->a = (double *)calloc((size_t)n * (size_t)w, sizeof(double));
  for (int i = 0; i < nnodes; i++)
    for (int p = rowp[i]; p < rowp[i + 1]; p++) {
      const double *blk = &vals[36 * (size_t)p];
      for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) {
          int R = f->pos_of[6 * i + r], Cc = f->pos_of[6 * cols[p] + c];
          f->a[(size_t)R * w + (Cc - R + f->bw)] = blk[6 * r + c];
        }
    }
  const int bw = f->bw;
  for (int k = 0; k < n; k++) {
    double *rk = &f->a[(size_t)k * w];
    const double piv = rk[bw];
    const int iend = (k + bw < n - 1) ? k + bw : n - 1;
    for (int i = k + 1; i <= iend; i++) {
      double *ri = &f->a[(size_t)i * w];
      double l = ri[k - i + bw];
      if (l == 0.0) continue;
      l /= piv;
      ri[k - i + bw] = l;
      const int jend = (k + bw < n - 1) ? k + bw : n - 1;
      for (int j = k + 1; j <= jend; j++) ri[j - i + bw] -= l * rk[j - k + bw];
    }
  }
  free(inv);
  return f;
}

void or_band_solve(const or_band *f, const double *b, double *x) {
  const int n = f->n, bw = f->bw, w = 2 * bw + 1;
  double *y = (double *)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++) y[f->pos_of[i]] = b[i];
  for (int i = 0; i < n; i++) {
    const double *ri = &f->a[(size_t)i * w];
    double s = y[i];
    int j0 = i - bw < 0 ? 0 : i - bw;
    for (int j = j0; j < i; j++) s -= ri[j - i + bw] * y[j];
    y[i] = s;
  }
  for (int i = n - 1; i >= 0; i--) {
    const double *ri = &f->a[(size_t)i * w];
    double s = y[i];
    int j1 = i + bw > n - 1 ? n - 1 : i + bw;
    for (int j = i + 1; j <= j1; j++) s -= ri[j - i + bw] * y[j];
    y[i] = s / ri[bw];
  }
  for (int i = 0; i < n; i++) x[i] = y[f->pos_of[i]];
  free(y);
}

void or_band_free(or_band *f) {
  if (!f) return;
  free(f->a);
  free(f->pos_of);
  free(f);
}

/* ---------------- symmetric tridiagonal eigenproblem ---------------------
 * Implicit QL with Wilkinson shifts (the EISPACK tql2 algorithm, Bowdler,
 * Martin, Reinsch & Wilkinson, Numer. Math. 11 (1968) 293-306). */
int or_tridiag_eig(int n, const double *d_in, const double *e_in, double *d, double *z) {
  double *e = (double *)calloc((size_t)n + 1, sizeof(double));
  for (int i = 0; i < n; i++) d[i] = d_in[i];
  for (int i = 0; i < n - 1; i++) e[i] = e_in[i];
  for (int i = 0; i < n * n; i++) z[i] = 0.0;
  for (int i = 0; i < n; i++) z[i * n + i] = 1.0; /* z[k*n + i]: comp i of vector k */
  for (int l = 0; l < n; l++) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; m++) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= 1e-16 * dd) break;
      }
      if (m != l) {
        if (iter++ == 60) {
          free(e);
          return -1;
        }
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; i--) {
          double f = s * e[i], b = c * e[i];
          r = hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          for (int k = 0; k < n; k++) {
            double t = z[(i + 1) * n + k];
            z[(i + 1) * n + k] = s * z[i * n + k] + c * t;
            z[i * n + k] = c * z[i * n + k] - s * t;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  /* sort ascending */
  for (int i = 0; i < n - 1; i++) {
    int k = i;
    for (int j = i + 1; j < n; j++)
      if (d[j] < d[k]) k = j;
    if (k != i) {
      double t = d[i];
      d[i] = d[k];
      d[k] = t;
      for (int q = 0; q < n; q++) {
        double tz = z[i * n + q];
        z[i * n + q] = z[k * n + q];
        z[k * n + q] = tz;
      }
    }
  }
  free(e);
  return 0;
}

/* ---------------- linear buckling ---------------- */
static double dotv(int n, const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* splitmix64 -> uniform [-1, 1) start vector (the reference uses rand(),
 * TACSBVec::setRand, src/bpmat/TACSBVec.cpp:462-491) */
static double rnd(unsigned long long *s) {
  unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return 2.0 * ((double)(z >> 11) / 9007199254740992.0) - 1.0;
}

int or_linear_buckling(const or_buckling_problem *p, double *lambda, double *phi, double *path, double *resid) {
  const int nn = p->nnodes, n = 6 * nn;
  int *rowp, *cols;
  int nnzb = or_node_csr(nn, p->nelems, p->conn, &rowp, &cols);
  double *K = (double *)malloc(sizeof(double) * 36 * (size_t)nnzb);
  double *G = (double *)malloc(sizeof(double) * 36 * (size_t)nnzb);
  double *A = (double *)malloc(sizeof(double) * 36 * (size_t)nnzb);
  int *perm = (int *)malloc(sizeof(int) * (size_t)nn);
  double *u = (double *)calloc((size_t)n, sizeof(double));
  double *rhs = (double *)malloc(sizeof(double) * (size_t)n);
  int nsteps = -1;

  /* TACSLinearBuckling::solve (src/TACSBuckling.cpp:236-271) with u0 = NULL */
  or_assemble_mat(0, OR_LINEAR, p->mat, nn, p->nelems, p->conn, p->Xpts, p->thick, p->temp, NULL, rowp, cols, K);
  or_bsr_apply_bcs(nn, rowp, cols, K, p->nbc, p->bc_nodes, p->bc_mask, 1.0);
  or_rcm(nn, rowp, cols, perm);
  or_band *fk = or_band_factor(nn, rowp, cols, K, perm);
  if (p->u0) {
    /* given path with the BC values set (TACSBVec::setBCs) */
    memcpy(u, p->u0, sizeof(double) * (size_t)n);
    or_vec_zero_bcs(u, p->nbc, p->bc_nodes, p->bc_mask);
    for (int b = 0; b < p->nbc && p->bc_vals; b++)
      for (int k = 0; k < 6; k++)
        if (p->bc_mask[b] >> k & 1) u[6 * p->bc_nodes[b] + k] = p->bc_vals[6 * b + k];
  } else {
    /* res = R(0) - f with BC rows u - ubar = -ubar; path = -K^{-1} res */
    or_assemble_res(OR_LINEAR, p->mat, nn, p->nelems, p->conn, p->Xpts, p->thick, p->temp, NULL, rhs);
    for (int i = 0; i < n; i++) rhs[i] -= p->force ? p->force[i] : 0.0;
    or_vec_zero_bcs(rhs, p->nbc, p->bc_nodes, p->bc_mask);
    for (int b = 0; b < p->nbc && p->bc_vals; b++)
      for (int k = 0; k < 6; k++)
        if (p->bc_mask[b] >> k & 1) rhs[6 * p->bc_nodes[b] + k] = -p->bc_vals[6 * b + k];
    or_band_solve(fk, rhs, u);
    for (int i = 0; i < n; i++) u[i] = -u[i];
  }
  or_band_free(fk);
  if (path) memcpy(path, u, sizeof(double) * (size_t)n);

  /* G at the path (central difference of the nonlinear tangent, accurate
   * step), aux = K + sigma G with BCs */
  or_assemble_mat(3, OR_NONLINEAR, p->mat, nn, p->nelems, p->conn, p->Xpts, p->thick, p->temp, u, rowp, cols, G);
  or_bsr_apply_bcs(nn, rowp, cols, G, p->nbc, p->bc_nodes, p->bc_mask, 1.0);
  for (size_t i = 0; i < 36 * (size_t)nnzb; i++) A[i] = K[i] + p->sigma * G[i];
  or_bsr_apply_bcs(nn, rowp, cols, A, p->nbc, p->bc_nodes, p->bc_mask, 1.0);
  or_band *fa = or_band_factor(nn, rowp, cols, A, perm);

  /* SEP::solve, full MGS in the K inner product (GSEP.cpp:485-599) */
  const int m = p->max_lanczos;
  double *Q = (double *)malloc(sizeof(double) * (size_t)n * (size_t)(m + 1));
  double *KQ = (double *)malloc(sizeof(double) * (size_t)n * (size_t)(m + 1));
  double *alpha = (double *)calloc((size_t)m + 1, sizeof(double));
  double *beta = (double *)calloc((size_t)m + 1, sizeof(double));
  double *w = (double *)malloc(sizeof(double) * (size_t)n);
  double *eig = (double *)malloc(sizeof(double) * (size_t)m);
  double *z = (double *)malloc(sizeof(double) * (size_t)m * (size_t)m);
  unsigned long long seed = p->seed;
  for (int i = 0; i < n; i++) Q[i] = rnd(&seed);
  or_vec_zero_bcs(Q, p->nbc, p->bc_nodes, p->bc_mask);
  or_bsr_matvec(nn, rowp, cols, K, Q, KQ);
  double nrm = sqrt(dotv(n, Q, KQ));
  for (int i = 0; i < n; i++) {
    Q[i] /= nrm;
    KQ[i] /= nrm;
  }
  int k;
  for (k = 0; k < m; k++) {
    double *qk = &Q[(size_t)k * n], *qn = &Q[(size_t)(k + 1) * n], *Kqn = &KQ[(size_t)(k + 1) * n];
    or_band_solve(fa, &KQ[(size_t)k * n], qn);
    or_vec_zero_bcs(qn, p->nbc, p->bc_nodes, p->bc_mask);
    for (int j = k; j >= 0; j--) {
      double h = dotv(n, qn, &KQ[(size_t)j * n]);
      for (int i = 0; i < n; i++) qn[i] -= h * Q[(size_t)j * n + i];
      if (j == k) alpha[k] = h;
    }
    /* second pass of orthogonalisation against all previous vectors */
    for (int j = k; j >= 0; j--) {
      double h = dotv(n, qn, &KQ[(size_t)j * n]);
      for (int i = 0; i < n; i++) qn[i] -= h * Q[(size_t)j * n + i];
      if (j == k) alpha[k] += h;
    }
    or_bsr_matvec(nn, rowp, cols, K, qn, Kqn);
    beta[k] = sqrt(dotv(n, qn, Kqn));
    for (int i = 0; i < n; i++) {
      qn[i] /= beta[k];
      Kqn[i] /= beta[k];
    }
    (void)qk;
  }
  nsteps = k;
  if (or_tridiag_eig(nsteps, alpha, beta, eig, z) != 0) {
    nsteps = -2;
    goto done;
  }
  {
    /* theta -> lambda = -theta sigma / (1 - theta) (GSEP.cpp:242-244), sort
     * by magnitude of lambda (SMALLEST_MAGNITUDE), keep nev */
    int *idx = (int *)malloc(sizeof(int) * (size_t)nsteps);
    double *lam = (double *)malloc(sizeof(double) * (size_t)nsteps);
    for (int i = 0; i < nsteps; i++) {
      idx[i] = i;
      lam[i] = -eig[i] * p->sigma / (1.0 - eig[i]);
    }
    for (int i = 0; i < nsteps; i++)
      for (int j = i + 1; j < nsteps; j++)
        if (fabs(lam[idx[j]]) < fabs(lam[idx[i]])) {
          int t = idx[i];
          idx[i] = idx[j];
          idx[j] = t;
        }
    double *v = (double *)malloc(sizeof(double) * (size_t)n);
    double *Kv = (double *)malloc(sizeof(double) * (size_t)n);
    double *Gv = (double *)malloc(sizeof(double) * (size_t)n);
    for (int e = 0; e < p->nev && e < nsteps; e++) {
      const int s = idx[e];
      lambda[e] = lam[s];
      memset(v, 0, sizeof(double) * (size_t)n);
      for (int j = 0; j < nsteps; j++) {
        const double c = z[(size_t)s * nsteps + j];
        for (int i = 0; i < n; i++) v[i] += c * Q[(size_t)j * n + i];
      }
      if (phi) memcpy(&phi[(size_t)e * n], v, sizeof(double) * (size_t)n);
      if (resid) {
        or_bsr_matvec(nn, rowp, cols, K, v, Kv);
        or_bsr_matvec(nn, rowp, cols, G, v, Gv);
        or_vec_zero_bcs(Kv, p->nbc, p->bc_nodes, p->bc_mask);
        or_vec_zero_bcs(Gv, p->nbc, p->bc_nodes, p->bc_mask);
        double num = 0.0, den = 0.0;
        for (int i = 0; i < n; i++) {
          double r = Kv[i] + lam[s] * Gv[i];
          num += r * r;
          den += Kv[i] * Kv[i];
        }
        resid[e] = sqrt(num / den);
      }
    }
    free(v);
    free(Kv);
    free(Gv);
    free(idx);
    free(lam);
  }
done:
  or_band_free(fa);
  free(Q);
  free(KQ);
  free(alpha);
  free(beta);
  free(w);
  free(eig);
  free(z);
  free(rowp);
  free(cols);
  free(K);
  free(G);
  free(A);
  free(perm);
  free(u);
  free(rhs);
  return nsteps;
}
