/*
 * oracle.h -- CPU FP64 restatement of the reference's shell-element assembly
 * and linear-buckling path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
 * library; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, and only as the checker / the timed CPU baseline.
 *
 * Every function follows the reference file:line cited at its definition.
 * Paths are relative to the reference repository (sean-engelstad/a2d-shells):
 *   SE.h  = src/elements/shell/TACSShellElement.h
 *   SU.h  = src/elements/shell/TACSShellUtilities.h
 *   SM.h  = src/elements/shell/TACSShellElementModel.h
 *   SD.h  = src/elements/shell/TACSDirector.h
 *   SB.h  = src/elements/shell/TACSShellElementQuadBasis.h
 *   ST.h  = src/elements/shell/TACSShellElementTransform.h
 *   EA.h  = src/elements/TACSElementAlgebra.h
 *   ISC   = src/constitutive/TACSIsoShellConstitutive.cpp
 *   TA    = src/TACSAssembler.cpp
 *
 * Parity status of each function is recorded in the "pinned by" line of its
 * comment in the .c files.
 */
#ifndef SHELL_ORACLE_H
#define SHELL_ORACLE_H

#if defined(__cplusplus) && !defined(OR_NO_EXTERN_C)
extern "C" {
#endif

/* Isotropic shell material + section (TACSMaterialProperties +
 * TACSIsoShellConstitutive, tOffset = 0, kcorr = 5/6). */
typedef struct {
  double E, nu, rho, cte;
  double k_drill; /* TACSShellConstitutive::DRILLING_REGULARIZATION (10.0) */
} or_material;

/* Element model: 0 = TACSShellLinearModel, 1 = TACSShellNonlinearModel */
enum { OR_LINEAR = 0, OR_NONLINEAR = 1 };

/* ---------------- element level (4-node MITC director shell) ------------- */

/* R_e(q) for statics (dvars = ddvars = 0), additive into res[24].
 * SE.h:303-449 */
void or_elem_residual(int model, const or_material *m, double t, double temp,
                      const double X[12], const double q[24], double res[24]);

/* R_e and K_e = alpha*dR/dq (beta = gamma = 0), additive into res, mat.
 * SE.h:455-672 */
void or_elem_jacobian(int model, const or_material *m, double t, double temp,
                      double alpha, const double X[12], const double q[24],
                      double res[24], double mat[576]);

/* getMatType: STIFFNESS (K with the element's own model at q) or the
 * geometric stiffness by the reference's central difference of the nonlinear
 * tangent (h = 1e-4).  mat is overwritten.  SE.h:675-771 */
void or_elem_stiffness(int model, const or_material *m, double t, double temp,
                       const double X[12], const double q[24],
                       double mat[576]);
/* getOutputData(TACS_BEAM_OR_SHELL_ELEMENT, STRAINS | STRESSES | EXTRAS):
 * 31 values at each of the 4 nodes (strain 9, stress 9, failure 6, design
 * 7), ys the yield stress.  SE.h:1402-1500; ISC:153-189, 254-297, 503-512 */
void or_elem_output(int model, const or_material *m, double t, double ys, const double X[12],
                    const double q[24], double out[124]);
/* getMatType(TACS_MASS_MATRIX): consistent mass of the element, mat
 * overwritten.  SE.h:613-648, 675-704; SD.h:368-430; ISC:120-129 */
void or_elem_mass(const or_material *m, double t, const double X[12], double mat[576]);
void or_elem_geometric_fd(const or_material *m, double t, double temp,
                          const double X[12], const double u[24],
                          double mat[576]);
/* same central-difference formula with step dh (dh = 1: round-off-level
 * accurate value of the same matrix, the tangent being quadratic in q) */
void or_elem_geometric_step(const or_material *m, double t, double temp,
                            const double X[12], const double u[24], double dh,
                            double mat[576]);

/* or_elem_geometric_step evaluated in long double (shell_elem_ld.cpp): the
 * accurate geometric stiffness where the double evaluation's round-off
 * matters (fine meshes) */
void or_elem_geometric_step_ld(const or_material *m, double t, double temp,
                               const double X[12], const double u[24], double dh,
                               double mat[576]);

/* dfdt += scale * psi^T dR_e/dt (thickness adjoint-residual product).
 * SE.h:774-868 with ISC:229-251. */
void or_elem_adj_res_dv(int model, const or_material *m, double t, double temp,
                        double scale, const double psi[24], const double X[12],
                        const double q[24], double *dfdt);

/* Strain energy U_e (computeEnergies, SE.h:202-297). */
double or_elem_energy(int model, const or_material *m, double t, double temp,
                      const double X[12], const double q[24]);

/* ---------------- assembly level ---------------- */

/* Node-to-node CSR with sorted, unique columns, diagonal kept.
 * TA:1731-1992 + src/utils/TacsUtilities.cpp:280-311.  rowp[nnodes+1];
 * cols allocated with malloc (caller frees). Returns nnzb. */
int or_node_csr(int nnodes, int nelems, const int *conn, int **rowp,
                int **cols);

/* First-appearance node renumbering of TACSCreator (src/TACSCreator.cpp:
 * 1141-1202, single partition).  new_of_old[nnodes]. */
void or_first_appearance(int nnodes, int nelems, const int *conn,
                         int *new_of_old);

/* Global assembly in element order with the reference's block scatter
 * (TACSAssembler.h:453-507 -> BCSRMat::addRowValues, BCSRMat.cpp:1778-1855).
 * vals is nnzb*36, block row-major.  kind: 0 = K (getMatType STIFFNESS),
 * 1 = G (geometric, FD step 1e-4), 2 = nonlinear Jacobian at u (alpha = 1),
 * 3 = G by the same formula with a unit step (accurate value when T = 0),
 * 4 = the same with step 1e5 (accurate value when a temperature dominates the
 * state norm), 5 = the mass matrix (getMatType(TACS_MASS_MATRIX)).
 * thick[nelems] per-element thickness; temp per-element temperature (may be
 * NULL).  u is the 6*nnodes state (may be NULL = 0).  Zeroes vals first. */
void or_assemble_mat(int kind, int model, const or_material *m, int nnodes,
                     int nelems, const int *conn, const double *Xpts,
                     const double *thick, const double *temp,
                     const double *u, const int *rowp, const int *cols,
                     double *vals);

/* Residual assembly (TA:4000-4070), additive-free: R is overwritten. */
void or_assemble_res(int model, const or_material *m, int nnodes, int nelems,
                     const int *conn, const double *Xpts, const double *thick,
                     const double *temp, const double *u, double *R);

/* BCSRMat::zeroRow (BCSRMat.cpp:2005-2034) for every BC node: rows of the
 * masked dofs are zeroed and the diagonal set to diag (1 for applyBCs). */
void or_bsr_apply_bcs(int nnodes, const int *rowp, const int *cols,
                      double *vals, int nbc, const int *bc_nodes,
                      const int *bc_mask, double diag);

/* TACSBVec::applyBCs (TACSBVec.cpp:546-575): zero masked entries, or with
 * vals != NULL set x[bc] = u[bc] - vals (the residual form). */
void or_vec_zero_bcs(double *x, int nbc, const int *bc_nodes,
                     const int *bc_mask);

/* y = A x, 6x6 blocks (BCSRMatVecMult6, BCSRMatMult6.cpp:82-125). */
void or_bsr_matvec(int nnodes, const int *rowp, const int *cols,
                   const double *vals, const double *x, double *y);

/* ---------------- solvers ---------------- */

/* Dense-band LU of a BSR matrix under a given node permutation (RCM from
 * or_rcm); used as the oracle's direct solver in place of the reference's
 * AMD-ordered complete block LU (TACSSchurPc with levFill = 1e6). */
typedef struct or_band or_band;
or_band *or_band_factor(int nnodes, const int *rowp, const int *cols,
                        const double *vals, const int *perm);
void or_band_solve(const or_band *f, const double *b, double *x);
void or_band_free(or_band *f);
void or_rcm(int nnodes, const int *rowp, const int *cols, int *perm);

/* Symmetric tridiagonal eigen-decomposition by implicit QL (textbook), in
 * place of LAPACK dstevr in SEP::checkConverged (GSEP.cpp:259-370).
 * d[n] diagonal, e[n-1] sub-diagonal, eigs[n] ascending, z[n*n] columns
 * z[k*n + i] = component i of eigenvector k. Returns 0 on success. */
int or_tridiag_eig(int n, const double *d, const double *e, double *eigs,
                   double *z);

/* Linear buckling (TACSLinearBuckling::solve, src/TACSBuckling.cpp:197-278) with shift-invert Lanczos, full MGS in the K
 * inner product (SEP::solve, GSEP.cpp:485-599), direct solves. */
typedef struct {
  int nnodes, nelems;
  const int *conn;
  const double *Xpts;
  const double *thick;
  const double *temp; /* may be NULL */
  const or_material *mat;
  int nbc;
  const int *bc_nodes, *bc_mask;
  const double *force; /* 6*nnodes */
  double sigma;
  int max_lanczos; /* Lanczos steps (the reference stops at num_eigvals) */
  int nev;
  unsigned long long seed; /* start vector */
  const double *u0; /* path (TACSLinearBuckling::solve with u0 != NULL); NULL:
                       static solve of K u = f */
  const double *bc_vals; /* 6 per BC entry (prescribed values); NULL = 0 */
} or_buckling_problem;

/* lambda[nev] (sorted by magnitude ascending, SEP sortEigenvalues
 * SMALLEST_MAGNITUDE), phi[nev*6*nnodes] (may be NULL), path u[6*nnodes]
 * (may be NULL), residual norms resid[nev] of K phi + lambda G phi (NULL ok).
 * Returns number of Lanczos steps taken, < 0 on error. */
int or_linear_buckling(const or_buckling_problem *p, double *lambda,
                       double *phi, double *path, double *resid);

#if defined(__cplusplus) && !defined(OR_NO_EXTERN_C)
}
#endif
#endif
