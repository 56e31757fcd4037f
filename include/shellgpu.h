/*
 * shellgpu.h -- C ABI of the MI355X (gfx950) shell-assembly and linear
 * buckling library.
 *
 * Plain C types only: device data are passed as raw device pointers, host
 * set-up data as host pointers that are copied.  Every function returns 0 on
 * success and a non-zero code on failure; shellgpu_last_error() returns the
 * message.  A context owns one HIP stream; calls on one context are not
 * thread-safe and are ordered on that stream.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository sean-engelstad/a2d-shells).
 */
#ifndef SHELLGPU_H
#define SHELLGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct shellgpu_ctx shellgpu_ctx;
typedef struct shellgpu_mat shellgpu_mat;
typedef struct shellgpu_factor shellgpu_factor;
typedef struct shellgpu_buckling shellgpu_buckling;
typedef struct shellgpu_frequency shellgpu_frequency;

/* matrix types (ElementMatrixType, src/elements/TACSElementTypes.h) */
enum {
  SHELLGPU_STIFFNESS = 0,        /* TACS_STIFFNESS_MATRIX, linear model     */
  SHELLGPU_GEOMETRIC = 1,        /* TACS_GEOMETRIC_STIFFNESS_MATRIX, exact  */
  SHELLGPU_GEOMETRIC_FD = 2,     /* same, reference central difference      */
  SHELLGPU_TANGENT_NONLINEAR = 3, /* addJacobian(alpha=1) of the nonlinear model */
  SHELLGPU_MASS = 4               /* TACS_MASS_MATRIX (getMatType, gamma = 1)  */
};
/* element models (src/elements/shell/TACSShellElementDefs.h:12-29) */
enum { SHELLGPU_LINEAR = 0, SHELLGPU_NONLINEAR = 1 };

/* ---- context ---------------------------------------------------------- */
/* TACSAssembler constructor (src/TACSAssembler.cpp:40-120), one rank */
int shellgpu_create(shellgpu_ctx **ctx, int device);
void shellgpu_destroy(shellgpu_ctx *ctx);
const char *shellgpu_last_error(const shellgpu_ctx *ctx);
int shellgpu_synchronize(shellgpu_ctx *ctx);
/* the context's hipStream_t */
void *shellgpu_get_stream(shellgpu_ctx *ctx);

/* device memory helpers (plumbing for callers without a HIP runtime) */
int shellgpu_malloc(shellgpu_ctx *ctx, void **dptr, size_t bytes);
int shellgpu_free(shellgpu_ctx *ctx, void *dptr);
int shellgpu_memcpy_h2d(shellgpu_ctx *ctx, void *dst, const void *src, size_t bytes);
int shellgpu_memcpy_d2h(shellgpu_ctx *ctx, void *dst, const void *src, size_t bytes);
int shellgpu_memcpy_d2d(shellgpu_ctx *ctx, void *dst, const void *src, size_t bytes);
int shellgpu_memset(shellgpu_ctx *ctx, void *dst, int value, size_t bytes);

/* ---- model set-up (host arrays, copied) -------------------------------- */
/* TACSAssembler::setElementConnectivity + setNodes (src/TACSAssembler.h:
 * 120-140): conn[4*nelems] node ids in tensor order (n0,n1,n2,n3) =
 * (xi,eta) = (-1,-1),(1,-1),(-1,1),(1,1); Xpts[3*nnodes].  Builds the
 * node-to-node BSR pattern (TACSAssembler.cpp:1839-1992) and the element
 * colouring. */
int shellgpu_set_mesh(shellgpu_ctx *ctx, int nnodes, int nelems, const int *conn,
                      const double *Xpts);
/* TACSAssembler::setNodes / getNodes(TACSBVec *) (src/TACSAssembler.h:
 * 156-157, src/TACSAssembler.cpp:912-929): replace or read the nodal
 * coordinates of the mesh, device arrays of 3*nnodes (x, y, z per node).
 * The connectivity, pattern, colouring and factorisation plans are kept;
 * matrices and factors are assembled again by the caller. */
int shellgpu_set_nodes(shellgpu_ctx *ctx, const double *Xpts);
int shellgpu_get_nodes(shellgpu_ctx *ctx, double *Xpts);
/* TACSMaterialProperties(rho, 0, E, nu, ys, cte, 0) + TACSIsoShellConstitutive
 * (tOffset = 0) + TACSShellConstitutive::setDrillingRegularization */
int shellgpu_set_material(shellgpu_ctx *ctx, double E, double nu, double rho,
                          double cte, double k_drill);
/* TACSAssembler::addBCs (src/TACSAssembler.cpp, TACSBcMap): mask bits 0-5
 * of the 6 dofs per node; vals[6*nbc] (may be NULL = 0) */
int shellgpu_set_bcs(shellgpu_ctx *ctx, int nbc, const int *nodes, const int *mask,
                     const double *vals);
/* TACSAssembler::setDesignVars with one thickness per element
 * (TACSIsoShellConstitutive tNum = element) for nbatch designs;
 * thick[nbatch*nelems]; temp[nbatch*nelems] or NULL (TACSShellElement::
 * setTemperature, TACSShellElement.h:76). */
int shellgpu_set_designs(shellgpu_ctx *ctx, int nbatch, const double *thick,
                         const double *temp);
int shellgpu_get_sizes(shellgpu_ctx *ctx, int *nnodes, int *nelems, int *nnzb,
                       int *ncolors, int *nbatch);
/* the BSR pattern (rowp[nnodes+1], cols[nnzb]) and per-element colour */
int shellgpu_get_pattern(shellgpu_ctx *ctx, int *rowp, int *cols);
int shellgpu_get_colors(shellgpu_ctx *ctx, int *elem_color);

/* ---- matrices (TACSMat, src/bpmat/KSM.h:146-200) ---------------------- */
/* nbatch matrices sharing the context's pattern; values on the device,
 * nbatch*nnzb*36 doubles, 6x6 blocks row-major (BCSRMat, BCSRMatImpl.h) */
int shellgpu_mat_create(shellgpu_ctx *ctx, int nbatch, shellgpu_mat **A);
void shellgpu_mat_destroy(shellgpu_mat *A);
/* device pointer to the values; the caller may write through it, so the
 * matrix no longer counts as holding equal copies for all designs (below) */
double *shellgpu_mat_values(shellgpu_mat *A);
int shellgpu_mat_get_values(shellgpu_mat *A, double *host_vals);
int shellgpu_mat_set_values(shellgpu_mat *A, const double *host_vals);
/* TACSMat::zeroEntries / copyValues / axpy (Y += alpha X) */
int shellgpu_mat_zero(shellgpu_mat *A);
int shellgpu_mat_copy(shellgpu_mat *dst, const shellgpu_mat *src);
int shellgpu_mat_axpy(shellgpu_mat *Y, double alpha, const shellgpu_mat *X);
/* TACSSchurMat::applyBCs -> BCSRMat::zeroRow (BCSRMat.cpp:2005-2034): masked
 * rows zeroed, diagonal of masked dofs set to diag (1.0 in the reference) */
int shellgpu_mat_apply_bcs(shellgpu_mat *A, double diag);
/* TACSMat::mult -> BCSRMatVecMult6 (BCSRMatMult6.cpp:82-125): y = A x for
 * every matrix of the batch; x, y device, nbatch*6*nnodes.  A matrix whose
 * copies are known equal (zeroed, or assembled from identical designs with
 * no per-design state: K, M, or any type at u = null; changed since only by
 * the uniform operations scale / axpy of such a matrix / add_diag /
 * apply_bcs / copy) is read once for all designs' vectors, with the same
 * result bit for bit */
int shellgpu_mat_mult(shellgpu_mat *A, const double *x, double *y);
/* TACSMat::multTranspose (src/bpmat/KSM.h:176): y = A^T x for every
 * matrix of the batch (the node pattern is structurally symmetric) */
int shellgpu_mat_mult_transpose(shellgpu_mat *A, const double *x, double *y);
/* TACSMat::scale (src/bpmat/KSM.h:178): A = alpha A */
int shellgpu_mat_scale(shellgpu_mat *A, double alpha);
/* TACSMat::addDiag (src/bpmat/KSM.h:180, BCSRMat::addDiag): every diagonal
 * entry += alpha */
int shellgpu_mat_add_diag(shellgpu_mat *A, double alpha);
/* TACSMat::addValues (src/bpmat/KSM.h:153-154) -> BCSRMat::addRowValues
 * (src/bpmat/BCSRMat.cpp:1778-1855) for design `design` of the batch: block
 * k (host vals[36 k .. 36 k + 36), row-major 6x6) is added to block
 * (brow[k], bcol[k]).  Negative indices are skipped.  Blocks outside the
 * pattern are reported (non-zero return) after the others are added; the
 * additions to one block are summed in input order. */
int shellgpu_mat_add_blocks(shellgpu_mat *A, int design, int nblk, const int *brow, const int *bcol,
                            const double *vals);

/* ---- assembly (device state vectors, nbatch*6*nnodes, NULL = zero) ----- */
/* TACSAssembler::assembleMatType (src/TACSAssembler.cpp:4186-4250): zero,
 * element loop, applyBCs (unit diagonal) */
int shellgpu_assemble_mattype(shellgpu_ctx *ctx, int mattype, const double *u,
                              shellgpu_mat *A);
/* TACSAssembler::assembleMatCombo (src/TACSAssembler.cpp:4264-4310):
 * A = sum_k scale[k] (matrix of type mattypes[k] at u), then BCs (unit
 * diagonal) */
int shellgpu_assemble_mat_combo(shellgpu_ctx *ctx, int nmats, const int *mattypes,
                                const double *scale, const double *u, shellgpu_mat *A);
/* TACSAssembler::addJacobianVecProduct (src/TACSAssembler.cpp:4334-4390):
 * y += scale J x with J = alpha dR/du + gamma M at u (element matrices,
 * nothing assembled), J^T when transpose != 0, then y[bc] = 0.  beta,
 * dudt and d2udt2 do not enter the shell's matrices. */
int shellgpu_add_jacobian_vec_product(shellgpu_ctx *ctx, int model, double scale, double alpha,
                                      double beta, double gamma, const double *u,
                                      const double *dudt, const double *d2udt2, const double *x,
                                      double *y, int transpose);
/* TACSAssembler::assembleRes (src/TACSAssembler.cpp:4000-4070): R = element
 * residuals, then R[bc] = u[bc] - ubar[bc] (TACSBVec::applyBCs) */
int shellgpu_assemble_res(shellgpu_ctx *ctx, int model, const double *u, double *R);
/* TACSAssembler::assembleJacobian(alpha, beta, gamma, res, A)
 * (src/TACSAssembler.cpp:4084-4178) -> TACSShellElement::addJacobian
 * (src/elements/shell/TACSShellElement.h:455-672) at the state u (NULL =
 * 0) with time derivatives dudt and d2udt2 (the reference's vars, dvars,
 * ddvars; NULL = 0; nbatch*6*nnodes device vectors each):
 *   A = alpha dR/du + gamma M,   R = R(u) + M d2udt2,
 * then R[bc] = u[bc] - ubar[bc] and BC rows of A zeroed with a unit
 * diagonal.  The residual and the matrix come from one element pass.  The
 * linearized-rotation director has no velocity terms, so beta and dudt do
 * not enter (TACSDirector.h:368-432).  R may be NULL. */
int shellgpu_assemble_jacobian(shellgpu_ctx *ctx, int model, double alpha, double beta,
                               double gamma, const double *u, const double *dudt,
                               const double *d2udt2, double *R, shellgpu_mat *A);

/* TACSAssembler::getElementOutputData(TACS_BEAM_OR_SHELL_ELEMENT,
 * TACS_OUTPUT_STRAINS | TACS_OUTPUT_STRESSES | TACS_OUTPUT_EXTRAS)
 * (src/TACSAssembler.cpp:4433-4468) -> TACSShellElement::getOutputData
 * (src/elements/shell/TACSShellElement.h:1402-1500): out (device, nbatch x
 * nelems x 4 x 31) receives, at the 4 nodes of every element (in element
 * order), the strains (9), the stresses (9), failure0..5 and dv1..7 (31
 * values, the element zone of a TACSToFH5 file) of state u (NULL = 0) for the
 * given model; ys is the yield stress of the von Mises failure index. */
int shellgpu_element_output(shellgpu_ctx *ctx, int model, const double *u, double ys, double *out);

/* ---- design sensitivities (thickness t_e per element; device outputs
 * accumulate, nbatch*nelems) ---------------------------------------------- */
/* TACSAssembler::addAdjointResProducts -> TACSShellElement::addAdjResProduct
 * (src/elements/shell/TACSShellElement.h:774-868): dfdt[d*nelems + e] +=
 * scale psi^T dR_e/dt_e at state u (NULL = zero) for the given model */
int shellgpu_add_adj_res_product(shellgpu_ctx *ctx, int model, double scale, const double *psi,
                                 const double *u, double *dfdt);
/* TACSAssembler::addMatDVSensInnerProduct(scale, matType, psi, phi, dfdx)
 * (element default src/elements/TACSElement.cpp:355-432, central
 * differences there, exact here): dfdt += scale psi^T (dA/dt_e) phi with
 * A = K (SHELLGPU_STIFFNESS, u ignored), G(u) (SHELLGPU_GEOMETRIC) or M
 * (SHELLGPU_MASS, u ignored).  psi NULL means psi = phi; psi != phi uses
 * the symmetry of A: psi^T A' phi = ((psi+phi)^T A' (psi+phi) -
 * (psi-phi)^T A' (psi-phi)) / 4. */
int shellgpu_add_mat_dv_sens_inner_product(shellgpu_ctx *ctx, int mattype, double scale,
                                           const double *psi, const double *phi, const double *u,
                                           double *dfdt);
/* TACSAssembler::evalMatSVSensInnerProduct(GEOMETRIC, psi, phi, dfdu)
 * (element default src/elements/TACSElement.cpp:509-560): dfdu =
 * d(psi^T G(u) phi)/du, which is independent of u (G is linear in the
 * path); psi NULL means psi = phi; no BCs applied */
int shellgpu_eval_mat_sv_sens_inner_product(shellgpu_ctx *ctx, const double *psi, const double *phi,
                                            double *dfdu);

/* ---- coordinate sensitivities (device outputs accumulate,
 * nbatch*3*nnodes, x/y/z per node) --------------------------------------- */
/* TACSAssembler::addMatXptSensInnerProduct(scale, matType, psi, phi, dfdX)
 * (element default src/elements/TACSElement.cpp:434-507, central
 * differences there, exact forward-mode derivatives here): dfdX += scale
 * d(psi^T A phi)/dX with A = K (SHELLGPU_STIFFNESS, u ignored), G(u)
 * (SHELLGPU_GEOMETRIC) or M (SHELLGPU_MASS, u ignored); psi NULL means
 * psi = phi (psi != phi by polarization, as above) */
int shellgpu_add_mat_xpt_sens_inner_product(shellgpu_ctx *ctx, int mattype, double scale,
                                            const double *psi, const double *phi, const double *u,
                                            double *dfdX);
/* TACSAssembler::addAdjointResXptSensProducts -> addAdjResXptProduct
 * (element default src/elements/TACSElement.cpp:270-332): dfdX += scale
 * d(psi^T R(u))/dX at state u (NULL = zero) for the given model */
int shellgpu_add_adj_res_xpt_product(shellgpu_ctx *ctx, int model, double scale, const double *psi,
                                     const double *u, double *dfdX);

/* ---- vector algebra on length-n device vectors (TACSBVec::axpby, scale,
 * dot, src/bpmat/TACSBVec.cpp:237-281 and the BLAS calls of TACSBVec);
 * results of shellgpu_vec_dot are returned on the host ---------------- */
int shellgpu_vec_axpby(shellgpu_ctx *ctx, size_t n, double alpha, double beta, const double *x,
                       double *y); /* y = alpha x + beta y */
int shellgpu_vec_scale(shellgpu_ctx *ctx, size_t n, double alpha, double *x);
int shellgpu_vec_dot(shellgpu_ctx *ctx, size_t n, const double *x, const double *y, double *out);

/* ---- vector BCs (TACSBVec, src/bpmat/TACSBVec.cpp:546-625) ------------- */
/* x[bc] = 0 for every design of the batch */
int shellgpu_vec_zero_bcs(shellgpu_ctx *ctx, double *x);
/* x[bc] = ubar[bc] (TACSBVec::setBCs) */
int shellgpu_vec_set_bcs(shellgpu_ctx *ctx, double *x);

/* ---- direct solver (TACSPc, src/bpmat/KSM.h:209-230) ------------------ */
/* Replaces TACSSchurPc's AMD-ordered complete block LU (src/bpmat/
 * TACSSchurMat.cpp:944-1652): nodes are ordered by breadth-first level sets
 * from a seed set, which makes the matrix block tridiagonal; the level
 * blocks are factored densely on the GPU for every design of the batch.
 * Seed nodes (e.g. one boundary ring) set the level structure; without a
 * seed a pseudo-peripheral node is used. */
int shellgpu_set_ordering_seed(shellgpu_ctx *ctx, int n, const int *nodes);
int shellgpu_factor_create(shellgpu_ctx *ctx, shellgpu_factor **F);
void shellgpu_factor_destroy(shellgpu_factor *F);
/* number of levels, largest level (dofs), factor memory (GiB, all designs) */
int shellgpu_factor_info(shellgpu_factor *F, int *nlevels, int *max_level_dofs, double *gib);
/* TACSPc::factor (TACSSchurPc::factor, TACSSchurMat.cpp:1430-1510) */
int shellgpu_factor_compute(shellgpu_factor *F, shellgpu_mat *A);
/* TACSPc::applyFactor: x = A^{-1} b for every design (device vectors) */
int shellgpu_factor_apply(shellgpu_factor *F, const double *b, double *x);

/* ---- Krylov (TACSKsm / GMRES, src/bpmat/KSM.cpp:547-955) --------------- */
/* Right-preconditioned restarted GMRES(msub) with nrestart restarts; stops
 * when |r| < atol or |r| < rtol |r0| (GMRES::solve, KSM.cpp:785-955); x is
 * zeroed first (zero_guess = 1).  F may be NULL (no preconditioner).
 * Returns the iteration count and the worst relative residual over designs. */
int shellgpu_gmres(shellgpu_ctx *ctx, shellgpu_mat *A, shellgpu_factor *F, const double *b,
                   double *x, int msub, int nrestart, double rtol, double atol, int *iters,
                   double *resnorm);
/* shellgpu_gmres without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops that reuse one factor, harnesses that forbid hipMalloc).  The
 * Krylov basis, the work vector and the reduction partials live in the
 * library's fixed device scratch array of 64 MiB (8,388,608 doubles), one
 * per GPU, held until the process ends: (msub + 2) vectors of
 * nbatch * 6 * nnodes doubles (each rounded up to a multiple of 32) plus
 * 258 * nbatch + 2 doubles must fit in it, or the call fails with an error.
 * F (may be NULL) must be factored; a factor of one design shared by
 * several designs must have been applied once with shellgpu_factor_apply
 * (which sizes its vector buffers).  Results equal shellgpu_gmres's bit for
 * bit.  Every call on a GPU uses the same array: calls that overlap on one
 * GPU (other threads, other contexts) race. */
int shellgpu_gmres_noalloc(shellgpu_ctx *ctx, shellgpu_mat *A, shellgpu_factor *F, const double *b,
                           double *x, int msub, int nrestart, double rtol, double atol, int *iters,
                           double *resnorm);

/* ---- nonlinear static path ------------------------------------------ */
/* Load-stepped Newton on the nonlinear model (TACSQuad4NonlinearShell).
 * The built reference has no nonlinear driver (the unbuilt
 * archive/TACSContinuation.{h,cpp} is an arc-length method); this solves
 * R_nl(u) = s f for s = k/nsteps, k = 1..nsteps, each load step by full
 * Newton from the previous state: R and the tangent by
 * shellgpu_assemble_jacobian(NONLINEAR), K_t du = R by GMRES preconditioned
 * by a factored tangent (refactored when GMRES stalls), u -= du.  A load
 * step ends when ||R - s f|| <= rtol ||f|| for every design (BC rows hold
 * u - ubar; f is zeroed on BC dofs).  u (device, nbatch*6*nnodes) is the
 * initial guess and the result; iters = Newton iterations over all steps;
 * relres (host, nbatch, may be NULL) = final relative residuals. */
int shellgpu_newton_solve(shellgpu_ctx *ctx, const double *force, int nsteps, int max_iters,
                          double rtol, double *u, int *iters, double *relres);
/* shellgpu_newton_solve with a factor object of the caller's
 * (shellgpu_factor_create on this context): Newton factors its tangents into
 * F and leaves the last one there.  From a zero initial state at zero
 * temperature, with no refactorisation, that is the tangent at zero state,
 * i.e. the stiffness K, which shellgpu_buckling_set_stiffness_factor can
 * share.  (No reference counterpart, as for shellgpu_newton_solve.) */
int shellgpu_newton_solve_factor(shellgpu_ctx *ctx, shellgpu_factor *F, const double *force,
                                 int nsteps, int max_iters, double rtol, double *u, int *iters,
                                 double *relres);

/* ---- linear buckling (TACSLinearBuckling, src/TACSBuckling.h:43-107) ---- */
/* TACSLinearBuckling::solve(rhs = force, u0 = NULL) (src/TACSBuckling.cpp:
 * 197-278) for every design: K, static path K u = f, G(u), K + sigma G,
 * shift-invert Lanczos (SEP::solve, src/bpmat/GSEP.cpp:485-599) with
 * max_lanczos steps at most, stopping when the nev eigenvalues of smallest
 * magnitude have relative Ritz residual <= tol.  sigma > 0 is the shift as
 * in the reference (halved while K + sigma G is indefinite); sigma <= 0
 * selects 0.98 x a lower bound of the first positive load per design, or,
 * for identical designs (load cases of one structure, nbatch > 1), first
 * tries buckling-mode Lanczos on their shared factor of K and reports
 * sigma_used 0 when it converges.  Outputs:
 * lambda[nbatch*nev] (host), phi (device, nbatch*nev*6*nnodes,
 * K-normalised; may be NULL), path (device, nbatch*6*nnodes; may be NULL),
 * Lanczos steps, relative Ritz residuals resid[nbatch*nev] (host, may be
 * NULL), shift used per design sigma_used[nbatch] (host, may be NULL).
 * force is device, nbatch*6*nnodes. */
int shellgpu_linear_buckling(shellgpu_ctx *ctx, double sigma, const double *force, int nev,
                             int max_lanczos, double tol, double *lambda, double *phi,
                             double *path, int *nsteps, double *resid, double *sigma_used);

/* The analysis object behind shellgpu_linear_buckling, for callers that
 * extract modes one at a time and need eigenvalue sensitivities.
 * TACSLinearBuckling constructor (src/TACSBuckling.cpp:51-137): model is
 * the element model whose residual defines a path passed as u0 (the
 * adjoint of a nonlinear path uses its tangent); sigma, nev, max_lanczos
 * and tol as in shellgpu_linear_buckling. */
int shellgpu_buckling_create(shellgpu_ctx *ctx, int model, double sigma, int nev,
                             int max_lanczos, double tol, shellgpu_buckling **B);
void shellgpu_buckling_destroy(shellgpu_buckling *B);
/* The matrices kmat, gmat, aux_mat and the preconditioner bound to aux_mat
 * that the TACSLinearBuckling constructor receives (src/TACSBuckling.cpp:
 * 51-137; K, G, A distinct, as the reference requires at :75-101): the
 * analysis assembles K and G into them and factors A = K + sigma G with F.
 * They stay owned by the caller.  Without this call the object uses its
 * own. */
int shellgpu_buckling_set_operators(shellgpu_buckling *B, shellgpu_mat *K, shellgpu_mat *G,
                                    shellgpu_mat *A, shellgpu_factor *F);
/* A factor of the caller's holding the stiffness K at zero state (BCs
 * applied) for this context's designs, such as the one
 * shellgpu_newton_solve_factor leaves.  The automatic shift estimate and
 * the GMRES of the static path and of the path adjoint use it as K^-1 (or
 * preconditioner) instead of the object factoring K.  The reference's
 * TACSLinearBuckling::solve(rhs, u0) factors only K + sigma G when u0 is
 * given and takes sigma from the user (src/TACSBuckling.cpp:238-275); the
 * automatic shift here needs K^-1, and this call shares the caller's factor
 * as TACS drivers share one preconditioner between analyses.  F stays the
 * caller's and must stay current (same designs).  A factor of a nearby
 * matrix only loosens the shift estimate: a shift above the first load is
 * caught by the Cholesky of K + sigma G and lowered. */
int shellgpu_buckling_set_stiffness_factor(shellgpu_buckling *B, shellgpu_factor *F);
/* GMRES settings of the static path solve (GMRES constructor and
 * setTolerances, src/bpmat/KSM.cpp:547-560, 727-730; default 10, 15, 1e-12,
 * 1e-30) */
int shellgpu_buckling_set_ksm(shellgpu_buckling *B, int msub, int nrestart, double rtol,
                              double atol);
/* TACSLinearBuckling::solve(rhs, u0) (src/TACSBuckling.cpp:197-278): K at
 * zero state; path = static solution of K u = f (u0 == NULL) or u0 (device,
 * nbatch*6*nnodes; BC values are imposed on it); G(path); K + sigma G;
 * Lanczos.  force (device) may be NULL. */
int shellgpu_buckling_solve(shellgpu_buckling *B, const double *force, const double *u0);
/* TACSLinearBuckling::extractEigenvalue / extractEigenvector
 * (src/TACSBuckling.cpp:283-297): eigenvalue n (0-based, sorted by
 * magnitude) of every design: lambda[nbatch] and error[nbatch] on the host
 * (either may be NULL), K-normalised phi on the device (nbatch*6*nnodes,
 * may be NULL). */
int shellgpu_buckling_extract(shellgpu_buckling *B, int n, double *lambda, double *phi,
                              double *error);
/* Lanczos steps taken, shift per design (host, nbatch; 0 when the
 * buckling-mode run on a shared factor of K converged), path (device,
 * nbatch*6*nnodes); any may be NULL. */
int shellgpu_buckling_info(shellgpu_buckling *B, int *nsteps, double *sigma_used,
                           double *path);
/* TACSLinearBuckling::addEigenDVSens / evalEigenDVSens and the Xpt
 * counterparts (archive/TACSBuckling.cpp:382-592; the built reference has
 * them commented out, src/TACSBuckling.cpp:383-594):
 *   dfdt[d*nelems + e] += sum_n coef[d*nev + n] d(lambda_n)/d(t_e)
 * with
 *   d lambda = -(phi^T dK phi + lambda phi^T dG phi - lambda psi^T dR)
 *              / (phi^T G phi),   (dR/du) psi = d(phi^T G phi)/d(path),
 * all products exact (the reference differentiates matrix inner products
 * by central differences, src/elements/TACSElement.cpp:355-507).  One
 * adjoint solve serves all n.  coef is host, nbatch*nev; dfdt is device,
 * nbatch*nelems; dfdX (device, nbatch*3*nnodes) receives the coordinate
 * sensitivities the same way; either output may be NULL.  The shifted
 * operator is refactored by a later solve. */
int shellgpu_buckling_add_eigen_sens(shellgpu_buckling *B, const double *coef, double *dfdt,
                                     double *dfdX);

/* ---- natural frequencies ----------------------------------------------- */
/* TACSFrequencyAnalysis with the Lanczos solver (src/TACSBuckling.h:128-200,
 * constructor src/TACSBuckling.cpp:614-702, solve :810-893, extract
 * :895-918): K x = lambda M x with lambda = omega^2, K and M assembled at
 * zero state with BCs (unit diagonal), A = K - sigma M factored, shift-invert
 * Lanczos on A^{-1} M in the M inner product (EPGeneralizedShiftInvert,
 * src/bpmat/GSEP.cpp:106-190): lambda = 1/theta + sigma, sorted by |lambda|
 * (SMALLEST_MAGNITUDE).  One object serves every design of the context.
 * The direct factorisation needs A positive definite: a shift above the
 * first eigenvalue of a design is halved until A factors (the reference's LU
 * accepts it).  Stops when the nev wanted Ritz pairs have relative residual
 * <= tol. */
int shellgpu_frequency_create(shellgpu_ctx *ctx, double sigma, int nev, int max_lanczos, double tol,
                              shellgpu_frequency **F);
void shellgpu_frequency_destroy(shellgpu_frequency *F);
int shellgpu_frequency_solve(shellgpu_frequency *F);
/* TACSFrequencyAnalysis::extractEigenvalue / extractEigenvector
 * (src/TACSBuckling.cpp:895-918): lambda[d] (host, nbatch), phi (device,
 * nbatch x 6 nnodes, M-normalised; may be NULL), error[d] = relative Ritz
 * residual (may be NULL) */
int shellgpu_frequency_extract(shellgpu_frequency *F, int n, double *lambda, double *phi, double *error);
int shellgpu_frequency_info(shellgpu_frequency *F, int *nsteps, double *sigma_used);
/* d(sum_n coef[n] lambda_n)/dt added into dfdt (device, nbatch*nelems) with
 * d lambda / dt = phi^T (dK/dt - lambda dM/dt) phi / (phi^T M phi)
 * (TACSFrequencyAnalysis::evalEigenDVSens, src/TACSBuckling.cpp:950-995, where
 * the element products are central differences; exact here); dfdX (device,
 * nbatch*3*nnodes) receives the coordinate sensitivities by the same formula
 * (TACSFrequencyAnalysis::evalEigenXptSens, src/TACSBuckling.cpp:997-1022,
 * commented out in the reference).  coef is host, nbatch*nev; either output
 * may be NULL. */
int shellgpu_frequency_add_eigen_sens(shellgpu_frequency *F, const double *coef, double *dfdt,
                                      double *dfdX);

/* KS aggregate of f[0..n) (ksAggregation / ksAggregationSens,
 * src/utils/TacsUtilities.cpp:971-1011): ks = m + log(sum exp(rho (f_i -
 * m))) / rho, m = max f; dks[i] = d ks / d f_i (may be NULL).  Host only. */
int shellgpu_ks_aggregate(int n, double rho, const double *f, double *ks, double *dks);

/* ---- schedules (no reference counterpart: the reference's element loop
 * has one schedule, TACSAssembler_thread.cpp:27-252) ---------------------- */
/* Autotuner: times, on the context's mesh and designs, every compiled
 * schedule of the matrix-assembly element kernels (stiffness, tangent,
 * geometric stiffness, Jacobian of either model: BSR block reads before
 * the quadrature loop or at the scatter, one- or two-wavefront register
 * budget, two elements per workgroup or wide workgroups of 4 with the
 * geometry spread over them) and, with flags
 * bit 0, the two element colourings (greedy first fit in element order, or
 * the elements of most neighbours first), and keeps the fastest of each for
 * the context.  Results are the same arithmetic under every schedule. */
int shellgpu_tune(shellgpu_ctx *ctx, int flags);
/* the tuner's choices and timings as a JSON object ("{}" before tuning) */
int shellgpu_get_schedule(shellgpu_ctx *ctx, char *buf, int len);
/* force a schedule (0-7: bit 0 block reads at the scatter, bit 1 the
 * two-wavefront budget, bit 2 wide workgroups) of one kernel class
 * (0 stiffness, 1 tangent, 2 geometric, 3 Jacobian linear, 4 Jacobian
 * nonlinear) */
int shellgpu_set_elem_schedule(shellgpu_ctx *ctx, int kernel_class, int variant);

/* ---- device identification (no reference counterpart: measurement
 * plumbing) ---------------------------------------------------------------- */
/* the context's GPU: marketing name (the runtime's, else the kernel
 * driver's product name of the PCI device) and architecture (e.g.
 * "gfx950"), PCI bus id ("0000:05:00.0"), compute units and HBM bytes; any
 * output may be NULL */
int shellgpu_device_info(shellgpu_ctx *ctx, char *name, int name_len, char *arch, int arch_len,
                         char *pci_bus_id, int pci_len, int *compute_units, size_t *hbm_bytes);

/* ---- timing (no reference counterpart: measurement plumbing) ---------- */
/* With timing enabled, HIP events on the context's stream bracket every
 * element-kernel pass (its colour launches, which run back to back) and
 * every SpMV launch; shellgpu_get_timing waits for the stream and returns
 * the summed durations (ms) and the launch counts since timing was
 * enabled, then resets them. */
int shellgpu_set_timing(shellgpu_ctx *ctx, int enable);
int shellgpu_get_timing(shellgpu_ctx *ctx, double *elem_ms, int *elem_launches,
                        double *spmv_ms, int *spmv_launches);
/* wall time between two points on the context's stream (HIP events) */
int shellgpu_timer_start(shellgpu_ctx *ctx);
int shellgpu_timer_stop(shellgpu_ctx *ctx, double *ms);

#ifdef __cplusplus
}
#endif
#endif
