// TACSShellGPU.cpp -- the TACS-shaped host classes over the shellgpu C ABI
// (elements, vectors, matrices, preconditioner, GMRES, assembler, buckling).
// The BDF reader is in TACSMeshLoader.cpp.
#include "TACSShellGPU.h"

#include <cmath>
#include <cstring>

namespace shellgpu {

namespace {
void report(shellgpu_ctx *ctx, int rc, const char *what) {
  if (rc) fprintf(stderr, "shellgpu: %s failed (%d): %s\n", what, rc, shellgpu_last_error(ctx));
}
}  // namespace

// ---- constitutive and elements --------------------------------------------

TACSIsoShellConstitutive::TACSIsoShellConstitutive(TACSMaterialProperties *props, TacsScalar t, int tNum,
                                                   TacsScalar, TacsScalar)
    : props(props), t(t), tNum(tNum) {
  props->incref();
}
TACSIsoShellConstitutive::~TACSIsoShellConstitutive() { props->decref(); }

TACSElement::TACSElement(TACSShellConstitutive *con, int nonlinear) : con(con), nonlinear(nonlinear), temperature(0.0) {
  con->incref();
}
TACSElement::~TACSElement() { con->decref(); }

TACSQuad4Shell::TACSQuad4Shell(TACSShellTransform *transform, TACSShellConstitutive *con)
    : TACSElement(con, 0), transform(transform) {
  transform->incref();
}
TACSQuad4Shell::~TACSQuad4Shell() { transform->decref(); }

TACSQuad4NonlinearShell::TACSQuad4NonlinearShell(TACSShellTransform *transform, TACSShellConstitutive *con)
    : TACSElement(con, 1), transform(transform) {
  transform->incref();
}
TACSQuad4NonlinearShell::~TACSQuad4NonlinearShell() { transform->decref(); }

TACSElement *TacsCreateShellByName(const char *name, TACSShellTransform *transform, TACSShellConstitutive *con) {
  if (!name) return NULL;
  if (strcmp(name, "CQUAD4") == 0 || strcmp(name, "CQUADR") == 0) return new TACSQuad4Shell(transform, con);
  return NULL;
}

// ---- vectors ----------------------------------------------------------------

TACSBVec::TACSBVec(TACSAssembler *assembler, size_t size)
    : assembler(assembler), ctx(assembler->getContext()), size(size), data(NULL) {
  void *p = NULL;
  int rc = shellgpu_malloc(ctx, &p, sizeof(double) * size);
  report(ctx, rc, "TACSBVec allocation");
  data = (double *)p;
  zeroEntries();
}
TACSBVec::~TACSBVec() {
  if (data) shellgpu_free(ctx, data);
}
int TACSBVec::getValues(std::vector<TacsScalar> &vals) {
  vals.resize(size);
  int rc = shellgpu_memcpy_d2h(ctx, vals.data(), data, sizeof(double) * size);
  report(ctx, rc, "TACSBVec::getValues");
  return rc;
}
int TACSBVec::setValues(const std::vector<TacsScalar> &vals) {
  if (vals.size() != size) {
    fprintf(stderr, "TACSBVec::setValues: size mismatch\n");
    return 1;
  }
  int rc = shellgpu_memcpy_h2d(ctx, data, vals.data(), sizeof(double) * size);
  report(ctx, rc, "TACSBVec::setValues");
  return rc;
}
TacsScalar TACSBVec::dot(TACSBVec *x) {
  double out = 0.0;
  report(ctx, shellgpu_vec_dot(ctx, size, data, x->data, &out), "TACSBVec::dot");
  return out;
}
TacsScalar TACSBVec::norm() { return std::sqrt(dot(this)); }
void TACSBVec::axpy(TacsScalar alpha, TACSBVec *x) {
  report(ctx, shellgpu_vec_axpby(ctx, size, alpha, 1.0, x->data, data), "TACSBVec::axpy");
}
void TACSBVec::axpby(TacsScalar alpha, TacsScalar beta, TACSBVec *x) {
  report(ctx, shellgpu_vec_axpby(ctx, size, alpha, beta, x->data, data), "TACSBVec::axpby");
}
void TACSBVec::scale(TacsScalar alpha) { report(ctx, shellgpu_vec_scale(ctx, size, alpha, data), "TACSBVec::scale"); }
void TACSBVec::copyValues(TACSBVec *x) {
  report(ctx, shellgpu_memcpy_d2d(ctx, data, x->data, sizeof(double) * size), "TACSBVec::copyValues");
}
void TACSBVec::zeroEntries() { report(ctx, shellgpu_memset(ctx, data, 0, sizeof(double) * size), "TACSBVec::zero"); }
void TACSBVec::setRand(double lower, double upper) {
  // splitmix64 stream with a fixed seed: the same values on every call
  std::vector<TacsScalar> v(size);
  unsigned long long z = 0x2545F4914F6CDD1DULL;
  for (size_t i = 0; i < size; i++) {
    z += 0x9E3779B97F4A7C15ULL;
    unsigned long long w = z;
    w = (w ^ (w >> 30)) * 0xBF58476D1CE4E5B9ULL;
    w = (w ^ (w >> 27)) * 0x94D049BB133111EBULL;
    w ^= w >> 31;
    v[i] = lower + (upper - lower) * ((double)(w >> 11) / 9007199254740992.0);
  }
  setValues(v);
}
void TACSBVec::applyBCs() { report(ctx, shellgpu_vec_zero_bcs(ctx, data), "TACSBVec::applyBCs"); }
void TACSBVec::setBCs() { report(ctx, shellgpu_vec_set_bcs(ctx, data), "TACSBVec::setBCs"); }

// ---- matrices, preconditioner, GMRES ----------------------------------------

TACSMat::TACSMat(TACSAssembler *assembler) : assembler(assembler), mat(NULL) {
  report(assembler->getContext(), shellgpu_mat_create(assembler->getContext(), 1, &mat), "TACSMat allocation");
}
TACSMat::~TACSMat() {
  if (mat) shellgpu_mat_destroy(mat);
}
void TACSMat::flush() {
  if (pend_row.empty()) return;
  shellgpu_ctx *ctx = assembler->getContext();
  report(ctx, shellgpu_mat_add_blocks(mat, 0, (int)pend_row.size(), pend_row.data(), pend_col.data(), pend_vals.data()),
         "TACSMat::addValues");
  pend_row.clear();
  pend_col.clear();
  pend_vals.clear();
}
shellgpu_mat *TACSMat::getHandle() {
  flush();
  return mat;
}
void TACSMat::mult(TACSBVec *x, TACSBVec *y) {
  flush();
  report(assembler->getContext(), shellgpu_mat_mult(mat, x->getDeviceArray(), y->getDeviceArray()), "TACSMat::mult");
}
void TACSMat::multTranspose(TACSBVec *x, TACSBVec *y) {
  flush();
  report(assembler->getContext(), shellgpu_mat_mult_transpose(mat, x->getDeviceArray(), y->getDeviceArray()),
         "TACSMat::multTranspose");
}
void TACSMat::zeroEntries() {
  // additions staged before the zeroing would have been zeroed with the rest
  pend_row.clear();
  pend_col.clear();
  pend_vals.clear();
  report(assembler->getContext(), shellgpu_mat_zero(mat), "TACSMat::zeroEntries");
}
void TACSMat::copyValues(TACSMat *m) {
  pend_row.clear();
  pend_col.clear();
  pend_vals.clear();
  report(assembler->getContext(), shellgpu_mat_copy(mat, m->getHandle()), "TACSMat::copyValues");
}
void TACSMat::axpy(TacsScalar alpha, TACSMat *m) {
  flush();
  report(assembler->getContext(), shellgpu_mat_axpy(mat, alpha, m->getHandle()), "TACSMat::axpy");
}
void TACSMat::scale(TacsScalar alpha) {
  flush();
  report(assembler->getContext(), shellgpu_mat_scale(mat, alpha), "TACSMat::scale");
}
void TACSMat::addDiag(TacsScalar alpha) {
  flush();
  report(assembler->getContext(), shellgpu_mat_add_diag(mat, alpha), "TACSMat::addDiag");
}
void TACSMat::applyBCs() {
  flush();
  report(assembler->getContext(), shellgpu_mat_apply_bcs(mat, 1.0), "TACSMat::applyBCs");
}
void TACSMat::addValues(int nrow, const int *row, int ncol, const int *col, int, int mv, const TacsScalar *values) {
  for (int i = 0; i < nrow; i++) {
    if (row[i] < 0) continue;
    for (int j = 0; j < ncol; j++) {
      if (col[j] < 0) continue;
      pend_row.push_back(row[i]);
      pend_col.push_back(col[j]);
      // block (i, j): rows 6 i .. 6 i + 5, columns 6 j .. 6 j + 5 of values
      for (int ii = 0; ii < 6; ii++)
        for (int jj = 0; jj < 6; jj++) pend_vals.push_back(values[(size_t)(6 * i + ii) * mv + 6 * j + jj]);
    }
  }
}
void TACSMat::endAssembly() { flush(); }
void TACSMat::getSize(int *nr, int *nc) {
  const int n = 6 * assembler->getNumNodes();
  if (nr) *nr = n;
  if (nc) *nc = n;
}
TACSMat *TACSMat::createDuplicate() { return new TACSMat(assembler); }

TACSSchurPc::TACSSchurPc(TACSMat *mat, int, double, int) : mat(mat), fac(NULL) {
  mat->incref();
  shellgpu_ctx *ctx = mat->getAssembler()->getContext();
  report(ctx, shellgpu_factor_create(ctx, &fac), "TACSSchurPc allocation");
}
TACSSchurPc::~TACSSchurPc() {
  if (fac) shellgpu_factor_destroy(fac);
  mat->decref();
}
void TACSSchurPc::factor() {
  report(mat->getAssembler()->getContext(), shellgpu_factor_compute(fac, mat->getHandle()), "TACSSchurPc::factor");
}
void TACSSchurPc::applyFactor(TACSBVec *x, TACSBVec *y) {
  report(mat->getAssembler()->getContext(), shellgpu_factor_apply(fac, x->getDeviceArray(), y->getDeviceArray()),
         "TACSSchurPc::applyFactor");
}

GMRES::GMRES(TACSMat *mat, TACSPc *pc, int m, int nrestart, int)
    : mat(mat), pc(pc), m(m), nrestart(nrestart), rtol(1e-8), atol(1e-30), last_iters(0) {
  mat->incref();
  if (pc) pc->incref();
}
GMRES::GMRES(TACSMat *mat, int m, int nrestart) : GMRES(mat, NULL, m, nrestart, 0) {}
GMRES::~GMRES() {
  mat->decref();
  if (pc) pc->decref();
}
void GMRES::setTolerances(double r, double a) {
  rtol = r;
  atol = a;
}
void GMRES::getOperators(TACSMat **m_, TACSPc **pc_) {
  if (m_) *m_ = mat;
  if (pc_) *pc_ = pc;
}
int GMRES::solve(TACSBVec *b, TACSBVec *x, int zero_guess) {
  shellgpu_ctx *ctx = mat->getAssembler()->getContext();
  int its = 0;
  double rn = 0.0;
  int rc = 0;
  if (zero_guess) {
    rc = shellgpu_gmres(ctx, mat->getHandle(), pc ? pc->getHandle() : NULL, b->getDeviceArray(),
                        x->getDeviceArray(), m, nrestart, rtol, atol, &its, &rn);
  } else {
    // correction form: A dx = b - A x, x += dx
    TACSBVec r(mat->getAssembler(), b->getSize()), dx(mat->getAssembler(), b->getSize());
    mat->mult(x, &r);
    r.axpby(1.0, -1.0, b);
    rc = shellgpu_gmres(ctx, mat->getHandle(), pc ? pc->getHandle() : NULL, r.getDeviceArray(),
                        dx.getDeviceArray(), m, nrestart, rtol, atol, &its, &rn);
    if (!rc) x->axpy(1.0, &dx);
  }
  report(ctx, rc, "GMRES::solve");
  last_iters = its;
  return (!rc && rn <= rtol) ? 1 : 0;
}

// ---- assembler ----------------------------------------------------------------

TACSAssembler::TACSAssembler(int num_nodes, int num_elements, const int *conn, const TacsScalar *Xpts, int device)
    : ctx(NULL), num_nodes(num_nodes), num_elements(num_elements), model(0), elements(num_elements, NULL),
      thick(num_elements, 1.0), vars(NULL), dvars(NULL), ddvars(NULL), ddvars_zero(true), designs_dirty(true),
      conn_h(conn, conn + 4 * (size_t)num_elements),
      comp_h(num_elements, 0), X_h(Xpts, Xpts + 3 * (size_t)num_nodes), num_comp(1) {
  int rc = shellgpu_create(&ctx, device);
  if (rc) {
    fprintf(stderr, "TACSAssembler: no GPU context (%d)\n", rc);
    return;
  }
  report(ctx, shellgpu_set_mesh(ctx, num_nodes, num_elements, conn, Xpts), "TACSAssembler::setMesh");
  vars = createVec();
  vars->incref();
}
TACSAssembler::~TACSAssembler() {
  if (vars) vars->decref();
  if (dvars) dvars->decref();
  if (ddvars) ddvars->decref();
  for (TACSElement *e : elements)
    if (e) e->decref();
  if (ctx) shellgpu_destroy(ctx);
}

void TACSAssembler::setComponents(int ncomp, const int *elem_comp) {
  num_comp = ncomp;
  comp_h.assign(elem_comp, elem_comp + num_elements);
}

int TACSAssembler::getElementOutputData(int write_flag, int *len, int *nvals, std::vector<TacsScalar> &data) {
  if (syncDesigns()) return 1;
  const bool want[3] = {(write_flag & TACS_OUTPUT_STRAINS) != 0, (write_flag & TACS_OUTPUT_STRESSES) != 0,
                        (write_flag & TACS_OUTPUT_EXTRAS) != 0};
  const int group_off[3] = {0, 9, 18}, group_len[3] = {9, 9, 13};
  int nv = 0;
  for (int g = 0; g < 3; g++) nv += want[g] ? group_len[g] : 0;
  *len = 4 * num_elements;
  *nvals = nv;
  data.assign((size_t)(*len) * nv, 0.0);
  if (nv == 0) return 0;
  TacsScalar ys = 1e11;
  if (elements[0] && elements[0]->getConstitutive()->getMaterialProperties())
    ys = elements[0]->getConstitutive()->getMaterialProperties()->ys;
  const size_t n31 = (size_t)(*len) * 31;
  double *dout = NULL;
  if (check(shellgpu_malloc(ctx, (void **)&dout, sizeof(double) * n31), "getElementOutputData")) return 1;
  std::vector<double> all(n31);
  int rc = check(shellgpu_element_output(ctx, model, vars->getDeviceArray(), ys, dout), "getElementOutputData");
  if (!rc) rc = check(shellgpu_memcpy_d2h(ctx, all.data(), dout, sizeof(double) * n31), "getElementOutputData");
  shellgpu_free(ctx, dout);
  if (rc) return rc;
  for (size_t p = 0; p < (size_t)(*len); p++) {
    int k = 0;
    for (int g = 0; g < 3; g++)
      if (want[g])
        for (int i = 0; i < group_len[g]; i++) data[p * nv + k++] = all[p * 31 + group_off[g] + i];
  }
  return 0;
}

int TACSAssembler::check(int rc, const char *what) {
  report(ctx, rc, what);
  return rc;
}

int TACSAssembler::setElements(TACSElement **elems) {
  int nl = -1;
  TACSMaterialProperties *mat = NULL;
  for (int e = 0; e < num_elements; e++) {
    if (!elems[e]) {
      fprintf(stderr, "TACSAssembler::setElements: element %d is NULL\n", e);
      return 1;
    }
    if (nl < 0) nl = elems[e]->isNonlinear();
    if (nl != elems[e]->isNonlinear()) {
      fprintf(stderr, "TACSAssembler::setElements: linear and nonlinear shells cannot be mixed\n");
      return 1;
    }
    TACSMaterialProperties *m = elems[e]->getConstitutive()->getMaterialProperties();
    if (!mat) mat = m;
    if (m->E != mat->E || m->nu != mat->nu || m->rho != mat->rho || m->alpha != mat->alpha) {
      fprintf(stderr, "TACSAssembler::setElements: all elements must share one isotropic material\n");
      return 1;
    }
  }
  for (int e = 0; e < num_elements; e++) {
    elems[e]->incref();
    if (elements[e]) elements[e]->decref();
    elements[e] = elems[e];
    thick[e] = elems[e]->getConstitutive()->getThickness();
  }
  model = nl > 0 ? 1 : 0;
  if (mat && check(shellgpu_set_material(ctx, mat->E, mat->nu, mat->rho, mat->alpha, 10.0), "setMaterial")) return 1;
  designs_dirty = true;
  return syncDesigns();
}

int TACSAssembler::addBCs(int nbcs, const int *nodes, const int *masks, const TacsScalar *vals) {
  return check(shellgpu_set_bcs(ctx, nbcs, nodes, masks, vals), "TACSAssembler::addBCs");
}

void TACSAssembler::setOrderingSeed(int n, const int *nodes) {
  check(shellgpu_set_ordering_seed(ctx, n, nodes), "TACSAssembler::setOrderingSeed");
}

int TACSAssembler::syncDesigns() {
  // temperatures live on the element objects (TACSQuad4Shell::setTemperature)
  std::vector<TacsScalar> temp(num_elements, 0.0);
  bool any = false;
  for (int e = 0; e < num_elements; e++) {
    if (elements[e]) temp[e] = elements[e]->getTemperature();
    any = any || temp[e] != 0.0;
  }
  if (!designs_dirty && thick == up_thick && temp == up_temp) return 0;
  designs_dirty = false;
  up_thick = thick;
  up_temp = temp;
  return check(shellgpu_set_designs(ctx, 1, thick.data(), any ? temp.data() : NULL), "TACSAssembler::setDesigns");
}

TACSBVec *TACSAssembler::createVec() { return new TACSBVec(this, 6 * (size_t)num_nodes); }
TACSBVec *TACSAssembler::createNodeVec() { return new TACSBVec(this, 3 * (size_t)num_nodes); }

void TACSAssembler::setNodes(TACSBVec *X) {
  if (!X || X->getSize() < 3 * (size_t)num_nodes) {
    fprintf(stderr, "TACSAssembler::setNodes: a node vector of 3 x %d values is required\n", num_nodes);
    return;
  }
  check(shellgpu_set_nodes(ctx, X->getDeviceArray()), "TACSAssembler::setNodes");
  std::vector<TacsScalar> x;
  X->getValues(x);
  X_h.assign(x.begin(), x.begin() + 3 * (size_t)num_nodes);
}

void TACSAssembler::getNodes(TACSBVec *X) {
  if (!X || X->getSize() < 3 * (size_t)num_nodes) return;
  check(shellgpu_get_nodes(ctx, X->getDeviceArray()), "TACSAssembler::getNodes");
}
TACSBVec *TACSAssembler::createDesignVec() { return new TACSBVec(this, (size_t)num_elements); }
TACSSchurMat *TACSAssembler::createSchurMat() { return new TACSSchurMat(this); }

void TACSAssembler::getDesignVars(TACSBVec *x) { x->setValues(thick); }
void TACSAssembler::setDesignVars(TACSBVec *x) {
  x->getValues(thick);
  syncDesigns();
}
void TACSAssembler::zeroVariables() { vars->zeroEntries(); }
void TACSAssembler::zeroDotVariables() {
  if (dvars) dvars->zeroEntries();
}
void TACSAssembler::zeroDDotVariables() {
  if (ddvars) ddvars->zeroEntries();
  ddvars_zero = true;
}
void TACSAssembler::setVariables(TACSBVec *q, TACSBVec *qdot, TACSBVec *qddot) {
  if (q) vars->copyValues(q);
  if (qdot) {
    if (!dvars) {
      dvars = createVec();
      dvars->incref();
    }
    dvars->copyValues(qdot);
  }
  if (qddot) {
    if (!ddvars) {
      ddvars = createVec();
      ddvars->incref();
    }
    ddvars->copyValues(qddot);
    ddvars_zero = false;
  }
}
void TACSAssembler::getVariables(TACSBVec *q, TACSBVec *qdot, TACSBVec *qddot) {
  if (q) q->copyValues(vars);
  if (qdot) {
    if (dvars)
      qdot->copyValues(dvars);
    else
      qdot->zeroEntries();
  }
  if (qddot) {
    if (ddvars)
      qddot->copyValues(ddvars);
    else
      qddot->zeroEntries();
  }
}
void TACSAssembler::applyBCs(TACSBVec *vec) { vec->applyBCs(); }
void TACSAssembler::applyBCs(TACSMat *mat) { mat->applyBCs(); }
void TACSAssembler::setBCs(TACSBVec *vec) { vec->setBCs(); }

void TACSAssembler::assembleRes(TACSBVec *residual, const TacsScalar) {
  syncDesigns();
  if (check(shellgpu_assemble_res(ctx, model, vars->getDeviceArray(), residual->getDeviceArray()),
            "TACSAssembler::assembleRes"))
    return;
  if (!ddvars_zero && ddvars) {
    // the inertial forces M q'' (zero on BC dofs; the BC rows keep u - ubar)
    TACSBVec *t = createVec();
    t->incref();
    check(shellgpu_add_jacobian_vec_product(ctx, model, 1.0, 0.0, 0.0, 1.0, vars->getDeviceArray(), NULL, NULL,
                                            ddvars->getDeviceArray(), t->getDeviceArray(), 0),
          "TACSAssembler::assembleRes (inertia)");
    residual->axpy(1.0, t);
    t->decref();
  }
}

void TACSAssembler::assembleJacobian(TacsScalar alpha, TacsScalar beta, TacsScalar gamma, TACSBVec *residual,
                                     TACSMat *A, MatrixOrientation, const TacsScalar) {
  syncDesigns();
  check(shellgpu_assemble_jacobian(ctx, model, alpha, beta, gamma, vars->getDeviceArray(),
                                   dvars ? dvars->getDeviceArray() : NULL,
                                   (ddvars && !ddvars_zero) ? ddvars->getDeviceArray() : NULL,
                                   residual ? residual->getDeviceArray() : NULL, A->getHandle()),
        "TACSAssembler::assembleJacobian");
}

// the shellgpu matrix type of getMatType(matType) for the assembler's model
static int mat_kind(ElementMatrixType matType, int model) {
  if (matType == TACS_GEOMETRIC_STIFFNESS_MATRIX) return SHELLGPU_GEOMETRIC;
  if (matType == TACS_MASS_MATRIX) return SHELLGPU_MASS;
  // getMatType(STIFFNESS) is the element's own Jacobian at the variables
  return model ? SHELLGPU_TANGENT_NONLINEAR : SHELLGPU_STIFFNESS;
}

void TACSAssembler::assembleMatCombo(ElementMatrixType matTypes[], TacsScalar scale[], int nmats, TACSMat *A,
                                     MatrixOrientation, const TacsScalar) {
  syncDesigns();
  std::vector<int> kinds(nmats);
  for (int k = 0; k < nmats; k++) kinds[k] = mat_kind(matTypes[k], model);
  check(shellgpu_assemble_mat_combo(ctx, nmats, kinds.data(), scale, vars->getDeviceArray(), A->getHandle()),
        "TACSAssembler::assembleMatCombo");
}

void TACSAssembler::addJacobianVecProduct(TacsScalar scale, TacsScalar alpha, TacsScalar beta, TacsScalar gamma,
                                          TACSBVec *x, TACSBVec *y, MatrixOrientation matOr, const TacsScalar) {
  syncDesigns();
  check(shellgpu_add_jacobian_vec_product(ctx, model, scale, alpha, beta, gamma, vars->getDeviceArray(), NULL, NULL,
                                          x->getDeviceArray(), y->getDeviceArray(), matOr == TACS_MAT_TRANSPOSE),
        "TACSAssembler::addJacobianVecProduct");
}

void TACSAssembler::assembleMatType(ElementMatrixType matType, TACSMat *A, MatrixOrientation, const TacsScalar) {
  syncDesigns();
  check(shellgpu_assemble_mattype(ctx, mat_kind(matType, model), vars->getDeviceArray(), A->getHandle()),
        "TACSAssembler::assembleMatType");
}

void TACSAssembler::addAdjointResProducts(TacsScalar scale, int n, TACSBVec **adjoint, TACSBVec **dfdx) {
  syncDesigns();
  for (int i = 0; i < n; i++)
    check(shellgpu_add_adj_res_product(ctx, model, scale, adjoint[i]->getDeviceArray(), vars->getDeviceArray(),
                                       dfdx[i]->getDeviceArray()),
          "TACSAssembler::addAdjointResProducts");
}

void TACSAssembler::addAdjointResXptSensProducts(TacsScalar scale, int n, TACSBVec **adjoint, TACSBVec **dfdX) {
  syncDesigns();
  for (int i = 0; i < n; i++)
    check(shellgpu_add_adj_res_xpt_product(ctx, model, scale, adjoint[i]->getDeviceArray(), vars->getDeviceArray(),
                                           dfdX[i]->getDeviceArray()),
          "TACSAssembler::addAdjointResXptSensProducts");
}

void TACSAssembler::addMatDVSensInnerProduct(TacsScalar scale, ElementMatrixType matType, TACSBVec *psi,
                                             TACSBVec *phi, TACSBVec *dfdx) {
  syncDesigns();
  const int kind = matType == TACS_GEOMETRIC_STIFFNESS_MATRIX ? SHELLGPU_GEOMETRIC
                   : matType == TACS_MASS_MATRIX              ? SHELLGPU_MASS
                                                              : SHELLGPU_STIFFNESS;
  check(shellgpu_add_mat_dv_sens_inner_product(ctx, kind, scale, psi->getDeviceArray(), phi->getDeviceArray(),
                                               vars->getDeviceArray(), dfdx->getDeviceArray()),
        "TACSAssembler::addMatDVSensInnerProduct");
}

void TACSAssembler::addMatXptSensInnerProduct(TacsScalar scale, ElementMatrixType matType, TACSBVec *psi,
                                              TACSBVec *phi, TACSBVec *dfdX) {
  syncDesigns();
  const int kind = matType == TACS_GEOMETRIC_STIFFNESS_MATRIX ? SHELLGPU_GEOMETRIC
                   : matType == TACS_MASS_MATRIX              ? SHELLGPU_MASS
                                                              : SHELLGPU_STIFFNESS;
  check(shellgpu_add_mat_xpt_sens_inner_product(ctx, kind, scale, psi->getDeviceArray(), phi->getDeviceArray(),
                                                vars->getDeviceArray(), dfdX->getDeviceArray()),
        "TACSAssembler::addMatXptSensInnerProduct");
}

void TACSAssembler::evalMatSVSensInnerProduct(ElementMatrixType matType, TACSBVec *psi, TACSBVec *phi,
                                              TACSBVec *res) {
  if (matType != TACS_GEOMETRIC_STIFFNESS_MATRIX) {
    fprintf(stderr, "TACSAssembler::evalMatSVSensInnerProduct: only the geometric stiffness depends on the state\n");
    return;
  }
  syncDesigns();
  check(shellgpu_eval_mat_sv_sens_inner_product(ctx, psi->getDeviceArray(), phi->getDeviceArray(),
                                                res->getDeviceArray()),
        "TACSAssembler::evalMatSVSensInnerProduct");
}

// ---- linear buckling ----------------------------------------------------------

TACSLinearBuckling::TACSLinearBuckling(TACSAssembler *assembler, TacsScalar sigma, TACSMat *gmat, TACSMat *kmat,
                                       TACSMat *aux_mat, TACSKsm *solver, int max_lanczos_vecs, int num_eigvals,
                                       double eig_tol)
    : assembler(assembler), gmat(gmat), kmat(kmat), aux_mat(aux_mat), solver(solver), sigma(sigma),
      max_lanczos_vecs(max_lanczos_vecs), num_eigvals(num_eigvals), eig_tol(eig_tol), buck(NULL) {
  assembler->incref();
  gmat->incref();
  kmat->incref();
  aux_mat->incref();
  solver->incref();
  shellgpu_ctx *ctx = assembler->getContext();
  TACSMat *smat = NULL;
  TACSPc *pc = NULL;
  solver->getOperators(&smat, &pc);
  if (gmat == kmat || gmat == aux_mat || kmat == aux_mat) {
    fprintf(stderr, "TACSLinearBuckling: K, G and the auxiliary matrix must be distinct\n");
    return;
  }
  if (smat != aux_mat || !pc || pc->getMat() != aux_mat) {
    fprintf(stderr, "TACSLinearBuckling: the solver and its preconditioner must be bound to aux_mat\n");
    return;
  }
  int rc = shellgpu_buckling_create(ctx, assembler->getModel(), sigma, num_eigvals, max_lanczos_vecs, eig_tol, &buck);
  if (!rc)
    rc = shellgpu_buckling_set_operators(buck, kmat->getHandle(), gmat->getHandle(), aux_mat->getHandle(),
                                         pc->getHandle());
  GMRES *g = dynamic_cast<GMRES *>(solver);
  if (!rc && g)
    rc = shellgpu_buckling_set_ksm(buck, g->getSubspaceSize(), g->getNumRestarts(), g->getRelTol(), g->getAbsTol());
  report(ctx, rc, "TACSLinearBuckling");
}

TACSLinearBuckling::~TACSLinearBuckling() {
  if (buck) shellgpu_buckling_destroy(buck);
  solver->decref();
  aux_mat->decref();
  kmat->decref();
  gmat->decref();
  assembler->decref();
}

void TACSLinearBuckling::solve(TACSBVec *rhs, TACSBVec *u0, KSMPrint *ksm_print) {
  if (!buck) return;
  shellgpu_ctx *ctx = assembler->getContext();
  assembler->syncDesigns();
  int rc = shellgpu_buckling_solve(buck, rhs ? rhs->getDeviceArray() : NULL, u0 ? u0->getDeviceArray() : NULL);
  report(ctx, rc, "TACSLinearBuckling::solve");
  if (rc) return;
  // the reference leaves the assembler at the path (setVariables(path))
  shellgpu_buckling_info(buck, NULL, NULL, assembler->getVariables()->getDeviceArray());
  if (ksm_print) {
    int steps = 0;
    shellgpu_buckling_info(buck, &steps, NULL, NULL);
    char line[128];
    snprintf(line, sizeof(line), "Lanczos steps %d\n", steps);
    ksm_print->print(line);
  }
}

TacsScalar TACSLinearBuckling::extractEigenvalue(int n, TacsScalar *error) {
  double lam = 0.0, err = 0.0;
  if (buck) report(assembler->getContext(), shellgpu_buckling_extract(buck, n, &lam, NULL, &err), "extractEigenvalue");
  if (error) *error = err;
  return lam;
}

TacsScalar TACSLinearBuckling::extractEigenvector(int n, TACSBVec *ans, TacsScalar *error) {
  double lam = 0.0, err = 0.0;
  if (buck)
    report(assembler->getContext(), shellgpu_buckling_extract(buck, n, &lam, ans->getDeviceArray(), &err),
           "extractEigenvector");
  if (error) *error = err;
  return lam;
}

void TACSLinearBuckling::evalEigenDVSens(int n, TACSBVec *dfdx) {
  dfdx->zeroEntries();
  std::vector<double> coef(num_eigvals, 0.0);
  if (n < 0 || n >= num_eigvals || !buck) return;
  coef[n] = 1.0;
  report(assembler->getContext(),
         shellgpu_buckling_add_eigen_sens(buck, coef.data(), dfdx->getDeviceArray(), NULL), "evalEigenDVSens");
}

void TACSLinearBuckling::evalEigenXptSens(int n, TACSBVec *dfdX) {
  dfdX->zeroEntries();
  std::vector<double> coef(num_eigvals, 0.0);
  if (n < 0 || n >= num_eigvals || !buck) return;
  coef[n] = 1.0;
  report(assembler->getContext(),
         shellgpu_buckling_add_eigen_sens(buck, coef.data(), NULL, dfdX->getDeviceArray()), "evalEigenXptSens");
}

// partial derivatives only, as the reference's addEigenDVSens /
// addEigenXptSens: -coef (phi^T dK phi + lambda phi^T dG phi) / (phi^T G phi)
void TACSLinearBuckling::addEigenDVSens(TacsScalar coef, int n, TACSBVec *dfdx) {
  TACSBVec *phi = assembler->createVec(), *t = assembler->createVec();
  phi->incref();
  t->incref();
  TacsScalar err;
  TacsScalar eig = extractEigenvector(n, phi, &err);
  gmat->mult(phi, t);
  TacsScalar scale = t->dot(phi);
  assembler->addMatDVSensInnerProduct(-coef / scale, TACS_STIFFNESS_MATRIX, phi, phi, dfdx);
  assembler->addMatDVSensInnerProduct(-eig * coef / scale, TACS_GEOMETRIC_STIFFNESS_MATRIX, phi, phi, dfdx);
  t->decref();
  phi->decref();
}

void TACSLinearBuckling::addEigenXptSens(TacsScalar coef, int n, TACSBVec *dfdX) {
  TACSBVec *phi = assembler->createVec(), *t = assembler->createVec();
  phi->incref();
  t->incref();
  TacsScalar err;
  TacsScalar eig = extractEigenvector(n, phi, &err);
  gmat->mult(phi, t);
  TacsScalar scale = t->dot(phi);
  assembler->addMatXptSensInnerProduct(-coef / scale, TACS_STIFFNESS_MATRIX, phi, phi, dfdX);
  assembler->addMatXptSensInnerProduct(-eig * coef / scale, TACS_GEOMETRIC_STIFFNESS_MATRIX, phi, phi, dfdX);
  t->decref();
  phi->decref();
}

void TACSLinearBuckling::checkEigenvector(int n) {
  TACSBVec *phi = assembler->createVec(), *t1 = assembler->createVec(), *t2 = assembler->createVec();
  phi->incref();
  t1->incref();
  t2->incref();
  TacsScalar err;
  TacsScalar eig = extractEigenvector(n, phi, &err);
  kmat->mult(phi, t1);
  gmat->mult(phi, t2);
  printf("|K*e| = %15.5e  \n|G*e| = %15.5e \n", t1->norm(), t2->norm());
  t1->axpy(eig, t2);
  printf("||K*e + lambda*G*e|| = %15.5e \n", t1->norm());
  t2->decref();
  t1->decref();
  phi->decref();
}

// ---- natural frequencies ---------------------------------------------------

TACSFrequencyAnalysis::TACSFrequencyAnalysis(TACSAssembler *assembler, TacsScalar sigma, TACSMat *mmat,
                                             TACSMat *kmat, TACSKsm *solver, int max_lanczos, int num_eigvals,
                                             double eig_tol)
    : assembler(assembler), mmat(mmat), kmat(kmat), solver(solver), sigma(sigma), max_lanczos(max_lanczos),
      num_eigvals(num_eigvals), eig_tol(eig_tol), freq(NULL) {
  assembler->incref();
  mmat->incref();
  kmat->incref();
  solver->incref();
  TACSMat *smat = NULL;
  TACSPc *pc = NULL;
  solver->getOperators(&smat, &pc);
  if (smat != kmat) fprintf(stderr, "Error, solver must be associated with the stiffness matrix\n");
  recreate();
}

void TACSFrequencyAnalysis::recreate() {
  if (freq) shellgpu_frequency_destroy(freq);
  freq = NULL;
  shellgpu_ctx *ctx = assembler->getContext();
  report(ctx, shellgpu_frequency_create(ctx, sigma, num_eigvals, max_lanczos, eig_tol, &freq),
         "TACSFrequencyAnalysis");
}

TACSFrequencyAnalysis::~TACSFrequencyAnalysis() {
  if (freq) shellgpu_frequency_destroy(freq);
  solver->decref();
  kmat->decref();
  mmat->decref();
  assembler->decref();
}

void TACSFrequencyAnalysis::setSigma(TacsScalar s) {
  sigma = s;
  recreate();
}

void TACSFrequencyAnalysis::solve(KSMPrint *ksm_print, int print_level) {
  (void)print_level;
  if (!freq) return;
  shellgpu_ctx *ctx = assembler->getContext();
  assembler->zeroVariables();
  assembler->syncDesigns();
  int rc = shellgpu_frequency_solve(freq);
  report(ctx, rc, "TACSFrequencyAnalysis::solve");
  if (rc) return;
  // the reference leaves M in mmat and the shifted operator in kmat
  double sig = sigma;
  shellgpu_frequency_info(freq, NULL, &sig);
  assembler->assembleMatType(TACS_STIFFNESS_MATRIX, kmat);
  assembler->assembleMatType(TACS_MASS_MATRIX, mmat);
  kmat->axpy(-sig, mmat);
  kmat->applyBCs();
  if (ksm_print) {
    int steps = 0;
    shellgpu_frequency_info(freq, &steps, NULL);
    char line[128];
    snprintf(line, sizeof(line), "Lanczos steps %d\n", steps);
    ksm_print->print(line);
  }
}

TacsScalar TACSFrequencyAnalysis::extractEigenvalue(int n, TacsScalar *error) {
  double lam = 0.0, err = 0.0;
  if (freq) report(assembler->getContext(), shellgpu_frequency_extract(freq, n, &lam, NULL, &err), "extractEigenvalue");
  if (error) *error = err;
  return lam;
}

TacsScalar TACSFrequencyAnalysis::extractEigenvector(int n, TACSBVec *ans, TacsScalar *error) {
  double lam = 0.0, err = 0.0;
  if (freq)
    report(assembler->getContext(), shellgpu_frequency_extract(freq, n, &lam, ans->getDeviceArray(), &err),
           "extractEigenvector");
  if (error) *error = err;
  return lam;
}

// (u^T M u) d lambda / dx = u^T (dK/dx - lambda dM/dx) u
void TACSFrequencyAnalysis::evalEigenDVSens(int n, TACSBVec *dfdx) {
  dfdx->zeroEntries();
  if (n < 0 || n >= num_eigvals || !freq) return;
  std::vector<double> coef(num_eigvals, 0.0);
  coef[n] = 1.0;
  report(assembler->getContext(),
         shellgpu_frequency_add_eigen_sens(freq, coef.data(), dfdx->getDeviceArray(), NULL), "evalEigenDVSens");
}

void TACSFrequencyAnalysis::evalEigenXptSens(int n, TACSBVec *dfdX) {
  dfdX->zeroEntries();
  if (n < 0 || n >= num_eigvals || !freq) return;
  std::vector<double> coef(num_eigvals, 0.0);
  coef[n] = 1.0;
  report(assembler->getContext(),
         shellgpu_frequency_add_eigen_sens(freq, coef.data(), NULL, dfdX->getDeviceArray()), "evalEigenXptSens");
}

void TACSFrequencyAnalysis::checkEigenvector(int n) {
  assembler->zeroVariables();
  assembler->assembleMatType(TACS_STIFFNESS_MATRIX, kmat);
  assembler->assembleMatType(TACS_MASS_MATRIX, mmat);
  TACSBVec *phi = assembler->createVec(), *t1 = assembler->createVec(), *t2 = assembler->createVec();
  phi->incref();
  t1->incref();
  t2->incref();
  TacsScalar err;
  TacsScalar eig = extractEigenvector(n, phi, &err);
  kmat->mult(phi, t1);
  mmat->mult(phi, t2);
  printf("|K*e| = %15.5e  \n|M*e| = %15.5e \n", t1->norm(), t2->norm());
  t1->axpy(-eig, t2);
  printf("||K*e - lambda*M*e|| = %15.5e \n", t1->norm());
  t2->decref();
  t1->decref();
  phi->decref();
}

// ---- .f5 output ---------------------------------------------------------------

TACSFH5File::TACSFH5File() : fp(NULL) {}
TACSFH5File::~TACSFH5File() { close(); }

int TACSFH5File::createFile(const char *file_name, int num_components, char **component_names) {
  if (fp) {
    fprintf(stderr, "[0] FH5: Error, cannot create new file\n");
    return 1;
  }
  fp = fopen(file_name, "wb");
  if (!fp) return 1;
  // number of components, then (length with the terminating zero, name)
  fwrite(&num_components, sizeof(int), 1, fp);
  for (int k = 0; k < num_components; k++) {
    const char *name = component_names[k] ? component_names[k] : "";
    int slen = (int)strlen(name) + 1;
    fwrite(&slen, sizeof(int), 1, fp);
    fwrite(name, 1, slen, fp);
  }
  return 0;
}

int TACSFH5File::writeZoneData(char *zone_name, char *var_names, FH5DataType data_name, int dim1, int dim2,
                               void *data) {
  if (!fp) return 0;
  int pre[5] = {(int)data_name, dim1, dim2, (int)strlen(zone_name) + 1, (int)strlen(var_names) + 1};
  fwrite(pre, sizeof(int), 5, fp);
  fwrite(zone_name, 1, pre[3], fp);
  fwrite(var_names, 1, pre[4], fp);
  const size_t sz = data_name == FH5_INT ? sizeof(int) : (data_name == FH5_FLOAT ? sizeof(float) : sizeof(double));
  fwrite(data, sz, (size_t)dim1 * dim2, fp);
  return 1;
}

void TACSFH5File::close() {
  if (fp) fclose(fp);
  fp = NULL;
}

namespace {
const char *kShellDisp[6] = {"u", "v", "w", "rotx", "roty", "rotz"};
const char *kShellLoads[6] = {"fx", "fy", "fz", "mx", "my", "mz"};
const char *kShellStrains[9] = {"ex0", "ey0", "exy0", "ex1", "ey1", "exy1", "eyz0", "exz0", "erot"};
const char *kShellStresses[9] = {"sx0", "sy0", "sxy0", "sx1", "sy1", "sxy1", "syz0", "sxz0", "srot"};
const char *kShellExtras[13] = {"failure0", "failure1", "failure2", "failure3", "failure4", "failure5", "dv1",
                                "dv2",      "dv3",      "dv4",      "dv5",      "dv6",      "dv7"};

std::string join(const char *const *names, int n) {
  std::string s;
  for (int i = 0; i < n; i++) {
    if (i) s += ",";
    s += names[i];
  }
  return s;
}
}  // namespace

TACSToFH5::TACSToFH5(TACSAssembler *assembler, ElementType elem_type, int write_flag)
    : assembler(assembler), elem_type(elem_type), write_flag(write_flag) {
  assembler->incref();
  element_write_flag =
      write_flag & (~(TACS_OUTPUT_CONNECTIVITY | TACS_OUTPUT_NODES | TACS_OUTPUT_DISPLACEMENTS | TACS_OUTPUT_LOADS));
  // getElementVarNames (src/io/TACSToFH5.cpp:480-545)
  nvals = 0;
  std::string v;
  if (element_write_flag & TACS_OUTPUT_STRAINS) {
    v += join(kShellStrains, 9);
    nvals += 9;
  }
  if (element_write_flag & TACS_OUTPUT_STRESSES) {
    v += (v.empty() ? "" : ",") + join(kShellStresses, 9);
    nvals += 9;
  }
  if (element_write_flag & TACS_OUTPUT_EXTRAS) {
    v += (v.empty() ? "" : ",") + join(kShellExtras, 13);
    nvals += 13;
  }
  variable_names = v;
  if (elem_type != TACS_BEAM_OR_SHELL_ELEMENT) {
    fprintf(stderr, "TACSToFH5: only TACS_BEAM_OR_SHELL_ELEMENT output is implemented\n");
    nvals = 0;
  }
  const int nc = assembler->getNumComponents();
  for (int k = 0; k < nc; k++) {
    char comp_name[128];
    snprintf(comp_name, sizeof(comp_name), "Component %d", k);
    component_names.push_back(comp_name);
  }
}

TACSToFH5::~TACSToFH5() { assembler->decref(); }

void TACSToFH5::setComponentName(int comp_num, const char *comp_name) {
  if (comp_num >= 0 && comp_num < (int)component_names.size()) component_names[comp_num] = comp_name;
}

int TACSToFH5::writeConnectivity(TACSFH5File *file) {
  // TACSToFH5::writeConnectivity (src/io/TACSToFH5.cpp:369-476), one rank
  const int ne = assembler->getNumElements();
  std::vector<int> comp(ne), ltypes(ne, TACS_QUAD_ELEMENT), ptr(ne + 1);
  for (int e = 0; e < ne; e++) comp[e] = assembler->getComponentNum(e);
  for (int e = 0; e <= ne; e++) ptr[e] = 4 * e;
  std::vector<int> conn = assembler->getConnectivity();
  char comp_name[] = "components", layout_name[] = "ltypes", ptr_name[] = "ptr", conn_name[] = "connectivity";
  file->writeZoneData(comp_name, comp_name, TACSFH5File::FH5_INT, ne, 1, comp.data());
  file->writeZoneData(layout_name, layout_name, TACSFH5File::FH5_INT, ne, 1, ltypes.data());
  file->writeZoneData(ptr_name, ptr_name, TACSFH5File::FH5_INT, ne + 1, 1, ptr.data());
  file->writeZoneData(conn_name, conn_name, TACSFH5File::FH5_INT, 4 * ne, 1, conn.data());
  return 0;
}

int TACSToFH5::writeToFile(const char *filename) {
  // TACSToFH5::writeToFile (src/io/TACSToFH5.cpp:126-362), one rank
  TACSFH5File *file = new TACSFH5File();
  file->incref();
  std::vector<char *> names;
  for (std::string &n : component_names) names.push_back(&n[0]);
  if (file->createFile(filename, (int)names.size(), names.data())) {
    file->decref();
    fprintf(stderr, "[0] TACSToFH5 error: Could not create file\n");
    return 1;
  }
  if (write_flag & TACS_OUTPUT_CONNECTIVITY) writeConnectivity(file);

  if (write_flag & (TACS_OUTPUT_NODES | TACS_OUTPUT_DISPLACEMENTS | TACS_OUTPUT_LOADS)) {
    std::string var_names;
    int dim2 = 0;
    if (write_flag & TACS_OUTPUT_NODES) {
      var_names = "X,Y,Z";
      dim2 += 3;
    }
    if (write_flag & TACS_OUTPUT_DISPLACEMENTS) {
      var_names += (var_names.empty() ? "" : ",") + join(kShellDisp, 6);
      dim2 += 6;
    }
    if (write_flag & TACS_OUTPUT_LOADS) {
      var_names += (var_names.empty() ? "" : ",") + join(kShellLoads, 6);
      dim2 += 6;
    }
    const int nn = assembler->getNumNodes();
    std::vector<TacsScalar> u, F;
    if (write_flag & TACS_OUTPUT_DISPLACEMENTS) assembler->getVariables()->getValues(u);
    if (write_flag & TACS_OUTPUT_LOADS) {
      // the residual without external loads infers the applied load
      TACSBVec *R = assembler->createVec();
      R->incref();
      assembler->assembleRes(R);
      R->getValues(F);
      R->decref();
    }
    const std::vector<TacsScalar> &X = assembler->getNodeLocations();
    std::vector<float> fd((size_t)nn * dim2);
    for (int i = 0; i < nn; i++) {
      int k = 0;
      if (write_flag & TACS_OUTPUT_NODES)
        for (int j = 0; j < 3; j++) fd[(size_t)dim2 * i + k++] = (float)X[3 * i + j];
      if (write_flag & TACS_OUTPUT_DISPLACEMENTS)
        for (int j = 0; j < 6; j++) fd[(size_t)dim2 * i + k++] = (float)u[6 * i + j];
      if (write_flag & TACS_OUTPUT_LOADS)
        for (int j = 0; j < 6; j++) fd[(size_t)dim2 * i + k++] = (float)F[6 * i + j];
    }
    char data_name[128];
    snprintf(data_name, sizeof(data_name), "continuous data t=%.10e", 0.0);
    file->writeZoneData(data_name, &var_names[0], TACSFH5File::FH5_FLOAT, nn, dim2, fd.data());
  }

  if (nvals > 0) {
    int len = 0, nv = 0;
    std::vector<TacsScalar> data;
    if (assembler->getElementOutputData(element_write_flag, &len, &nv, data) == 0) {
      std::vector<float> fd(data.size());
      for (size_t i = 0; i < data.size(); i++) fd[i] = (float)data[i];
      char data_name[128];
      snprintf(data_name, sizeof(data_name), "element data t=%.10e", 0.0);
      file->writeZoneData(data_name, &variable_names[0], TACSFH5File::FH5_FLOAT, len, nv, fd.data());
    }
  }
  file->close();
  file->decref();
  return 0;
}

}  // namespace shellgpu
