// TACSShellGPU.h -- C++ host classes shaped like the reference's TACS API
// (TACSAssembler, TACSBVec, TACSMat/TACSSchurMat, TACSPc/TACSSchurPc,
// GMRES, TACSLinearBuckling, TACSMeshLoader and the shell element / material
// objects), implemented over the shellgpu C ABI.  Driver code written
// against the reference (examples/cylinder-buckling/mechBuckling.cpp) ports
// to these classes call for call; the differences are listed per class.
//
// Common conventions of the reference kept here: intrusive reference
// counts (TACSObject incref/decref, src/utils/TACSObject.h:112-135; new
// objects start at 0, holders incref their inputs); setup calls return an
// int (0 = success); run-time failures print to stderr.  All data live on
// one GPU; a TACSAssembler owns one shellgpu context (one HIP stream).
#pragma once

#include <cstdio>
#include <string>
#include <vector>

#include "shellgpu.h"

// MPI, when its header is on the include path: the reference's drivers call
// MPI_Init / MPI_Comm_rank / MPI_Finalize and pass MPI_COMM_WORLD to the
// mesh loader.  This library runs one rank; the communicator overloads
// below check that.
#if defined(__has_include)
#if __has_include(<mpi.h>)
#include <mpi.h>
#endif
#endif

namespace shellgpu {

typedef double TacsScalar;

// ElementMatrixType (src/elements/TACSElementTypes.h)
// Output flags, element types and layouts (src/elements/TACSElementTypes.h:
// 37-43, 49-60, 65-77), same values
const int TACS_OUTPUT_CONNECTIVITY = 1;
const int TACS_OUTPUT_NODES = 2;
const int TACS_OUTPUT_DISPLACEMENTS = 4;
const int TACS_OUTPUT_STRAINS = 8;
const int TACS_OUTPUT_STRESSES = 16;
const int TACS_OUTPUT_EXTRAS = 32;
const int TACS_OUTPUT_LOADS = 64;
enum ElementType { TACS_ELEMENT_NONE = 0, TACS_BEAM_OR_SHELL_ELEMENT = 4 };
enum ElementLayout { TACS_LAYOUT_NONE = 0, TACS_QUAD_ELEMENT = 8 };

// MatrixOrientation (src/bpmat/KSM.h:34), same values
enum MatrixOrientation { TACS_MAT_NORMAL, TACS_MAT_TRANSPOSE };

#ifdef MPI_VERSION
// the one-rank check of the communicator overloads
inline void tacs_check_one_rank(MPI_Comm comm, const char *who) {
  int size = 1;
  MPI_Comm_size(comm, &size);
  if (size != 1)
    fprintf(stderr, "%s: this library runs one rank per GPU context; the communicator has %d ranks\n", who,
            size);
}
#endif

// ElementMatrixType (src/elements/TACSElementTypes.h:101-107), same values
enum ElementMatrixType {
  TACS_JACOBIAN_MATRIX,
  TACS_STIFFNESS_MATRIX,
  TACS_MASS_MATRIX,
  TACS_GEOMETRIC_STIFFNESS_MATRIX,
  TACS_STIFFNESS_PRODUCT_DERIVATIVE
};

class TACSObject {
 public:
  TACSObject() : ref_count(0) {}
  virtual ~TACSObject() {}
  void incref() { ref_count++; }
  void decref() {
    if (--ref_count <= 0) delete this;
  }
  int refcount() const { return ref_count; }

 private:
  int ref_count;
};

// TACSMaterialProperties (src/constitutive/TACSMaterialProperties.h),
// isotropic constructor
class TACSMaterialProperties : public TACSObject {
 public:
  TACSMaterialProperties(TacsScalar rho, TacsScalar specific_heat, TacsScalar E, TacsScalar nu,
                         TacsScalar ys, TacsScalar alpha, TacsScalar kappa)
      : rho(rho), specific_heat(specific_heat), E(E), nu(nu), ys(ys), alpha(alpha), kappa(kappa) {}
  TacsScalar rho, specific_heat, E, nu, ys, alpha, kappa;
};

// Shell transforms (src/elements/shell/TACSShellElementTransform.h).  Only
// the natural transform is implemented on the GPU.
class TACSShellTransform : public TACSObject {};
class TACSShellNaturalTransform : public TACSShellTransform {};

// TACSIsoShellConstitutive(props, t, tNum, tlb, tub)
// (src/constitutive/TACSIsoShellConstitutive.cpp:28-81); k_drill = 10,
// tOffset = 0
class TACSShellConstitutive : public TACSObject {
 public:
  virtual ~TACSShellConstitutive() {}
  virtual TACSMaterialProperties *getMaterialProperties() = 0;
  virtual TacsScalar getThickness() = 0;
};
class TACSIsoShellConstitutive : public TACSShellConstitutive {
 public:
  TACSIsoShellConstitutive(TACSMaterialProperties *props, TacsScalar t = 1.0, int tNum = -1,
                           TacsScalar tlb = 0.0, TacsScalar tub = 1.0);
  ~TACSIsoShellConstitutive();
  TACSMaterialProperties *getMaterialProperties() { return props; }
  TacsScalar getThickness() { return t; }
  int getThicknessNum() { return tNum; }

 private:
  TACSMaterialProperties *props;
  TacsScalar t;
  int tNum;
};

// Shell elements (src/elements/shell/TACSShellElementDefs.h:12-29): the
// 4-node MITC director shell with the linearised rotation, linear or
// nonlinear strain model
class TACSElement : public TACSObject {
 public:
  TACSElement(TACSShellConstitutive *con, int nonlinear);
  ~TACSElement();
  void setTemperature(TacsScalar T) { temperature = T; }
  TacsScalar getTemperature() { return temperature; }
  TACSShellConstitutive *getConstitutive() { return con; }
  int isNonlinear() const { return nonlinear; }
  int getVarsPerNode() { return 6; }
  int getNumNodes() { return 4; }

 private:
  TACSShellConstitutive *con;
  int nonlinear;
  TacsScalar temperature;
};
class TACSQuad4Shell : public TACSElement {
 public:
  TACSQuad4Shell(TACSShellTransform *transform, TACSShellConstitutive *con);
  ~TACSQuad4Shell();

 private:
  TACSShellTransform *transform;
};
class TACSQuad4NonlinearShell : public TACSElement {
 public:
  TACSQuad4NonlinearShell(TACSShellTransform *transform, TACSShellConstitutive *con);
  ~TACSQuad4NonlinearShell();

 private:
  TACSShellTransform *transform;
};
// TacsCreateShellByName (src/elements/shell/TACSShellElementDefs.cpp):
// "CQUAD4" / "CQUADR" -> TACSQuad4Shell; NULL for element types this library
// does not implement
TACSElement *TacsCreateShellByName(const char *name, TACSShellTransform *transform,
                                   TACSShellConstitutive *con);

class TACSAssembler;

// TACSBVec (src/bpmat/TACSBVec.h): a device vector of vars_per_node = 6
// values per node (or 3 per node for coordinate vectors, 1 per element for
// design vectors)
class TACSBVec : public TACSObject {
 public:
  TACSBVec(TACSAssembler *assembler, size_t size);
  ~TACSBVec();
  size_t getSize() const { return size; }
  double *getDeviceArray() { return data; }
  // host copies of the values
  int getValues(std::vector<TacsScalar> &vals);
  int setValues(const std::vector<TacsScalar> &vals);
  TacsScalar norm();
  TacsScalar dot(TACSBVec *x);
  void axpy(TacsScalar alpha, TACSBVec *x);
  void axpby(TacsScalar alpha, TacsScalar beta, TACSBVec *x);
  void scale(TacsScalar alpha);
  void copyValues(TACSBVec *x);
  void zeroEntries();
  // uniform values in [lower, upper] from a fixed host generator
  // (TACSBVec::setRand, src/bpmat/TACSBVec.cpp:462-491)
  void setRand(double lower = -1.0, double upper = 1.0);
  // zero the BC entries / set them to the BC values (TACSBVec.cpp:546-625)
  void applyBCs();
  void setBCs();

 private:
  TACSAssembler *assembler;
  shellgpu_ctx *ctx;
  size_t size;
  double *data;
};

// TACSMat (src/bpmat/KSM.h:146-208) over the shared 6x6-block BSR pattern;
// TACSSchurMat is the name the examples use for the assembler's matrix
class TACSMat : public TACSObject {
 public:
  explicit TACSMat(TACSAssembler *assembler);
  ~TACSMat();
  void mult(TACSBVec *x, TACSBVec *y);
  // y = A^T x (TACSMat::multTranspose, KSM.h:176)
  void multTranspose(TACSBVec *x, TACSBVec *y);
  void zeroEntries();
  void copyValues(TACSMat *mat);
  void axpy(TacsScalar alpha, TACSMat *mat);
  // A = alpha A; diagonal += alpha (KSM.h:178-180)
  void scale(TacsScalar alpha);
  void addDiag(TacsScalar alpha);
  void applyBCs();
  // TACSMat::addValues (KSM.h:153-154, TACSSchurMat::addValues,
  // TACSSchurMat.cpp:453-526): values is the dense (6 nrow) x mv row-major
  // matrix of the blocks (row[i], col[j]) (block (i, j) at row 6 i, column
  // 6 j); negative indices are skipped, blocks outside the pattern are
  // reported on stderr.  Additions are staged on the host and applied in one
  // launch at endAssembly, or before any other use of the matrix.
  void addValues(int nrow, const int *row, int ncol, const int *col, int nv, int mv, const TacsScalar *values);
  void beginAssembly() {}
  void endAssembly();
  // scalar rows and columns
  void getSize(int *nr, int *nc);
  // a matrix with the same pattern (values zero)
  TACSMat *createDuplicate();
  // the device matrix (staged additions applied first)
  shellgpu_mat *getHandle();
  TACSAssembler *getAssembler() { return assembler; }

 private:
  void flush();
  TACSAssembler *assembler;
  shellgpu_mat *mat;
  std::vector<int> pend_row, pend_col;
  std::vector<TacsScalar> pend_vals;
};
class TACSSchurMat : public TACSMat {
 public:
  explicit TACSSchurMat(TACSAssembler *assembler) : TACSMat(assembler) {}
};

// TACSPc (src/bpmat/KSM.h:209-230).  TACSSchurPc(mat, levFill, fill,
// reorder) (src/bpmat/TACSSchurMat.h:59-214) factors the matrix it is
// bound to; here that factorisation is the level-set block Cholesky of the
// shellgpu library whatever levFill says (a complete factorisation, which
// is what levFill = 1e6 gives in the reference).
class TACSPc : public TACSObject {
 public:
  virtual ~TACSPc() {}
  virtual void factor() = 0;
  virtual void applyFactor(TACSBVec *x, TACSBVec *y) = 0;
  virtual TACSMat *getMat() = 0;
  virtual shellgpu_factor *getHandle() = 0;
};
class TACSSchurPc : public TACSPc {
 public:
  TACSSchurPc(TACSMat *mat, int levFill = 1000000, double fill = 10.0, int reorder_schur = 1);
  ~TACSSchurPc();
  void factor();
  void applyFactor(TACSBVec *x, TACSBVec *y);
  TACSMat *getMat() { return mat; }
  shellgpu_factor *getHandle() { return fac; }

 private:
  TACSMat *mat;
  shellgpu_factor *fac;
};

// KSMPrint (src/bpmat/KSM.h:232-262)
class KSMPrint : public TACSObject {
 public:
  virtual ~KSMPrint() {}
  virtual void print(const char *cstr) = 0;
};
class KSMPrintStdout : public KSMPrint {
 public:
  KSMPrintStdout(const char *descript, int rank, int freq) : descript(descript), rank(rank), freq(freq) {}
  void print(const char *cstr) {
    if (rank == 0) printf("%s: %s", descript.c_str(), cstr);
  }

 private:
  std::string descript;
  int rank, freq;
};

// TACSKsm / GMRES (src/bpmat/KSM.h:264-278, KSM.cpp:547-955):
// right-preconditioned restarted GMRES; solve returns 1 when converged
class TACSKsm : public TACSObject {
 public:
  virtual ~TACSKsm() {}
  virtual int solve(TACSBVec *b, TACSBVec *x, int zero_guess = 1) = 0;
  virtual void setTolerances(double rtol, double atol) = 0;
  virtual void getOperators(TACSMat **mat, TACSPc **pc) = 0;
  virtual void setMonitor(KSMPrint *print) {}
};
class GMRES : public TACSKsm {
 public:
  GMRES(TACSMat *mat, TACSPc *pc, int m, int nrestart, int isFlexible);
  GMRES(TACSMat *mat, int m, int nrestart);
  ~GMRES();
  int solve(TACSBVec *b, TACSBVec *x, int zero_guess = 1);
  void setTolerances(double rtol, double atol);
  void getOperators(TACSMat **mat, TACSPc **pc);
  int getSubspaceSize() const { return m; }
  int getNumRestarts() const { return nrestart; }
  double getRelTol() const { return rtol; }
  double getAbsTol() const { return atol; }
  int getIterCount() const { return last_iters; }

 private:
  TACSMat *mat;
  TACSPc *pc;
  int m, nrestart;
  double rtol, atol;
  int last_iters;
};

// TACSAssembler (src/TACSAssembler.h:59-431), one rank, the GPU context
// behind it.  Per-element thickness is the design vector (one value per
// element, TACSIsoShellConstitutive with tNum = element index); elements
// are set per component by TACSMeshLoader (or setElements).
class TACSAssembler : public TACSObject {
 public:
  // nodes Xpts[3*num_nodes] and 4-node connectivity conn[4*num_elements]
  // in the element node order of the shell basis (tensor order)
  TACSAssembler(int num_nodes, int num_elements, const int *conn, const TacsScalar *Xpts,
                int device = 0);
  ~TACSAssembler();

  int getNumNodes() const { return num_nodes; }
  int getNumElements() const { return num_elements; }
  int getVarsPerNode() const { return 6; }
  shellgpu_ctx *getContext() { return ctx; }

  // one element object per element (the same object may be shared)
  int setElements(TACSElement **elements);
  TACSElement *getElement(int elem) { return elements[elem]; }
  // boundary conditions: node, bitmask of dofs, values[6] per node
  int addBCs(int nbcs, const int *nodes, const int *masks, const TacsScalar *vals);

  TACSBVec *createVec();
  TACSBVec *createNodeVec();
  TACSBVec *createDesignVec();
  // nodal coordinates (TACSAssembler.h:156-157): moving the nodes keeps the
  // connectivity, the matrix pattern and the colouring
  void setNodes(TACSBVec *X);
  void getNodes(TACSBVec *X);
  TACSSchurMat *createSchurMat();
  TACSMat *createMat() { return createSchurMat(); }

  void getDesignVars(TACSBVec *x);
  void setDesignVars(TACSBVec *x);
  // the state and its first and second time derivatives
  // (TACSAssembler.h:183-191)
  void zeroVariables();
  void zeroDotVariables();
  void zeroDDotVariables();
  void setVariables(TACSBVec *q, TACSBVec *qdot = NULL, TACSBVec *qddot = NULL);
  void getVariables(TACSBVec *q, TACSBVec *qdot = NULL, TACSBVec *qddot = NULL);
  TACSBVec *getVariables() { return vars; }
  void applyBCs(TACSBVec *vec);
  void applyBCs(TACSMat *mat);
  void setBCs(TACSBVec *vec);

  // element loops at the current variables (TACSAssembler.h:213-228,
  // TACSAssembler.cpp:4000-4390).  lambda scales the contributions of
  // auxiliary elements in the reference; this path has none (loads are
  // force vectors), so it does not enter.  The shell's element matrices
  // (K, K_t, G, M) are symmetric, so TACS_MAT_TRANSPOSE assembles the same
  // matrices as TACS_MAT_NORMAL; addJacobianVecProduct forms the transposed
  // element products.
  void assembleRes(TACSBVec *residual, const TacsScalar lambda = 1.0);
  void assembleJacobian(TacsScalar alpha, TacsScalar beta, TacsScalar gamma, TACSBVec *residual, TACSMat *A,
                        MatrixOrientation matOr = TACS_MAT_NORMAL, const TacsScalar lambda = 1.0);
  void assembleMatType(ElementMatrixType matType, TACSMat *A, MatrixOrientation matOr = TACS_MAT_NORMAL,
                       const TacsScalar lambda = 1.0);
  void assembleMatCombo(ElementMatrixType matTypes[], TacsScalar scale[], int nmats, TACSMat *A,
                        MatrixOrientation matOr = TACS_MAT_NORMAL, const TacsScalar lambda = 1.0);
  void addJacobianVecProduct(TacsScalar scale, TacsScalar alpha, TacsScalar beta, TacsScalar gamma, TACSBVec *x,
                             TACSBVec *y, MatrixOrientation matOr = TACS_MAT_NORMAL,
                             const TacsScalar lambda = 1.0);

  // sensitivities (the reference's archived assembler API)
  void addAdjointResProducts(TacsScalar scale, int num_adjoints, TACSBVec **adjoint,
                             TACSBVec **dfdx);
  void addAdjointResXptSensProducts(TacsScalar scale, int num_adjoints, TACSBVec **adjoint,
                                    TACSBVec **dfdX);
  // psi != phi through the symmetry of the element matrices (polarization)
  void addMatDVSensInnerProduct(TacsScalar scale, ElementMatrixType matType, TACSBVec *psi,
                                TACSBVec *phi, TACSBVec *dfdx);
  void addMatXptSensInnerProduct(TacsScalar scale, ElementMatrixType matType, TACSBVec *psi,
                                 TACSBVec *phi, TACSBVec *dfdX);
  void evalMatSVSensInnerProduct(ElementMatrixType matType, TACSBVec *psi, TACSBVec *phi,
                                 TACSBVec *res);

  // the element model of all elements (0 linear, 1 nonlinear)
  int getModel() const { return model; }

  // host copies of the mesh and the component of each element (output)
  const std::vector<int> &getConnectivity() const { return conn_h; }
  const std::vector<TacsScalar> &getNodeLocations() const { return X_h; }
  void setComponents(int num_components, const int *elem_comp);
  int getNumComponents() const { return num_comp; }
  int getComponentNum(int elem) const { return comp_h[elem]; }
  // TACSAssembler::getElementOutputData (src/TACSAssembler.cpp:4433-4468)
  // for TACS_BEAM_OR_SHELL_ELEMENT at the current variables: nvals values at
  // each of the 4 nodes of every element (strains, stresses and extras as
  // selected by write_flag); returns 0 on success
  int getElementOutputData(int write_flag, int *len, int *nvals, std::vector<TacsScalar> &data);
  // upload design-dependent data (thickness, temperatures, material) if
  // the element objects or design variables changed
  int syncDesigns();
  void setOrderingSeed(int n, const int *nodes);

 private:
  int check(int rc, const char *what);
  shellgpu_ctx *ctx;
  int num_nodes, num_elements, model;
  std::vector<TACSElement *> elements;
  std::vector<TacsScalar> thick;
  std::vector<TacsScalar> up_thick, up_temp;  // values last uploaded
  TACSBVec *vars, *dvars, *ddvars;  // dvars, ddvars created on first use
  bool ddvars_zero;                 // ddvars is known to be zero
  bool designs_dirty;
  std::vector<int> conn_h, comp_h;
  std::vector<TacsScalar> X_h;
  int num_comp;
  friend class TACSBVec;
};

// TACSLinearBuckling (src/TACSBuckling.h:43-107, TACSBuckling.cpp:51-350
// and the archived sensitivities archive/TACSBuckling.cpp:382-592)
class TACSLinearBuckling : public TACSObject {
 public:
  TACSLinearBuckling(TACSAssembler *assembler, TacsScalar sigma, TACSMat *gmat, TACSMat *kmat,
                     TACSMat *aux_mat, TACSKsm *solver, int max_lanczos_vecs, int num_eigvals,
                     double eig_tol);
  ~TACSLinearBuckling();
  void solve(TACSBVec *rhs = NULL, TACSBVec *u0 = NULL, KSMPrint *ksm_print = NULL);
  TacsScalar extractEigenvalue(int n, TacsScalar *error);
  TacsScalar extractEigenvector(int n, TACSBVec *ans, TacsScalar *error);
  void evalEigenDVSens(int n, TACSBVec *dfdx);
  void evalEigenXptSens(int n, TACSBVec *dfdX);
  void addEigenDVSens(TacsScalar coef, int n, TACSBVec *dfdx);
  void addEigenXptSens(TacsScalar coef, int n, TACSBVec *dfdX);
  void checkEigenvector(int n);
  int getNumEigs() const { return num_eigvals; }
  TacsScalar getSigma() const { return sigma; }

 private:
  TACSAssembler *assembler;
  TACSMat *gmat, *kmat, *aux_mat;
  TACSKsm *solver;
  TacsScalar sigma;
  int max_lanczos_vecs, num_eigvals;
  double eig_tol;
  shellgpu_buckling *buck;
};

// TACSFrequencyAnalysis, Lanczos constructor (src/TACSBuckling.h:128-200,
// src/TACSBuckling.cpp:614-702, 793-1062): K u = lambda M u, lambda = omega^2,
// by shift-invert Lanczos on (K - sigma M)^{-1} M in the M inner product.
// The solver must be bound to kmat, as in the reference.  solve() leaves
// mmat = M and kmat = K - sigma M with BCs, as the reference does; the GPU
// object factors its own copy of the shifted operator.  evalEigenDVSens and
// evalEigenXptSens (declared at src/TACSBuckling.h:153-154) are exact (the
// reference's versions, commented out there, take element central
// differences).
class TACSFrequencyAnalysis : public TACSObject {
 public:
  TACSFrequencyAnalysis(TACSAssembler *assembler, TacsScalar sigma, TACSMat *mmat, TACSMat *kmat, TACSKsm *solver,
                        int max_lanczos, int num_eigvals, double eig_tol);
  ~TACSFrequencyAnalysis();
  TacsScalar getSigma() const { return sigma; }
  void setSigma(TacsScalar s);
  void solve(KSMPrint *ksm_print = NULL, int print_level = 0);
  TacsScalar extractEigenvalue(int n, TacsScalar *error);
  TacsScalar extractEigenvector(int n, TACSBVec *ans, TacsScalar *error);
  void evalEigenDVSens(int n, TACSBVec *dfdx);
  void evalEigenXptSens(int n, TACSBVec *dfdX);
  void checkEigenvector(int n);
  int getNumEigs() const { return num_eigvals; }

 private:
  void recreate();
  TACSAssembler *assembler;
  TACSMat *mmat, *kmat;
  TACSKsm *solver;
  TacsScalar sigma;
  int max_lanczos, num_eigvals;
  double eig_tol;
  shellgpu_frequency *freq;
};

// TACSFH5File writer (src/io/TACSFH5.cpp:22-283), one rank: a header
// (number of components, then for each the length and text of its name with
// the terminating zero) followed by zones, each a header of 5 ints (data
// type, dim1, dim2, zone-name length, variable-names length), the two names
// with their terminating zeros, and dim1 x dim2 values, all in native byte
// order.  Data types: FH5_INT, FH5_DOUBLE, FH5_FLOAT.
class TACSFH5File : public TACSObject {
 public:
  enum FH5DataType { FH5_INT = 0, FH5_DOUBLE = 1, FH5_FLOAT = 2 };
  TACSFH5File();
  ~TACSFH5File();
  int createFile(const char *file_name, int num_components, char **component_names);
  int writeZoneData(char *zone_name, char *var_names, FH5DataType data_name, int dim1, int dim2, void *data);
  void close();

 private:
  FILE *fp;
};

// TACSToFH5 (src/io/TACSToFH5.{h,cpp}) for TACS_BEAM_OR_SHELL_ELEMENT:
// writeToFile writes the connectivity zones (components, ltypes, ptr,
// connectivity), the continuous nodal zone (X,Y,Z / u..rotz / fx..mz as
// FH5_FLOAT) and the element zone (strains, stresses, extras at the element
// nodes, FH5_FLOAT) selected by write_flag, in the reference's order and
// names.  Element values come from the GPU (shellgpu_element_output).
class TACSToFH5 : public TACSObject {
 public:
  TACSToFH5(TACSAssembler *assembler, ElementType elem_type, int write_flag);
  ~TACSToFH5();
  void setComponentName(int comp_num, const char *comp_name);
  int writeToFile(const char *filename);

 private:
  int writeConnectivity(TACSFH5File *file);
  TACSAssembler *assembler;
  ElementType elem_type;
  int write_flag, element_write_flag, nvals;
  std::vector<std::string> component_names;
  std::string variable_names;
};

// TACSMeshLoader (src/io/TACSMeshLoader.{h,cpp}): Nastran bulk-data reader
// for GRID / GRID* / CQUAD4 (and CQUADR) / SPC / SPC* in short, long and
// free-field formats, nodes and elements sorted by their file ids, CQUAD4
// corners permuted to the tensor order [0, 1, 3, 2] (TACSMeshLoader.cpp:
// 569-1098); createTACS applies the first-appearance node renumbering of
// TACSCreator (src/TACSCreator.cpp:1141-1202).
class TACSMeshLoader : public TACSObject {
 public:
  TACSMeshLoader();
#ifdef MPI_VERSION
  // TACSMeshLoader(MPI_Comm) (src/io/TACSMeshLoader.h:63), one rank
  explicit TACSMeshLoader(MPI_Comm comm) : TACSMeshLoader() { tacs_check_one_rank(comm, "TACSMeshLoader"); }
#endif
  ~TACSMeshLoader();
  int scanBDFFile(const char *file_name);
  int getNumComponents() const { return num_components; }
  const char *getElementDescript(int comp);
  void setElement(int comp, TACSElement *elem);
  int getNumNodes() const { return (int)Xpts.size() / 3; }
  int getNumElements() const { return (int)elem_comp.size(); }
  // file-ordered data (after the id sort), before renumbering
  const std::vector<TacsScalar> &getNodes() const { return Xpts; }
  const std::vector<int> &getConnectivity() const { return conn; }
  const std::vector<int> &getComponents() const { return elem_comp; }
  void getBCs(std::vector<int> &nodes, std::vector<int> &masks, std::vector<TacsScalar> &vals) const;
  TACSAssembler *createTACS(int vars_per_node, int device = 0);

 private:
  int num_components;
  std::vector<std::string> comp_elems, comp_descript;
  std::vector<TACSElement *> comp_element;
  std::vector<TacsScalar> Xpts;
  std::vector<int> conn, elem_comp;
  // one entry per SPC card: node, dof index, value
  std::vector<int> bc_node, bc_dof;
  std::vector<TacsScalar> bc_val;
};

}  // namespace shellgpu
