/*
 * hg_oracle.h -- CPU restatement of HyTeG's matrix-free P1 operator / GMG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * CPU baseline of the hyteg_amd build.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker.  The
 * product library (hyteg_amd/csrc) never links or calls it.
 *
 * Data model (shared contract with the product's host<->device transfer):
 * a P1 function on level L is a flat FP64 array in *canonical order*:
 *   [ macro-vertices | macro-edge interiors | macro-face interiors | macro-cell interiors ]
 * each block ordered by primitive index (= HyTeG PrimitiveID creation order,
 * SetupPrimitiveStorage.cpp:98-380) and, inside a primitive, lexicographically
 * in the primitive's own lattice (x fastest), which is the order HyTeG's SOR
 * kernels sweep (P1Operator.hpp:1352-1420, sor_3D_macrocell_P1.cpp:54-75).
 */
#ifndef HG_ORACLE_H
#define HG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_mesh or_mesh;
typedef struct or_p1op or_p1op;

/* DoFType bits (types/types.hpp:36-44). */
enum { OR_INNER = 1, OR_DIRICHLET = 2, OR_NEUMANN = 4, OR_FREESLIP = 8, OR_ALL = 15 };

/* Mesh from vertex coordinates and cells given as 4 vertex ids in the cell's
 * local vertex order.  Cells are sorted by their vertex vector (MeshInfo keys
 * cells_ by it), edges/faces are derived and sorted like MeshInfo::addEdge /
 * addFace (MeshInfo.cpp:37-84).  Boundary flags follow
 * setMeshBoundaryFlagsOnBoundary(1,0,true) + create0123BC: primitives on the
 * boundary are Dirichlet, the rest Inner. */
or_mesh* or_mesh_create( int nv, const double* coords, int nc, const int64_t* cells );
/* MeshInfo::meshCuboid(lowerLeft, upperRight, nx, ny, nz) (MeshGenCuboid.cpp:34-160). */
or_mesh* or_mesh_cuboid( double x0, double y0, double z0, double x1, double y1, double z1, int nx, int ny, int nz );
void     or_mesh_free( or_mesh* m );
void     or_mesh_counts( const or_mesh* m, int64_t* out4 ); /* nv, ne, nf, nc */
void     or_mesh_cell_vertices( const or_mesh* m, int64_t* out ); /* nc*4, sorted cell order */

int64_t or_p1_ndofs( const or_mesh* m, int level );
int64_t or_p1_nbnd( const or_mesh* m, int level );
void    or_p1_coords( const or_mesh* m, int level, double* xyz );
void    or_p1_flags( const or_mesh* m, int level, int32_t* flags );
/* canonical id of cell c's lattice point (x,y,z) */
int64_t or_p1_locate( const or_mesh* m, int level, int c, int x, int y, int z );

or_p1op* or_p1op_create( const or_mesh* m, int level );
void     or_p1op_free( or_p1op* op );
int64_t  or_p1op_nnz( const or_p1op* op );
/* full assembled matrix in CSR (canonical numbering) */
void or_p1op_csr( const or_p1op* op, int64_t* rowptr, int64_t* cols, double* vals );
void or_p1op_cell_stencil( const or_p1op* op, int c, double* out15 );

/* P1Operator::apply (P1Operator.hpp:192-320): dst (=|+=) A src on DoFs whose
 * type matches flag; other dst entries untouched. */
void or_p1_apply( const or_p1op* op, const double* src, double* dst, int flag, int add );
/* P1Operator::smooth_sor forward (P1Operator.hpp:340-418), HyTeG data flow. */
void or_p1_sor( const or_p1op* op, double* x, const double* b, int flag, double relax );

/* P1toP1LinearRestriction::restrict3D with the generated pull kernel
 * (P1toP1LinearRestriction.cpp:169-345, restrict_3D_macrocell_P1_pull_additive.cpp) */
void or_p1_restrict( const or_mesh* m, int fineLevel, const double* fine, double* coarse, int flag );
/* P1toP1LinearProlongation::prolongate3DAdditively (P1toP1LinearProlongation.cpp:194-411) */
void or_p1_prolongate( const or_mesh* m, int coarseLevel, const double* coarse, double* fine, int flag, int add );

double or_dot( int64_t n, const double* a, const double* b, const int32_t* flags, int flag );

/* CGSolver::solve (CGSolver.hpp:94-203), identity preconditioner. returns iterations */
int or_p1_cg( const or_p1op* op, double* x, const double* b, int flag, int maxIter, double relTol, double absTol );

/* GeometricMultigridSolver (GeometricMultigridSolver.hpp:180-308) with
 * GaussSeidelSmoother (SOR relax=1) and CGSolver on minLevel.  Runs `cycles`
 * V-cycles (or until ||r||/||r0|| <= tol when tol>0), writes the residual
 * 2-norm (flag Inner) after each cycle into res_hist[0..cycles] (res_hist[0] =
 * initial).  Returns the number of cycles done. */
int or_p1_gmg( const or_mesh* m, int minLevel, int maxLevel, double* x, const double* b, int pre, int post,
               int cycles, double tol, double* res_hist, int coarseMaxIter, double coarseTol );

/* vertex ids of the macro-primitive owning canonical DoF id (1 vertex, 2 edge
 * vertices, 3 face vertices or 4 cell vertices); returns their number */
int or_p1_owner_vertices( const or_mesh* m, int level, int64_t id, int* vs );
/* One SOR sweep over an assembled CSR matrix with the rows visited colour by
 * colour (colour[p] in 0..ncol-1; backward: colours in reverse), rows with
 * mask[p] == 0 left alone; every row reads the latest values (the multicolour
 * orderings of the GPU Gauss-Seidel sweeps). */
void or_sor_mc_csr( int64_t n, const int64_t* rowptr, const int64_t* cols, const double* vals, const int32_t* colour,
                    int ncol, const unsigned char* mask, double* x, const double* b, double relax, int backward );
/* or_p1_gmg for nvec right-hand sides (vector k at B + k*ndofs), OpenMP over the
 * vectors; cycles_done[k] per vector */
void or_p1_gmg_many( const or_mesh* m, int minLevel, int maxLevel, int nvec, double* X, const double* B, int pre,
                     int post, int cycles, double tol, int coarseMaxIter, double coarseTol, int* cycles_done );

/* ---- P2 on the icosahedral shell ----------------------------------------- */
typedef struct or_p2op or_p2op;

/* MeshInfo::meshSphericalShell(ntan, nrad, rmin, rmax) (SHELLMESH_CLASSIC) with
 * IcosahedralShellMap blending on every cell */
or_mesh* or_mesh_shell( int ntan, int nrad, double rmin, double rmax );
/* per cell: ray[3], ref[3], thr[3], for[3], radRay, radRef */
void or_mesh_blend( const or_mesh* m, double* out14 );
/* F(x) and DF(x) (row-major 3x3, may be NULL) of one cell's shell map */
void or_shell_map( const double* blend14, const double* x, double* Fx, double* DF );
/* the 11 quadrature points (barycentric, 11x4) and weights (reference tet, sum 1/6) */
void or_quadrature( double* pts44, double* w11 );
/* 10x10 element matrix of div(k grad) with k = exp(sum_m a_m sin(om_m.x + ph_m))
 * (nm = 0: k = 1); p = 4 vertex coordinates in computational space; blend may be NULL */
void or_p2_element_matrix( const double* p, const double* blend, int nm, const double* ka, const double* kom,
                           const double* kph, double* A );

/* The reduced element form of the low-precision applies: the affine
 * tetrahedron of the four blended vertices, k constant (k at the blended
 * centroid, rounded to binary16), 4-point degree-2 rule */
void or_p2_reduced_element_matrix( const double* p, const double* blend, int nm, const double* ka, const double* kom,
                                   const double* kph, double* A );
/* the operator's element matrices: 0 the reference form (default), 1 the reduced one */
void or_p2op_set_reduced( or_p2op* op, int on );
/* P2ElementwiseOperator<p2_div_k_grad_blending_q4> on P2 level `level`; its
 * DoFs are the canonical P1 DoFs of level+1 */
or_p2op* or_p2op_create( const or_mesh* m, int level, int nm, const double* ka, const double* kom, const double* kph );
void     or_p2op_free( or_p2op* op );
void     or_p2_apply( const or_p2op* op, const double* src, double* dst, int flag, int add );
void     or_p2_inv_diag( const or_p2op* op, double* out );
int64_t  or_p2op_nelem( const or_p2op* op );
void     or_p2op_coo( const or_p2op* op, int64_t* rows, int64_t* cols, double* vals );

/* ---- P2 multigrid ------------------------------------------------------------- */
/* P2toP2QuadraticProlongation::prolongate (add = 0) / prolongateAndAdd (add = 1):
 * fine = level coarseLevel+1; fine cell-interior DoFs always, macro
 * vertex/edge/face DoFs where their flag is in `flag` */
void or_p2_prolongate( const or_mesh* m, int coarseLevel, const double* coarse, double* fine, int flag, int add );
/* P2toP2QuadraticRestriction::restrict: coarse = P^T fine on coarse interior DoFs and
 * flagged macro vertex/edge/face DoFs */
void or_p2_restrict( const or_mesh* m, int fineLevel, const double* fine, double* coarse, int flag );
/* ChebyshevSmoother::setupCoefficientsInternal(order, lower, upper): coeffs[order] */
int or_chebyshev_coefficients( int order, double lower, double upper, double* coeffs );
/* one ChebyshevSmoother::solve step with the operator's inverse diagonal */
void or_p2_chebyshev_smooth( const or_p2op* op, const double* invDiag, int order, const double* coeffs, double* x,
                             const double* b, int flag );
/* chebyshev::estimateRadius: power iteration on D^-1 A (flag All); x is the start
 * vector and returns the last iterate */
double or_p2_power_iteration( const or_p2op* op, const double* invDiag, double* x, int its );
/* CGSolver::solve with identity (jacobi 0) or inverse-diagonal (jacobi 1) preconditioner */
int or_p2_cg( const or_p2op* op, double* x, const double* b, int flag, int maxIter, double relTol, double absTol,
              int jacobi );
/* V(pre,post) cycles with Chebyshev smoothing of `order` (bounds 0.3 rho, 1.2 rho on
 * every level) and CG at minLevel, until the Inner residual drops by `tol`;
 * res_hist[0..k] (k = cycles done, returned) */
int or_p2_gmg( const or_mesh* m, int minLevel, int maxLevel, int nm, const double* ka, const double* kom,
               const double* kph, double* x, const double* b, int pre, int post, int order, double rho, int cycles,
               double tol, double* res_hist, int coarseMaxIter, double coarseTol, int coarseJacobi );

/* ---- P2-P1 Stokes (Taylor-Hood) ------------------------------------------------ */
typedef struct or_stokesop or_stokesop;
/* element matrices of the blended full-viscous operator, eta = exp(c0 + c1 |F(x)|):
 * A 30x30 (index 10 c + i: component c, P2 function i; may be NULL) and
 * B 4x30 = -(psi_i, div u) (may be NULL); p = 4 computational vertices, blend may be NULL */
void         or_stokes_element( const double* p, const double* blend, double c0, double c1, double* A, double* B );
/* 4x4 P1 mass with k = 1/eta (4-point degree-2 rule, blended) */
void         or_stokes_kmass_element( const double* p, const double* blend, double c0, double c1, double* M );
or_stokesop* or_stokesop_create( const or_mesh* m, int level, double c0, double c1 );
void         or_stokesop_free( or_stokesop* op );
/* mask bit 1: yu = A u, 2: yu += B^T p, 4: yp = B u; u = [ux|uy|uz] (level+1 numbering),
 * p on level; add: UpdateType Add */
void or_stokes_apply( const or_stokesop* op, const double* u, const double* p, double* yu, double* yp, int flag, int mask,
                      int add );
void or_stokes_inv_diag( const or_stokesop* op, double* out );
/* P1PSPGInvDiagOperator and P1LumpedInvMassOperator (affine geometry) of one level */
void or_p1_pspg_inv_diag( const or_mesh* m, int level, double* out );
void or_p1_lumped_inv_mass( const or_mesh* m, int level, double* out );
/* boundary flags by surface: boundary faces with mean vertex radius < rmid get innerFlag */
void or_mesh_set_shell_flags( or_mesh* m, double rmid, int innerFlag, int outerFlag );
/* P2ProjectNormalOperator::project with the radial normal x/|x| on DoFs of type flag */
void or_p2_project_normal( const or_mesh* m, int level, double* u, int flag );

#ifdef __cplusplus
}
#endif
#endif
