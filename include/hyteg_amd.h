/*
 * hyteg_amd.h -- C ABI of the MI355X-native matrix-free operator / multigrid
 * library.  Plain pointers, sizes and integer handles only; no device pointer
 * and no C++ or torch type crosses this boundary.
 *
 * Every call returns 0 on success and a negative status otherwise; the message
 * of the last failure is available from hg_last_error().  A context owns one
 * HIP device, one HIP stream and every object created through it.  Contexts
 * are not re-entrant: calls on one context must not overlap.
 *
 * Each entry point names the HyTeG interface it replaces (paths relative to
 * the HyTeG source tree).
 */
#ifndef HYTEG_AMD_H
#define HYTEG_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hg_ctx hg_ctx;
typedef int64_t       hg_handle;

/* status codes */
enum
{
   HG_OK            = 0,
   HG_ERR_ARG       = -1,
   HG_ERR_HIP       = -2,
   HG_ERR_HANDLE    = -3,
   HG_ERR_STATE     = -4,
   HG_ERR_NOMEM     = -5,
   HG_ERR_UNSUPPORT = -6
};

/* DoFType bits, types/types.hpp:36-44 */
enum
{
   HG_INNER     = 1,
   HG_DIRICHLET = 2,
   HG_NEUMANN   = 4,
   HG_FREESLIP  = 8,
   HG_ALL       = 15
};
/* UpdateType, types/types.hpp:29-33 */
enum
{
   HG_REPLACE = 0,
   HG_ADD     = 1
};
/* storage value types of a function */
enum
{
   HG_F64 = 0,
   HG_F32 = 1,
   HG_F16 = 2
};

/* ---- context (replaces walberla::Environment / MPIManager set-up) -------- */
int         hg_ctx_create( int device, hg_ctx** out );
int         hg_ctx_destroy( hg_ctx* ctx );
const char* hg_last_error( const hg_ctx* ctx );
int         hg_ctx_synchronize( hg_ctx* ctx );
/* HIP graph capture of the work a context enqueues (no counterpart in the
 * reference): everything called between begin and end on this host thread is
 * recorded instead of run (calls that return scalars to the host cannot be
 * captured); the instantiated graph replays it with one launch.  The batched
 * multigrid solvers capture their V-cycle this way unless disabled with
 * hg_ctx_set_graphs(ctx, 0). */
int hg_graph_begin_capture( hg_ctx* ctx );
int hg_graph_end_capture( hg_ctx* ctx, hg_handle* graph );
int hg_graph_launch( hg_ctx* ctx, hg_handle graph );
int hg_graph_destroy( hg_ctx* ctx, hg_handle graph );
int hg_ctx_set_graphs( hg_ctx* ctx, int enable );
int hg_ctx_graphs_enabled( hg_ctx* ctx, int* enabled );
/* bytes of device memory held by the context */
int hg_ctx_memory( hg_ctx* ctx, int64_t* bytes );

/* ---- storage: MeshInfo + SetupPrimitiveStorage + PrimitiveStorage -------- */
/* MeshInfo::meshCuboid(lowerLeft, upperRight, nx, ny, nz)  (mesh/MeshInfo.hpp:562,
 * mesh/MeshGenCuboid.cpp:34-160), then SetupPrimitiveStorage(meshInfo, 1) with
 * setMeshBoundaryFlagsOnBoundary(1, 0, true)
 * (primitivestorage/SetupPrimitiveStorage.hpp:54-56, .cpp:977-988). */
int hg_storage_cuboid( hg_ctx* ctx, const double* lower3, const double* upper3, int nx, int ny, int nz, hg_handle* out );
/* MeshInfo built from vertex coordinates and tetrahedra (local vertex order as
 * given; vertex ids 0..nv-1 in MeshInfo id order), e.g. a Gmsh file
 * (mesh/GmshReaderForMSH22.hpp:60-170), same boundary flagging. */
int hg_storage_create( hg_ctx* ctx, int nv, const double* coords, int nc, const int64_t* cells, hg_handle* out );
/* MeshInfo::meshSphericalShell(ntan, nrad, rmin, rmax) with SHELLMESH_CLASSIC
 * (mesh/MeshInfo.hpp:512-516, mesh/MeshGenSphericalShell.cpp:642-700) and
 * IcosahedralShellMap::setMap on all cells (geometry/IcosahedralShellMap.hpp:491) */
int hg_storage_shell( hg_ctx* ctx, int ntan, int nrad, double rmin, double rmax, hg_handle* out );
/* the same shell with the boundary flags of its two surfaces: the primitives of
 * the inner surface get innerFlag, those of the outer one outerFlag, each one of
 * HG_DIRICHLET, HG_NEUMANN, HG_FREESLIP (the setMeshBoundaryFlagsOnBoundary pair
 * with MeshInfo::hollowFlag::flagInnerBoundary / flagOuterBoundary and a
 * BoundaryCondition mapping them, mesh/MeshInfo.hpp:52-58) */
int hg_storage_shell_bc( hg_ctx* ctx, int ntan, int nrad, double rmin, double rmax, int innerFlag, int outerFlag,
                         hg_handle* out );
/* per macro-cell shell map data: rayVertex, refVertex, thrVertex, forVertex,
 * radRayVertex, radRefVertex (14 doubles per cell, IcosahedralShellMap.hpp:608-731) */
int hg_storage_blend( hg_ctx* ctx, hg_handle storage, double* out14 );
/* macro-cell vertex ids in local order (nc*4) and vertex coordinates (nv*3) */
int hg_storage_cell_vertices( hg_ctx* ctx, hg_handle storage, int64_t* out );
int hg_storage_vertex_coords( hg_ctx* ctx, hg_handle storage, double* out );
/* nv, ne, nf, nc of the macro mesh (PrimitiveStorage::getNumberOf*) */
int hg_storage_counts( hg_ctx* ctx, hg_handle storage, int64_t* out4 );
/* number of P1 DoFs on a level in canonical order (numberOfGlobalDoFs,
 * p1functionspace/VertexDoFFunction.cpp) */
int hg_p1_ndofs( hg_ctx* ctx, hg_handle storage, int level, int64_t* out );
/* coordinates (nd*3) and DoFType (nd) of every canonical P1 DoF */
int hg_p1_dof_coords( hg_ctx* ctx, hg_handle storage, int level, double* xyz );
int hg_p1_dof_flags( hg_ctx* ctx, hg_handle storage, int level, int32_t* flags );

/* P2 DoFs of level L: the micro-vertices and micro-edge midpoints of level L,
 * numbered like the P1 DoFs of level L+1 (P2Function = vertex + edge DoFs,
 * p2functionspace/P2Function.hpp:35-74) */
int hg_p2_ndofs( hg_ctx* ctx, hg_handle storage, int level, int64_t* out );
int hg_p2_dof_coords( hg_ctx* ctx, hg_handle storage, int level, double* xyz );
int hg_p2_dof_flags( hg_ctx* ctx, hg_handle storage, int level, int32_t* flags );

/* ---- functions: P1Function<T> (p1functionspace/VertexDoFFunction.hpp:58-512) */
/* P1Function<T>(name, storage, minLevel, maxLevel), all values zero */
int hg_p1_function_create( hg_ctx* ctx, hg_handle storage, int minLevel, int maxLevel, int dtype, hg_handle* out );
/* A P1 function holding nvec (a power of two up to 64) independent vectors
 * interleaved per DoF: the values of DoF i are stored together (multi-vector
 * layout for batched solves; no counterpart in the reference, which solves one
 * system per call).  upload/download then take host arrays [ndofs][nvec];
 * assign/add/mult/interpolate/copyFrom act on all vectors; interpolate_random
 * gives vector k the seed seed+k. */
int hg_p1_function_create_batched( hg_ctx* ctx, hg_handle storage, int minLevel, int maxLevel, int dtype, int nvec,
                                   hg_handle* out );
/* number of interleaved vectors of a function */
int hg_function_nvec( hg_ctx* ctx, hg_handle fn, int* nvec );
/* vector kdst of dst = vector ksrc of src (all DoFs of `level`, halos included) */
int hg_function_copy_vector( hg_ctx* ctx, hg_handle dst, int kdst, hg_handle src, int ksrc, int level );
/* P2Function<T>(name, storage, minLevel, maxLevel) (p2functionspace/P2Function.hpp:35-74) */
int hg_p2_function_create( hg_ctx* ctx, hg_handle storage, int minLevel, int maxLevel, int dtype, hg_handle* out );
/* A P2 function holding nvec (a power of two up to 64) independent vectors
 * interleaved per DoF, as hg_p1_function_create_batched (no counterpart in the
 * reference): the operand of a batched P2 operator, the P2 transfers and the
 * BLAS-1, halo and dot routines. */
int hg_p2_function_create_batched( hg_ctx* ctx, hg_handle storage, int minLevel, int maxLevel, int dtype, int nvec,
                                   hg_handle* out );
int hg_function_destroy( hg_ctx* ctx, hg_handle fn );
/* storage handle a function lives on (Function::getStorage, functions/Function.hpp:35-80) */
int hg_function_storage( hg_ctx* ctx, hg_handle fn, hg_handle* storage );
/* host canonical FP64 array <-> device (interpolate from host data / getDoFs) */
int hg_function_upload( hg_ctx* ctx, hg_handle fn, int level, const double* canonical );
int hg_function_download( hg_ctx* ctx, hg_handle fn, int level, double* canonical );
/* assign(scalars, functions, level, flag): dst = sum_i s_i f_i (VertexDoFFunction.cpp:1131) */
int hg_function_assign( hg_ctx* ctx, hg_handle dst, int nsrc, const double* scalars, const hg_handle* srcs, int level, int flag );
/* add(scalars, functions, level, flag): dst += sum_i s_i f_i (VertexDoFFunction.cpp:1224) */
int hg_function_add( hg_ctx* ctx, hg_handle dst, int nsrc, const double* scalars, const hg_handle* srcs, int level, int flag );
/* interpolate(constant, level, flag) (VertexDoFFunction.cpp:384) */
int hg_function_interpolate_const( hg_ctx* ctx, hg_handle fn, double value, int level, int flag );
/* interpolate seeded uniform[lo, hi) values on DoFs of matching type: the value
 * of canonical DoF i is SplitMix64(seed + (i+1) * 0x9E3779B97F4A7C15) mapped to
 * [lo, hi), independent of the storage layout (random interpolation,
 * cf. VertexDoFFunction::interpolate with a random expression) */
int hg_function_interpolate_random( hg_ctx* ctx, hg_handle fn, uint64_t seed, double lo, double hi, int level, int flag );
/* dotGlobal(other, level, flag) -- FP64 accumulation (VertexDoFFunction.cpp:1710-1716) */
int hg_function_dot( hg_ctx* ctx, hg_handle a, hg_handle b, int level, int flag, double* out );
/* per-vector dotGlobal of interleaved functions: out[nvec] */
int hg_function_dot_batched( hg_ctx* ctx, hg_handle a, hg_handle b, int level, int flag, double* out );
/* hg_function_dot_batched without allocating device memory.  No counterpart in
 * the reference (its dotGlobal allocates nothing on a device); it is for callers
 * that cannot allocate: code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc.  The partial sums and the results go to a
 * file-scope device array of 65,600 doubles (513 KiB) per GPU, held until the
 * process ends, instead of the context's reduction buffer.  A dot product needs
 * at most 1024 x nvec + nvec of them; larger requests are refused with
 * HG_ERR_ARG.  Every call on a GPU shares that array and it keeps the last
 * call's contents: calls that overlap on one GPU (several contexts or threads)
 * race.  The function's level must already be set up (it is once the function
 * exists). */
int hg_function_dot_batched_noalloc( hg_ctx* ctx, hg_handle a, hg_handle b, int level, int flag, double* out );
/* Device-resident scalars for Krylov loops (no counterpart in the reference,
 * whose dotGlobal returns every dot product to the host): each context holds
 * 256 FP64 slots in device memory.  hg_slots_alloc reserves n consecutive slots
 * (first index in *first), hg_slots_free returns them, hg_slots_read copies n
 * slot values to the host (the one call here that waits for the stream).
 * hg_function_dot_slot writes dotGlobal(a, b, level, flag) of each of the
 * function's vectors into slot, slot+1, ... without waiting.
 * hg_function_lincomb_slots is assign (add = 0) or add (add = 1) of
 * hg_function_assign / hg_function_add with the scalars c_i = signs[i] *
 * (slot[num[i]] / slot[den[i]]) read on the device (a negative index stands
 * for 1; a zero denominator gives c_i = 0); CGSolver uses them to keep alpha and beta on the device
 * (CGSolver.hpp:149-190 computes them on the host). */
int hg_slots_alloc( hg_ctx* ctx, int n, int* first );
int hg_slots_free( hg_ctx* ctx, int first, int n );
int hg_slots_read( hg_ctx* ctx, int first, int n, double* out );
/* slots first .. first+n-1 = values (waits for the copy; refused while the
 * calling thread captures a graph) */
int hg_slots_write( hg_ctx* ctx, int first, int n, const double* values );
int hg_function_dot_slot( hg_ctx* ctx, hg_handle a, hg_handle b, int level, int flag, int slot );
int hg_function_lincomb_slots( hg_ctx* ctx, hg_handle dst, int nsrc, const double* signs, const int* num, const int* den,
                               const hg_handle* srcs, int level, int flag, int add );
/* hg_function_lincomb_slots with every c_i = 0 while slot[gate] != 0 (gate < 0: no
 * gate), and the stopping test of CGSolver::solve (CGSolver.hpp:180-186) evaluated
 * on the device: init != 0 sets slot[done] = (sqrt(slot[rr0]) < absTol); otherwise,
 * while slot[done] == 0, slot[done] = (sqrt(slot[rr]) / sqrt(slot[rr0]) < relTol or
 * sqrt(slot[rr]) < absTol).  A CG loop of fixed length gated by `done` leaves its
 * iterate where the host-tested loop would stop, without waiting for the device
 * (no counterpart in the reference; used by captured V-cycles). */
int hg_function_lincomb_slots_gated( hg_ctx* ctx, hg_handle dst, int nsrc, const double* signs, const int* num,
                                     const int* den, const hg_handle* srcs, int level, int flag, int add, int gate );
int hg_slots_cg_stop( hg_ctx* ctx, int rr, int rr0, double relTol, double absTol, int done, int init );
/* vector k of fn scaled by scales[k] on flagged DoFs */
int hg_function_scale_batched( hg_ctx* ctx, hg_handle fn, const double* scales, int level, int flag );
/* one step of ChebyshevSmoother::solve (ChebyshevSmoother.hpp:186-231) fused
 * into one pass: with r = b - s (b != 0, the first step, s = A x) or r = s (b = 0,
 * s = A t), t = invDiag * r on the flagged DoFs and 0 on all others, then
 * x += c * t on the flagged DoFs; bit-identical to the assign / setToZero /
 * multElementwise / add sequence it replaces */
int hg_function_chebyshev_step( hg_ctx* ctx, hg_handle x, hg_handle t, hg_handle s, hg_handle b, hg_handle invDiag,
                                double c, int level, int flag );
/* the same step with a step size per interleaved vector: vector k takes
 * c = slot[slot0 + k], read on the device when the step runs (a captured graph
 * picks up new values written between its launches; no counterpart in the
 * reference) */
int hg_function_chebyshev_step_slots( hg_ctx* ctx, hg_handle x, hg_handle t, hg_handle s, hg_handle b,
                                      hg_handle invDiag, int slot0, int level, int flag );
/* multElementwise({a, b}, level, flag): dst = a * b (VertexDoFFunction.cpp:1408) */
int hg_function_mult( hg_ctx* ctx, hg_handle dst, hg_handle a, hg_handle b, int level, int flag );
/* copyFrom<otherT>(other, level): cast copy between storage types (VertexDoFFunction.hpp:151) */
int hg_function_copy_from( hg_ctx* ctx, hg_handle dst, hg_handle src, int level );
/* sumGlobal(level, flag) of a single-vector FP64 function (VertexDoFFunction.hpp:310) */
int hg_function_sum( hg_ctx* ctx, hg_handle fn, int level, int flag, double* out );
/* add(scalar, level, flag): fn += value on flagged DoFs (VertexDoFFunction.hpp:195);
 * with hg_function_sum this is vertexdof::projectMean (VertexDoFFunction.hpp:586-592) */
int hg_function_add_const( hg_ctx* ctx, hg_handle fn, double value, int level, int flag );
/* v = 1/v on the flagged DoFs, zeros kept (VertexDoFFunction::invertElementwise,
 * VertexDoFFunction.hpp:252); single-vector FP64 */
int hg_function_invert( hg_ctx* ctx, hg_handle fn, int level, int flag );
/* communicate<Vertex,Edge>..<Face,Cell>: refresh every macro-cell halo copy from
 * its owner (VertexDoFFunction.hpp:366-392, VertexDoFPackInfo.hpp:534-610) */
int hg_function_communicate( hg_ctx* ctx, hg_handle fn, int level );

/* ---- P1 constant-stencil operator (constant_stencil_operator/P1ConstantOperator.hpp:167-168) */
/* P1ConstantLaplaceOperator(storage, minLevel, maxLevel) */
int hg_p1_laplace_create( hg_ctx* ctx, hg_handle storage, int minLevel, int maxLevel, hg_handle* out );
int hg_operator_destroy( hg_ctx* ctx, hg_handle op );
/* Operator::apply(src, dst, level, flag, updateType) (operators/Operator.hpp:74-118,
 * p1functionspace/P1Operator.hpp:192-320) */
int hg_p1_apply( hg_ctx* ctx, hg_handle op, hg_handle src, hg_handle dst, int level, int flag, int update );
/* SORSmoothable::smooth_sor(dst, rhs, relax, level, flag) (solvers/Smoothables.hpp:34-120,
 * P1Operator.hpp:340-418).  Multicolour ordering: macro-vertex/edge/face DoFs
 * in graph colours, then the 8 parity colours of every macro-cell lattice.
 * smooth_gs == relax 1. */
int hg_p1_smooth_sor( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, double relax, int level, int flag );
/* SORBackwardsSmoothable::smooth_sor_backwards(dst, rhs, relax, level, flag)
 * (solvers/Smoothables.hpp:99-110, P1Operator.hpp:331-336,420-427; GSBackwardsSmoothable
 * smooth_gs_backwards == relax 1): the colours of hg_p1_smooth_sor in reverse order,
 * the cell interiors first. */
int hg_p1_smooth_sor_backwards( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, double relax, int level,
                                int flag );
/* The sweep order of hg_p1_smooth_sor on `level` (no counterpart in the reference,
 * which sweeps lexicographically): colour[i] for every canonical DoF i, the owned
 * vertex/edge/face DoFs in 0..nb-1 (*nBoundaryColours = nb), the cell interiors in
 * nb + (x&1) + 2(y&1) + 4(z&1) of their cell-lattice coordinates.  The forward
 * sweep visits the colours in increasing order; no two coupled DoFs share one. */
int hg_p1_sor_colours( hg_ctx* ctx, hg_handle op, int level, int32_t* colour, int* nBoundaryColours );
/* OperatorWithInverseDiagonal::getInverseDiagonalValues (solvers/Smoothables.hpp:112-120,
 * P1Operator.hpp:458-480): 1/a_ii of every DoF of `level` into the single-vector FP64
 * P1 function dst (computed on first use and kept by the operator) */
int hg_p1_inverse_diagonal( hg_ctx* ctx, hg_handle op, int level, hg_handle dst );
/* WeightedJacobiSmoothable::smooth_jac(dst, rhs, src, relax, level, flag)
 * (solvers/Smoothables.hpp:34-57, P1Operator.hpp:428-447):
 * dst = src + relax * D^-1 (rhs - A src) on the DoFs of type flag */
int hg_p1_smooth_jac( hg_ctx* ctx, hg_handle op, hg_handle dst, hg_handle rhs, hg_handle src, double relax, int level,
                      int flag );
/* Operator::gemv(alpha, src, beta, dst, level, flag) (operators/Operator.hpp:101-118):
 * dst = alpha A src + beta dst on the DoFs of type flag; applyScaled(alpha, src, dst,
 * level, flag, updateType) (Operator.hpp:83-99) is beta = 0 (Replace) or 1 (Add) */
int hg_p1_gemv( hg_ctx* ctx, hg_handle op, double alpha, hg_handle src, double beta, hg_handle dst, int level,
                int flag );
/* stencil weights of the cell interior (15, same order as
 * apply_3D_macrocell_vertexdof_to_vertexdof_replace.cpp:36-50) */
int hg_p1_cell_stencil( hg_ctx* ctx, hg_handle op, int level, int cell, double* out15 );

/* ---- P2 elementwise div(k grad) operator with blending -------------------- */
/* P2ElementwiseOperator<forms::p2_div_k_grad_blending_q4>(storage, minLevel, maxLevel, form)
 * (elementwiseoperators/P2ElementwiseOperator.hpp:61,475; form
 * forms/form_hyteg_generated/p2/p2_div_k_grad_blending_q4.hpp) with the
 * coefficient k(x) = exp(sum_m a[m] sin(omega[3m..3m+2] . x + phi[m])), nm <= 32
 * (nm = 0: k = 1), evaluated at the blended quadrature points. */
int hg_p2_divkgrad_create( hg_ctx* ctx, hg_handle storage, int minLevel, int maxLevel, int nm, const double* a,
                           const double* omega, const double* phi, hg_handle* out );
/* Batched form of hg_p2_divkgrad_create (no counterpart in the reference, which
 * solves one system per call): nslot (a power of two, 2..64) coefficient slots,
 * each with its own k (initially k = 1) and its own stored quadrature-point
 * table, for functions of nslot interleaved vectors (hg_p2_function_create_batched):
 * vector s is multiplied by the operator of slot s, with the kernels and the
 * per-vector results of the single-vector operator.  FP64, update Replace with
 * Inner in the flag; hg_p2_inverse_diagonal gives the slots' diagonals
 * interleaved.  hg_p2_divkgrad_set_slot replaces the k of one slot (its tables
 * and inverse diagonals are rebuilt on the operator's stream). */
int hg_p2_divkgrad_create_batched( hg_ctx* ctx, hg_handle storage, int minLevel, int maxLevel, int nslot,
                                   hg_handle* out );
int hg_p2_divkgrad_set_slot( hg_ctx* ctx, hg_handle op, int slot, int nm, const double* a, const double* omega,
                             const double* phi );
/* active = 0: applies of the batched operator skip slot `slot` and give it a zero
 * result (for slots whose vectors are zero, as the idle slots of a serving
 * batch); active = 1 (every slot's initial state) computes it again.  Read on
 * the device when the apply runs (a captured graph follows it). */
int hg_p2_divkgrad_slot_active( hg_ctx* ctx, hg_handle op, int slot, int active );
/* Operator::apply = gemv(1, src, 0|1, dst, level, flag) (P2ElementwiseOperator.cpp:110-328) */
int hg_p2_apply( hg_ctx* ctx, hg_handle op, hg_handle src, hg_handle dst, int level, int flag, int update );
/* P2ElementwiseOperator::gemv(alpha, src, beta, dst, level, flag): dst = alpha A src + beta dst
 * on the DoFs of type flag (P2ElementwiseOperator.cpp:110-233, Operator.hpp:101-118);
 * applyScaled(alpha, src, dst, level, flag, updateType) is beta 0 / 1 (.cpp:331-339) */
int hg_p2_gemv( hg_ctx* ctx, hg_handle op, double alpha, hg_handle src, double beta, hg_handle dst, int level,
                int flag );
/* WeightedJacobiSmoothable::smooth_jac(dst, rhs, src, relax, level, flag)
 * (P2ElementwiseOperator.cpp:342-371): dst = src + relax * D^-1 (rhs - A src) on the
 * DoFs of type flag */
int hg_p2_smooth_jac( hg_ctx* ctx, hg_handle op, hg_handle dst, hg_handle rhs, hg_handle src, double relax, int level,
                      int flag );
/* P2ElementwiseOperator::computeAndStoreLocalElementMatrices counterpart
 * (elementwiseoperators/P2ElementwiseOperator.hpp:169,375-385): instead of the
 * 10x10 element matrices (100 values per micro-cell) the operator keeps k at the
 * 11 quadrature points of every micro-cell (11 values, FP64), built on the first
 * apply of a level.  enable = 1 (default) stores, 0 evaluates k in every apply
 * and frees the tables.  Results are identical either way. */
int hg_p2_store_coefficients( hg_ctx* ctx, hg_handle op, int enable );
/* Element form of the applies on FP32 and FP16 functions (the preconditioner of the
 * mixed-precision refinement; no counterpart in the reference).  reduced = 1 (the
 * default, from level 2): on levels >= fromLevel every micro-cell is integrated as
 * the affine tetrahedron of its four blended vertices, with k constant on it (its
 * value at the blended centroid, stored in FP16) and the 4-point degree-2 rule, in
 * FP32 arithmetic with FP32 accumulation (FP16 results are rounded once); the
 * operator differs from the reference's by O(h) (0.2 % of an element matrix on the
 * shell at level 6).  reduced = 0: the reference form of hg_p2_apply in FP32
 * arithmetic on every level.  FP64 applies always use the reference form. */
int hg_p2_low_precision_form( hg_ctx* ctx, hg_handle op, int reduced, int fromLevel );
/* OperatorWithInverseDiagonal::getInverseDiagonalValues on `level`, written
 * into the P2 function dst (P2ElementwiseOperator.cpp:418-568) */
int hg_p2_inverse_diagonal( hg_ctx* ctx, hg_handle op, int level, hg_handle dst );

/* ---- grid transfer ------------------------------------------------------ */
/* P1toP1LinearRestriction::restrict(f, sourceLevel, flag)
 * (gridtransferoperators/P1toP1LinearRestriction.cpp:169-345) */
int hg_p1_restrict( hg_ctx* ctx, hg_handle fn, int sourceLevel, int flag );
/* P1toP1LinearProlongation::prolongate / prolongateAndAdd(f, sourceLevel, flag)
 * (gridtransferoperators/P1toP1LinearProlongation.hpp:34-58, .cpp:194-411) */
int hg_p1_prolongate( hg_ctx* ctx, hg_handle fn, int sourceLevel, int flag, int update );

/* P2toP2QuadraticRestriction::restrict(f, sourceLevel, flag)
 * (gridtransferoperators/P2toP2QuadraticRestriction.cpp:131-290) */
int hg_p2_restrict( hg_ctx* ctx, hg_handle fn, int sourceLevel, int flag );
/* P2toP2QuadraticProlongation::prolongate / prolongateAndAdd(f, sourceLevel, flag)
 * (gridtransferoperators/P2toP2QuadraticProlongation.cpp:37-63,217-425) */
int hg_p2_prolongate( hg_ctx* ctx, hg_handle fn, int sourceLevel, int flag, int update );
/* Inspection of the translation-invariant P2 transfer stencils (no
 * counterpart in the reference; used by tests).  kind 0: pull stencil of the
 * fine class cls = (x&3) + 4(y&3) + 16(z&3), offsets from 2*(p>>2) in coarse
 * lattice units; kind 1: restriction stencil of the coarse class
 * cls = (x&1) + 2(y&1) + 4(z&1), offsets from 2*p in fine lattice units.
 * Writes min(count, max) entries. */
int hg_p2_transfer_stencil( int kind, int cls, int max, int32_t* offsets, double* weights, int* count );

/* ---- solvers (host control flow, device kernels) ------------------------ */
/* CGSolver(storage, minLevel, maxLevel, maxIter, relTol, absTol).solve(A, x, b, level)
 * (solvers/CGSolver.hpp:36-203).  *iterations receives the iteration count. */
int hg_p1_cg_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int level, int maxIter, double relTol, double absTol,
                    int* iterations );
/* GeometricMultigridSolver(storage, GaussSeidelSmoother, CGSolver(minLevel), P1toP1LinearRestriction,
 * P1toP1LinearProlongation, minLevel, maxLevel, pre, post).solve(A, x, b, maxLevel), called
 * `cycles` times or until ||b - Ax||_Inner <= tol * ||b - Ax0||_Inner (tol > 0).
 * res_hist (cycles+1 entries, may be NULL) receives the residual norms.  When
 * the context allows graphs (hg_ctx_set_graphs, the default) the cycles after the
 * first are replayed from a HIP graph and the coarse CG runs a fixed count,
 * min(coarseMaxIter, number of free coarse DoFs), with its coarseTol stopping test
 * evaluated on the device (the iterations after it holds leave the iterate alone).
 * (solvers/GeometricMultigridSolver.hpp:126-308, solvers/GaussSeidelSmoother.hpp:31-55) */
int hg_p1_gmg_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel, int pre, int post,
                     int cycles, double tol, int coarseMaxIter, double coarseTol, double* res_hist, int* cycles_done );

/* ---- timing tree (walberla::WcTimingTree as HyTeG's solvers use it,
 * GeometricMultigridSolver.hpp:200-306): nested named regions on the context's
 * stream, each bracketed by two HIP events; while enabled, the C++ solvers open
 * "Geometric Multigrid Solver", "Level <L>", "Smoother", "Restriction",
 * "Prolongation" and "Coarse Grid Solver" like the reference.  The report lists
 * one line per region path ("A/B/C"): path, tab, count, tab, total ms (device
 * time), newline; *needed receives the report's size including the final zero.
 * Regions opened during a graph capture are not recorded. */
int hg_timing_enable( hg_ctx* ctx, int enable );
int hg_timing_start( hg_ctx* ctx, const char* name );
int hg_timing_stop( hg_ctx* ctx, const char* name );
int hg_timing_report( hg_ctx* ctx, char* buf, int64_t size, int64_t* needed );
int hg_timing_reset( hg_ctx* ctx );

/* ---- device time (ms, HIP events on the context stream) of the dominant
 * kernel of the last entry point, recorded while event timing is enabled ---- */
int hg_ctx_event_timing( hg_ctx* ctx, int enable );
int hg_ctx_last_kernel_ms( hg_ctx* ctx, double* ms );

/* CGSolver<P2ElementwiseOperator>(storage, level, level, maxIter, relTol, absTol,
 * prec) on `level` (solvers/CGSolver.hpp:94-203); jacobi = 1 uses the operator's
 * inverse diagonal as preconditioner (ApplyPreconditioner over the
 * inverse-diagonal multiplication, SubstitutePreconditioner.hpp:120-158) */
int hg_p2_cg_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int level, int maxIter, double relTol,
                    double absTol, int jacobi, int* iterations );
/* `steps` ChebyshevSmoother::solve calls on `level` (solvers/ChebyshevSmoother.hpp:112-232)
 * after setupCoefficients(order, spectralRadius) with bounds 0.3/1.2 x radius */
int hg_p2_chebyshev_smooth( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int level, int order,
                            double spectralRadius, int steps );
/* chebyshev::estimateRadius(A, level, iterations, storage, x, tmp)
 * (ChebyshevSmoother.hpp:656-665, numerictools/SpectrumEstimation.hpp:56-90):
 * power iteration on D^-1 A from the start vector x (overwritten) */
int hg_p2_estimate_radius( hg_ctx* ctx, hg_handle op, hg_handle x, int level, int iterations, double* radius );
/* GeometricMultigridSolver<P2ElementwiseOperator> with ChebyshevSmoother
 * (setupCoefficients(order, spectralRadius), bounds 0.3/1.2 x radius on every
 * level), P2 quadratic transfer and (preconditioned) CG on minLevel, run as V(pre,
 * post) cycles until the Inner residual norm drops by `tol` or `cycles` are done;
 * res_hist[0..done] receives the residual norms */
int hg_p2_gmg_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel, int pre, int post,
                     int chebyshevOrder, double spectralRadius, int cycles, double tol, int coarseMaxIter,
                     double coarseTol, int coarseJacobi, double* res_hist, int* cycles_done );

/* Mixed-precision iterative refinement around the P2 V-cycle
 * (tests/hyteg/mixedPrecision/mixedPrecisionIterativeRefinement.cpp:163-172):
 * each outer step forms r = b - A x in FP64, casts it (scaled so that D^-1 r has
 * unit RMS) to
 * lowDtype (HG_F32 or HG_F16 storage, FP32 arithmetic), runs innerCycles
 * V(pre,post) cycles of the low-precision hierarchy (Chebyshev(order) with the
 * given radius, quadratic transfer, FP64 Jacobi-PCG on minLevel to coarseTol) and
 * adds the correction in FP64.  Stops when the FP64 Inner residual norm drops by
 * `tol` or after maxOuter steps; res_hist[0..done] gets the residual norms. */
int hg_p2_ir_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel, int lowDtype,
                    int pre, int post, int chebyshevOrder, double spectralRadius, int innerCycles, int maxOuter,
                    double tol, int coarseMaxIter, double coarseTol, double* res_hist, int* outer_done );

/* Batched solves of the P1 Laplacian for all vectors of interleaved functions x, b:
 * GMG V(pre,post) cycles with Gauss-Seidel smoothing (GeometricMultigridSolver,
 * P1toP1Linear transfer), coarseSweeps Gauss-Seidel sweeps on minLevel, until every
 * vector's Inner residual drops by tol; res_hist[(cycle) * nvec + k] */
int hg_p1_batched_gmg_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel, int pre,
                             int post, int coarseSweeps, int cycles, double tol, double* res_hist, int* cycles_done );
/* The same inside FP64 iterative refinement, the V-cycles running on lowDtype
 * (HG_F32 / HG_F16 storage, FP32 arithmetic) copies; each vector's residual is
 * scaled separately before the cast */
int hg_p1_batched_ir_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel,
                            int lowDtype, int pre, int post, int coarseSweeps, int innerCycles, int maxOuter, double tol,
                            double* res_hist, int* outer_done );

/* ---- serving --------------------------------------------------------------- */
/* Runs a stream of n solve requests on `device` with continuous batching (no
 * counterpart in the reference).  Request i of kind[i] (0: P1 Poisson on the unit
 * cube, 1: P1 Poisson on the shell macro mesh, 2: P2 div(k grad) on the blended
 * shell, 3: P2 div(k grad) on the cube, 4: isoviscous P2-P1 Stokes on the blended
 * shell with no-slip surfaces and the buoyancy of 12 Gaussian blobs, levels 2-4;
 * k, blob centres and the RHS generated from seed[i])
 * at level[i] arrives arrival_s[i] seconds after the call starts and is solved to
 * relative residual `tol` (Stokes: at most 150 Uzawa V-cycles; its residual floors
 * near 1e-7 to 5e-7, so a tol below 1e-6 runs to that cap).  P1 requests of
 * one (kind, level) share an interleaved hierarchy of `slots` vectors (a power of
 * two up to 64), P2 requests one of 8 vectors with a coefficient slot each, and
 * join / leave it at V-cycle boundaries; every class has its own stream (run the
 * process with GPU_MAX_HW_QUEUES=16 set before HIP initialises, so that the
 * classes' streams do not share HIP's default 4 hardware queues); a class
 * hierarchy is built only while the device memory of the
 * live ones stays below memBudgetGB (an estimate until a worker has built its
 * hierarchy, then the bytes its context actually holds).  Outputs per request:
 * latency (arrival to completion), reached relative residual, checksum x.x of the
 * (velocity) solution, V-cycles; wall_s = total time. */
int hg_serve_stream( int device, int64_t n, const int* kind, const int* level, const uint64_t* seed,
                     const double* arrival_s, double tol, int slots, double memBudgetGB, double* latency_s,
                     double* rel_residual, double* checksum, int* cycles, double* wall_s );
/* hg_serve_stream that also copies request i's solution (canonical order, FP64, the
 * ndofs of its kind and level: P1 DoFs of `level` for kinds 0/1, P2 DoFs for 2/3,
 * for kind 4 the three P2 velocity components followed by the P1 pressure) into
 * solutions[i] when that pointer is not null (solutions itself may be null) */
int hg_serve_stream_solutions( int device, int64_t n, const int* kind, const int* level, const uint64_t* seed,
                               const double* arrival_s, double tol, int slots, double memBudgetGB, double* latency_s,
                               double* rel_residual, double* checksum, int* cycles, double* wall_s,
                               double* const* solutions );
/* the coefficient k(x) = exp(sum_m a_m sin(omega_m . x + phi_m)), m < 16, the stream
 * generates for a diffusion request of the given seed */
int hg_serve_request_coefficient( uint64_t seed, double* a16, double* omega48, double* phi16 );
/* the 12 blob centres (x, y, z each) the stream generates for a Stokes request of
 * the given seed */
int hg_serve_request_blobs( uint64_t seed, double* centres36 );
/* message of the last hg_serve_stream failure on this thread */
const char* hg_serve_last_error( void );

/* ---- P2-P1 Taylor-Hood Stokes -------------------------------------------------------
 * P2P1ElementwiseBlendingFullViscousStokesOperator(storage, minLevel, maxLevel, viscosity)
 * (elementwiseoperators/P2P1ElementwiseBlendingStokesOperator.hpp:125-216) with
 * viscosity eta(x) = exp(viscC0 + viscC1 |x|) at the physical (blended) point. */
int hg_stokes_create( hg_ctx* ctx, hg_handle storage, int minLevel, int maxLevel, double viscC0, double viscC1,
                      hg_handle* out );
/* Blocks of the Stokes apply: src/dst = {u_x, u_y, u_z, p} (FP64 P2 velocity
 * components and FP64 P1 pressure of one storage); mask bit 1: dst.uvw = A src.uvw
 * (P2ElementwiseBlendingFullViscousOperator::apply), bit 2: dst.uvw += B^T src.p
 * (P1ToP2ElementwiseBlendingDivTOperator), bit 4: dst.p = B src.uvw
 * (P2ToP1ElementwiseBlendingDivOperator); mask 7 is the operator's apply
 * (P2P1ElementwiseBlendingStokesOperator.hpp:176-184).  Mask 8 (alone):
 * dst.p = M_k src.p, the P1 mass with k = 1/eta at the blended points (the
 * Schur-complement operator of the free-slip wrapper, a P1 "K-mass",
 * terraneo/operators/P2P1StokesOperatorWithWrapper.hpp:388-395), 4-point
 * degree-2 rule.  Handles of unused
 * components are ignored.  The pressure has no boundary DoFs (createAllInnerBC):
 * it is written when flag contains Inner. */
int hg_stokes_apply( hg_ctx* ctx, hg_handle op, const hg_handle* src, const hg_handle* dst, int level, int flag, int mask,
                     int update );
/* inverse diagonal of the velocity block (VectorToVectorOperator::
 * computeInverseDiagonalOperatorValues) into dst = {d_x, d_y, d_z} */
int hg_stokes_velocity_inverse_diagonal( hg_ctx* ctx, hg_handle op, int level, const hg_handle* dst );
/* kind 0: P1PSPGInvDiagOperator values (the Uzawa smoother's pressure scaling),
 * kind 1: P1LumpedInvMassOperator values (the MINRES pressure preconditioner)
 * (constant_stencil_operator/P1ConstantOperator.hpp:195-215), affine geometry */
int hg_stokes_pressure_diagonal( hg_ctx* ctx, hg_handle op, int level, int kind, hg_handle dst );
/* P2ProjectNormalOperator::project (p2functionspace/P2ProjectNormalOperator.hpp:34-61)
 * with the radial normal n = x/|x|: u -= (u.n) n on the P2 DoFs of u = {u_x, u_y, u_z}
 * (FP64) whose boundary type is in flag */
int hg_p2_project_normal( hg_ctx* ctx, const hg_handle* u, int level, int flag );
/* the free-slip wrapper of the Stokes operator (P2P1FullStokesFreeslipWrapperTemplate
 * with P2ProjectNormalOperator, terraneo/operators/P2P1StokesOperatorWithWrapper.hpp:41-386),
 * P = the normal projection on FreeslipBoundary DoFs: mask 1: dst.uvw = P A P src.uvw,
 * 2: dst.uvw = P B^T src.p, 4: dst.p = B P src.uvw, 7: the full apply on (P u, p),
 * P on the velocity result, and the pressure mean of dst projected out */
int hg_stokes_freeslip_apply( hg_ctx* ctx, hg_handle op, const hg_handle* src, const hg_handle* dst, int level,
                              int flag, int mask, int update );
/* In the solver calls below freeslip != 0 selects the free-slip wrapper as the
 * operator, and with it the projections of stokesGMGUzawaFSSolver: Chebyshev
 * corrections (ChebyshevSmoother<., P2ProjectNormalOperator>), the Uzawa velocity
 * residual (UzawaSmootherWithFreeSlipProjection) and both grid transfers
 * (P2P1StokesToP2P1Stokes{Restriction,Prolongation}WithFreeSlipProjection). */
/* chebyshev::estimateRadius on the velocity block (ChebyshevSmoother.hpp:656-665):
 * x = {x_x, x_y, x_z} is the start vector and returns the last iterate */
int hg_stokes_estimate_radius( hg_ctx* ctx, hg_handle op, const hg_handle* x, const hg_handle* tmp, int level,
                               int iterations, int freeslip, double* radius );
/* one UzawaSmoother::solve (solvers/UzawaSmoother.hpp:330-470, or
 * UzawaSmootherWithFreeSlipProjection, UzawaSmoother.hpp:480-613) with a Chebyshev
 * velocity smoother of the given order and radius inside
 * FullStokesVelocityBlockBlockDiagonalPreconditioner */
int hg_stokes_uzawa_smooth( hg_ctx* ctx, hg_handle op, const hg_handle* x, const hg_handle* b, int level, double omega,
                            int velocityIterations, int chebyshevOrder, double spectralRadius, int projectPressureMean,
                            int freeslip );
/* MinResSolver with StokesPressureBlockPreconditioner<P1LumpedInvMassOperator>
 * (solvers/MinresSolver.hpp:117-245, solvertemplates/StokesSolverTemplates.hpp:54-69) */
int hg_stokes_minres_solve( hg_ctx* ctx, hg_handle op, const hg_handle* x, const hg_handle* b, int level, int maxIter,
                            double relTol, double absTol, int freeslip, int* iterations );
/* GeometricMultigridSolver<Stokes> as stokesGMGUzawaFSSolver builds it
 * (solvertemplates/StokesFSGMGUzawaSolverTemplate.hpp:63-190): Uzawa smoothing,
 * P2P1StokesToP2P1Stokes transfers, V(pre,post) with smoothIncrement more steps
 * per coarser level, MINRES on minLevel.  Runs V-cycles (projecting the pressure
 * mean after each when asked) until the residual 2-norm over
 * Inner | Neumann | FreeSlip drops by tol or `cycles` ran; res_hist[0..done] */
int hg_stokes_gmg_solve( hg_ctx* ctx, hg_handle op, const hg_handle* x, const hg_handle* b, int minLevel, int maxLevel,
                         int pre, int post, int smoothIncrement, double omega, int velocityIterations, int chebyshevOrder,
                         double spectralRadius, int coarseMaxIter, double coarseTol, int projectPressureMean,
                         int freeslip, int cycles, double tol, double* res_hist, int* cycles_done );
/* FGMRESSolver (solvers/FGMRESSolver.hpp:55-500) with the block-factorisation
 * preconditioner of StokesMCFGMRESSolver (terraneo/solvers/StokesMCFGMRESSolver.hpp:50-160,
 * BlockFactorisationPreconditioner, StokesBlockPreconditioners.hpp:366-448):
 * velocity block by one GMG V(aPre, aPost) cycle (Chebyshev of the given order
 * on [0.1, 1.2] x radius, P2toP2QuadraticVector transfers (...WithProjection
 * on the FreeSlip DoFs when freeslip), MINRES with identity
 * preconditioner on minLevel), Schur complement by CG on the 1/eta pressure
 * mass (mask 8 of hg_stokes_apply) preconditioned with its inverse lumped
 * diagonal.  Iterates until the approximate residual drops by relTol or maxIter
 * iterations ran (restart every `restart`); res_hist[0..iterations] holds the
 * initial and the approximate residual 2-norms over Inner | Neumann | FreeSlip.
 * velocityDtype HG_F32 / HG_F16 runs the velocity V-cycle in FP32 arithmetic on
 * FP32 / FP16 vectors (coarsest level in FP64) inside the FP64 iteration: the
 * mixed-precision variant; HG_F64 runs it in FP64. */
int hg_stokes_fgmres_solve( hg_ctx* ctx, hg_handle op, const hg_handle* x, const hg_handle* b, int minLevel,
                            int maxLevel, int aPre, int aPost, int chebyshevOrder, double spectralRadius,
                            int aCoarseMaxIter, double aCoarseTol, int schurMaxIter, double schurTol, int freeslip,
                            int velocityDtype, int maxIter, int restart, double relTol, double* res_hist,
                            int* iterations );

/* ---- P2 constant-stencil Laplacian ------------------------------------------------------
 * P2ConstantLaplaceOperator(storage, minLevel, maxLevel)
 * (constant_stencil_operator/P2ConstantOperator.cpp:100-1350): the P2 Laplacian of the
 * affine macro-cells (blending is not applied), one element matrix per micro-cell type
 * and macro-cell. */
int hg_p2_const_create( hg_ctx* ctx, hg_handle storage, int minLevel, int maxLevel, hg_handle* out );
/* Operator::apply on FP64 P2 functions (P2ConstantOperator::apply) */
int hg_p2_const_apply( hg_ctx* ctx, hg_handle op, hg_handle src, hg_handle dst, int level, int flag, int update );
/* SORSmoothable::smooth_sor (P2ConstantOperator.cpp:1267-1330; relax 1 is smooth_gs):
 * the owned macro vertex/edge/face DoFs in a greedy multicolour order, then the cell
 * interiors in 27 colours */
int hg_p2_const_smooth_sor( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int level, int flag, double relax );
/* SORBackwardsSmoothable::smooth_sor_backwards (P2ConstantOperator.hpp:95-106; relax 1 is
 * smooth_gs_backwards, :67-77): the colours of hg_p2_const_smooth_sor in reverse order */
int hg_p2_const_smooth_sor_backwards( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int level, int flag,
                                      double relax );
/* the sweep order of hg_p2_const_smooth_sor (no counterpart in the reference): colour[i]
 * for every canonical P2 DoF i, the owned vertex/edge/face DoFs in 0..nb-1
 * (*nBoundaryColours = nb), the cell interiors in nb + (X%3) + 3(Y%3) + 9(Z%3) of their
 * level-(L+1) lattice coordinates */
int hg_p2_const_sor_colours( hg_ctx* ctx, hg_handle op, int level, int32_t* colour, int* nBoundaryColours );
/* OperatorWithInverseDiagonal values into dst (every DoF) */
int hg_p2_const_inverse_diagonal( hg_ctx* ctx, hg_handle op, int level, hg_handle dst );
/* GeometricMultigridSolver<P2ConstantLaplaceOperator> with GaussSeidelSmoother, CGSolver
 * (no preconditioner) on minLevel and quadratic transfers, the set-up of
 * P2GMG3DConvergenceTest.cpp:60-100; V(pre,post) cycles until the Inner residual drops
 * by tol or `cycles` ran; res_hist[0..done] */
int hg_p2_const_gmg_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel, int pre,
                           int post, int coarseMaxIter, double coarseTol, int cycles, double tol, double* res_hist,
                           int* cycles_done );

/* ---- VTK output ----------------------------------------------------------------------
 * VTKOutput::write for P1 or P2 functions (dataexport/VTKOutput/VTKOutput.cpp,
 * VTKP1Writer.cpp:34-120, VTKP2Writer.cpp:34-160): one XML unstructured-grid file
 * with every macro-cell's lattice points (mapped by the blending) and the level's
 * micro-tetrahedra (VTK_TETRA for P1, VTK_QUADRATIC_TETRA for P2), one point-data
 * array per function named names[i].  The functions are single-vector functions
 * of one space on the storage; binary != 0 writes VTK's inline base64 format. */
int hg_vtk_write( hg_ctx* ctx, const char* path, hg_handle storage, int level, int nfun, const hg_handle* functions,
                  const char* const* names, int binary );

#ifdef __cplusplus
}
#endif
#endif
