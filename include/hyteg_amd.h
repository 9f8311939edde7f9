/*
 * hyteg_amd.h — C ABI of the MI355X (gfx950) matrix-free P1 multigrid library.
 *
 * Every entry point returns an int status (HYG_OK = 0) and, on failure, stores a message
 * retrievable with hyg_last_error() (thread-local). No exception crosses this boundary.
 * Streams are passed as `void*` holding a hipStream_t (NULL = the default stream).
 * Handles are opaque. Functions and operators live on the device selected by hyg_init().
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * reference repository root, src/hyteg/...).
 */
#ifndef HYTEG_AMD_H
#define HYTEG_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define HYG_OK 0
#define HYG_ERR_ARG 1     /* bad handle, level out of range, size mismatch, unsupported option */
#define HYG_ERR_HIP 2     /* a HIP runtime call failed */
#define HYG_ERR_IO 3      /* mesh file missing or malformed */
#define HYG_ERR_NUMERIC 4 /* non-finite value detected (zero diagonal, NaN residual) */
#define HYG_ERR_STATE 5   /* library not initialised or object in the wrong state */

/* DoFType bit flags, types/types.hpp:29-59 */
#define HYG_NONE 0
#define HYG_INNER 1
#define HYG_DIRICHLET 2
#define HYG_NEUMANN 4
#define HYG_FREESLIP 8
#define HYG_ALL 15

/* UpdateType, types/types.hpp:61-72 */
#define HYG_REPLACE 0
#define HYG_ADD 1

/* value types of a function */
#define HYG_F64 8
#define HYG_F32 4

typedef struct hyg_mesh_s* hyg_mesh;
typedef struct hyg_func_s* hyg_func;
typedef struct hyg_op_s*   hyg_op;
typedef struct hyg_gmg_s*  hyg_gmg;
typedef struct hyg_batch_s* hyg_batch;

/* Last error message of the calling thread ("" if none). */
const char* hyg_last_error( void );

/* Select the device and create the library's per-device state. Must precede all other calls. */
int hyg_init( int device );

/* ---------------------------------------------------------------- mesh
 * Replaces MeshInfo::fromGmshFile (mesh/MeshGenFromGmshFile.cpp:32) followed by
 * SetupPrimitiveStorage(meshInfo, 1) + setMeshBoundaryFlagsOnBoundary(1, 0, true)
 * (primitivestorage/SetupPrimitiveStorage.cpp:40-365, 781-792) and PrimitiveStorage.
 * Boundary primitives get mesh flag 1 = Dirichlet (boundary/BoundaryConditions.hpp create0123BC). */
int hyg_mesh_create_from_msh( const char* path, hyg_mesh* out );
/* MeshInfo::fromGmshFile + SetupPrimitiveStorage with the file's physical tags as mesh boundary flags
 * (mesh/MeshGenFromGmshFile.cpp:100-116, 160-200; keep_tags = 1: no setMeshBoundaryFlagsOnBoundary): the edge of
 * a line element and the face of a triangle element get the element's first tag, every other primitive 0; flags
 * map to DoF types as BoundaryCondition::create0123BC (boundary/BoundaryConditions.cpp:26-33): 1 Dirichlet,
 * 2 Neumann, 3 free-slip, others Inner. keep_tags = 0 is hyg_mesh_create_from_msh. */
int hyg_mesh_create_from_msh_tags( const char* path, int keep_tags, hyg_mesh* out );
/* Replaces MeshInfo::meshSphericalShell( ntan, nrad, rmin, rmax ) (mesh/MeshGenSphericalShell.cpp, default
 * SHELLMESH_CLASSIC: arc-midpoint refinement of the icosahedron's diamonds, ntan - 1 a power of two,
 * nrad >= 2 equidistant radii) followed by the same storage setup and boundary flags as above: the inner
 * and outer spheres are Dirichlet. (ntan, nrad) = (5, 2) gives 960 macro-cells. */
int hyg_mesh_create_spherical_shell( int ntan, int nrad, double rmin, double rmax, hyg_mesh* out );
/* Replaces MeshInfo::meshSymmetricCuboid( lowerLeftFront, upperRightBack, nx, ny, nz )
 * (mesh/MeshGenSymmetricCuboid.cpp:35-253): nx x ny x nz sub-cubes of 24 tetrahedra each (corners plus
 * face centres), point-symmetric about the centre, followed by the same storage setup and boundary flags
 * as above (the whole surface is Dirichlet). */
int hyg_mesh_create_symmetric_cuboid( const double lower[3], const double upper[3], int nx, int ny, int nz, hyg_mesh* out );
/* MeshInfo::meshCuboid(lowerLeftFront, upperRightBack, nx, ny, nz) (mesh/MeshGenCuboid.cpp:35-174), six tetrahedra
 * per grid cube, followed by the same storage setup and boundary flags as hyg_mesh_create_from_msh. */
int hyg_mesh_create_cuboid( const double lower[3], const double upper[3], int nx, int ny, int nz, hyg_mesh* out );
/* MeshInfo::meshCubedDomain(cubeCoordinates, cubeType) (mesh/MeshGenCubedDomain.cpp:37-383): unit cubes at the
 * integer positions cubes[3 i .. 3 i + 2] (i < ncubes), split into 6 (cube_type 0) or 24 (cube_type 1) tetrahedra,
 * followed by the same storage setup and boundary flags as hyg_mesh_create_from_msh. */
int hyg_mesh_create_cubed_domain( const int* cubes, int ncubes, int cube_type, hyg_mesh* out );
int hyg_mesh_destroy( hyg_mesh mesh );
/* out[0..6] = #vertices, #edges, #faces, #cells, #boundary vertices, #boundary edges, #boundary faces */
int hyg_mesh_info( hyg_mesh mesh, int64_t* out );
/* Number of vertex DoFs at `level` in the canonical order (numberOfGlobalDoFs<P1FunctionTag>). */
int hyg_mesh_num_dofs( hyg_mesh mesh, int level, int64_t* out );
/* xyz[3*n], boundary[n] (1 = on a Dirichlet primitive) of every DoF in canonical order:
 * macro-vertices, then edge interiors, face interiors, cell interiors, each primitive in
 * storage order and its points in the reference's own linearisation. Either pointer may be NULL. */
int hyg_mesh_dof_coords( hyg_mesh mesh, int level, double* xyz, uint8_t* boundary );

/* ---------------------------------------------------------------- functions
 * Replaces P1Function<ValueType>(name, storage, minLevel, maxLevel)
 * (p1functionspace/VertexDoFFunction.cpp:84-169; memory/FunctionMemory.hpp:46-240). */
int hyg_func_create( hyg_mesh mesh, int min_level, int max_level, int dtype, hyg_func* out );
/* Replaces P2Function<ValueType>(name, storage, minLevel, maxLevel) (p2functionspace/P2Function.hpp;
 * its VertexDoFFunction and EdgeDoFFunction, edgedofspace/EdgeDoFIndexing.hpp:919-1000). The vertex DoFs of
 * level L and the DoFs at the midpoints of the level-L micro-edges are exactly the vertices of level L+1,
 * so a P2 function of level L exchanges its values in the canonical order of level L+1
 * (hyg_mesh_num_dofs / hyg_mesh_dof_coords with level + 1). Every call below takes P2 levels for a P2
 * function; P1 and P2 handles cannot be mixed in one call. max_level <= 9. */
int hyg_func_create_p2( hyg_mesh mesh, int min_level, int max_level, int dtype, hyg_func* out );
int hyg_func_destroy( hyg_func f );
/* Host <-> device copy of one level in canonical order (double on the host side).
 * Upload also writes every copy of a value that is shared between macro-cells. */
int hyg_func_upload( hyg_func f, int level, const double* canonical, void* stream );
int hyg_func_download( hyg_func f, int level, double* canonical, void* stream );

/* VertexDoFFunction::interpolate(constant, level, flag)  (VertexDoFFunction.cpp:382-394) */
int hyg_interpolate_const( hyg_func f, double value, int level, int flag, void* stream );
/* VertexDoFFunction::assign({scalars}, {functions}, level, flag), 1 to 3 sources
 * (VertexDoFFunction.cpp:1107-1199) */
int hyg_assign( hyg_func dst, int n, const double* scalars, const hyg_func* srcs, int level, int flag, void* stream );
/* VertexDoFFunction::add({scalars}, {functions}, level, flag): dst += sum_k s_k f_k (VertexDoFFunction.cpp:1366-1444) */
int hyg_add( hyg_func dst, int n, const double* scalars, const hyg_func* srcs, int level, int flag, void* stream );
/* VertexDoFFunction::multElementwise({a, b}, level, flag) (VertexDoFFunction.cpp:1445-1523) */
int hyg_mult_elementwise( hyg_func dst, hyg_func a, hyg_func b, int level, int flag, void* stream );
/* VertexDoFFunction::dotGlobal(other, level, flag) (VertexDoFFunction.cpp:1668-1753); synchronises `stream`. */
int hyg_dot( hyg_func a, hyg_func b, int level, int flag, double* out, void* stream );
/* No counterpart in the reference. hyg_dot for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the same kernels and the same
 * result, written asynchronously to the caller's device address d_out (one double) on `stream`, with no
 * host synchronisation. The block partial sums go to a fixed scratch array of HYG_NOALLOC_DOT_PARTIALS
 * doubles (2 MiB) that every device holds for the life of the process (its address is resolved by
 * hyg_init); a level needing more partial sums is rejected with HYG_ERR_ARG. Every call on a device
 * shares that device's scratch: calls that overlap on one device (two streams, two threads) race. */
#define HYG_NOALLOC_DOT_PARTIALS 262144
int hyg_dot_noalloc( hyg_func a, hyg_func b, int level, int flag, double* d_out, void* stream );
/* P1Function<T>::copyFrom(P1Function<U>) between value types (VertexDoFFunction.hpp:511-530, FunctionMemory.hpp:117) */
int hyg_convert( hyg_func dst, hyg_func src, int level, void* stream );

/* ---------------------------------------------------------------- halos in the reference's layout
 * The library keeps macro-face, -edge and -vertex values as replicas in every incident cell array, so no halo
 * exchange precedes an operator. These calls produce and consume the reference's per-primitive memory of one level
 * instead (p1functionspace/VertexDoFMemory.hpp:44-64), for code that holds data in that layout or a multi-GPU
 * extension that ships halos. Primitives in storage order (vertices, edges, faces); per primitive its own values,
 * then its ghost layers as VertexDoFFunction::communicate fills them (VertexDoFPackInfo.hpp: edge -> vertex
 * :204-245, face -> edge with the face's cell ghost layers :295-408, cell -> face :508-660):
 *   vertex: own, then per incident edge the edge point next to the vertex;
 *   edge:   own w = N+1 values from its lower-id vertex, per incident face the N points one row into the face,
 *           per incident cell the N-1 points one step towards both other cell vertices;
 *   face:   own w(w+1)/2 values, per incident cell (cell 0: VERTEX_TC layer, cell 1: VERTEX_BC) the N(N+1)/2 points
 *           one layer into that cell.
 * hyg_halo_layout writes nV+nE+nF+1 element offsets (the last is the total). hyg_halo_pack gathers that memory
 * into `out` (device, the function's value type); hyg_halo_unpack writes the own values in `in` into all replicas
 * (the downward communication vertex -> edge -> face -> cell, VertexDoFPackInfo.hpp:161-505: a point takes the value
 * of the primitive it lies inside, so vertex values win over edge ends and edge values over face borders);
 * hyg_halo_sum_replicas replaces every replica of a point on a macro-face, -edge or -vertex whose DoF type is in
 * flag by the sum over its cells (communicateAdditively<Cell, Face/Edge/Vertex>, VertexDoFFunction.hpp:389-415,
 * VertexDoFAdditivePackInfo.hpp:378-571, followed by the downward copy). P1 functions. */
int hyg_halo_layout( hyg_mesh mesh, int level, int64_t* offsets );
int hyg_halo_pack( hyg_func f, int level, void* out, void* stream );
int hyg_halo_unpack( hyg_func f, int level, const void* in, void* stream );
int hyg_halo_sum_replicas( hyg_func f, int level, int flag, void* stream );

/* ---------------------------------------------------------------- operators
 * Replaces P1ConstantLaplaceOperator(storage, minLevel, maxLevel)
 * (p1functionspace/P1ConstantOperator.cpp:77, 602-655) incl.
 * computeInverseDiagonalOperatorValues() (P1Operator.hpp:465-469, 645-915). */
int hyg_op_create_p1_const( hyg_mesh mesh, int min_level, int max_level, hyg_op* out );
int hyg_op_destroy( hyg_op op );
/* Replaces P1ElementwiseOperator<forms::p1_div_k_grad_affine_q3>(storage, minLevel, maxLevel, form)
 * (elementwiseoperators/P1ElementwiseOperator.cpp:64-415; forms/form_hyteg_generated/p1/
 * p1_div_k_grad_affine_q3.cpp:179-381) for -div(k grad u) with the coefficient
 *   k(x) = k0 exp( sum_m a_m cos(2 pi kappa_m . x + phi_m) ),
 * modes[5*m .. 5*m+4] = kappa_x, kappa_y, kappa_z, phi, a (0 <= n_modes <= 64; n_modes = 0: k = k0).
 * k is evaluated at the 5 quadrature points of every element on the fly in each apply. The operator
 * supports hyg_apply, hyg_smooth_jac, hyg_residual, hyg_inverse_diagonal, hyg_power_iteration and GMG
 * in FP64 and FP32. */
int hyg_op_create_p1_divkgrad( hyg_mesh mesh, int min_level, int max_level, double k0, const double* modes, int n_modes,
                               hyg_op* out );
/* Replaces IcosahedralShellMap::setMap(setupStorage) (geometry/IcosahedralShellMap.hpp:40-128, 281-515) on a
 * spherical-shell mesh: every cell gets the shell map (two vertices on one ray: the lone vertex of the sphere with one
 * vertex is `ref`, its ray partner `ray`; cells with two vertices on each sphere take the map of the cell across their
 * first local face without both ray vertices), F(x) = x (r_ray + lambda(x) (r_ref - r_ray)) / |x| with lambda the
 * fraction of the way from the plane (ray, thr, for) to ref. hyg_mesh_dof_coords returns blended positions from
 * then on (edge and face points: the map of their first cell); operators created afterwards are unaffected except
 * hyg_op_create_p1_blending_laplace. HYG_ERR_ARG if a cell is not a shell cell. */
int hyg_mesh_set_shell_map( hyg_mesh mesh );
/* Replaces P1ElementwiseBlendingLaplaceOperator(storage, minLevel, maxLevel) (P1ElementwiseOperator<
 * forms::p1_diffusion_blending_q3 >, elementwiseoperators/P1ElementwiseOperator.hpp:332; form
 * p1_diffusion_blending_q3.cpp:287-...) on a mesh with the shell map: element matrices
 * a_ij = |det J| sum_q w_q |det DF_q| (DF_q^-T g_i).(DF_q^-T g_j) over the 5-point degree-3 rule. The 15 stencil weights
 * of every cell point are summed once at creation (15 FP64 + 15 FP32 values per point and cell); supports what
 * hyg_op_create_p1_divkgrad supports except surrogates and batches. */
int hyg_op_create_p1_blending_laplace( hyg_mesh mesh, int min_level, int max_level, hyg_op* out );
/* The same operator with a polynomial coefficient k(x) = sum_i coeffs[i] m_i(x) over the (degree+1)(degree+2)(degree+3)/6
 * monomials m_i = x^a y^b z^c of degree <= degree (0..12) in the order of the reference's MonomialBasis3D
 * (polynomial/MonomialBasis3D.hpp: by total degree, within a degree by decreasing a, then decreasing b), as the
 * forms::p1_div_k_grad_affine_q3(k, k) of tests/hyteg/operators/P1SurrogateOperatorTest.cpp:37-45 take it. */
int hyg_op_create_p1_divkgrad_poly( hyg_mesh mesh, int min_level, int max_level, int degree, const double* coeffs, hyg_op* out );
/* Replaces P1SurrogateOperator< p1_div_k_grad_affine_q3 >::interpolateStencils(polyDegree, interpolationLevel)
 * (p1functionspace/P1SurrogateOperator.hpp:97-110, 232-335) on a variable-coefficient P1 operator: from then on the
 * operator is the surrogate. On every level L the 15 stencil weights of the interior points of each macro-cell are
 * polynomials of degree `degree` (0..12) in the cell-local coordinates (x, y, z) / 2^L, least-squares fits
 * (LSQPInterpolator3D, column-pivoted QR) to the exact stencils at the interior points of level
 * min(L, interpolation_level) (>= 2) scaled to level L; face, edge and vertex points keep the exact stencils. Applies,
 * smoothers, residuals, the inverse diagonal (interior: 1 / the polynomial centre weight) and GMG use the surrogate.
 * Keeps the coefficient integrals (as hyg_op_store_local_matrices) for the face, edge and vertex points. Calling it
 * again refits. Surrogate operators cannot be batched. */
int hyg_op_interpolate_stencils( hyg_op op, int degree, int interpolation_level );
/* Replaces P2ConstantLaplaceOperator(storage, minLevel, maxLevel) (P2ConstantOperator< p2_tet_diffusion >,
 * p2functionspace/P2ConstantOperator.cpp:99-113, with its vertex/edge sub-operators and
 * computeInverseDiagonalOperatorValues), applied element by element with the element matrices of the six
 * micro-cell types of each macro-cell (the sums the reference's stencils hold). Works on P2 functions; with
 * hyg_restrict / hyg_prolongate on P2 functions (quadratic transfers) it supports hyg_apply, hyg_smooth_jac,
 * hyg_residual, hyg_inverse_diagonal, hyg_power_iteration, GMG and defect correction in FP64 and FP32. */
int hyg_op_create_p2_const( hyg_mesh mesh, int min_level, int max_level, hyg_op* out );
/* Replaces P2ElementwiseOperator< forms::p2_div_k_grad_affine_q4 >(storage, minLevel, maxLevel, form)
 * (elementwiseoperators/P2ElementwiseOperator.cpp:100-310; forms/form_hyteg_generated/p2/p2_div_k_grad_affine_q4.cpp,
 * Xiao-Gimbutas 11-point rule) for -div(k grad u) on P2 functions, with the coefficient
 *   kind 0: k(x) = k0 exp(alpha (1 - |x|)) (1 + beta xi_e), xi_e uniform in [-1, 1) per micro-element from
 *           SplitMix64(seed + (id + 1) 0x9E3779B97F4A7C15), id = (6 cell + type) S3(Nc + 1) + anchor index;
 *   kind 1: k(x) = 1 + x/2 + y^2/4.
 * k is evaluated at the 11 quadrature points of every element once, at creation; each apply assembles the
 * element rows on the fly. Supports what hyg_op_create_p2_const supports. */
int hyg_op_create_p2_divkgrad( hyg_mesh mesh, int min_level, int max_level, int kind, double k0, double alpha, double beta,
                               uint64_t seed, hyg_op* out );
/* Replaces computeAndStoreLocalElementMatrices of the elementwise operators (elementwiseoperators/
 * P2ElementwiseOperator.cpp:503-580, P1ElementwiseOperator.cpp:418). For the variable-coefficient P2
 * operator: assembles, on every level, the stencil of every cell-interior point and the one-sided stencil of
 * every cell-face point from the element matrices (about 29 values per point on average, FP64 plus an FP32 copy)
 * and applies those from then on; points on cell edges and vertices stay element by element. For the
 * variable-coefficient P1 operator: keeps the coefficient integral of every micro-element (the element matrix is
 * that integral times the cell's matrix of its type) instead of evaluating k in every apply. Trades HBM for the
 * per-apply work; results are unchanged up to rounding (P2) or bit for bit (P1). */
int hyg_op_store_local_matrices( hyg_op op );
/* Cell stencil (15 weights in the order C W E N S NW SE TC TW TS TSE BC BN BE BNW) of cell `cell`
 * (constant-coefficient operators only). */
int hyg_op_cell_stencil( hyg_op op, int level, int cell, double* out15 );
/* Operator::apply(src, dst, level, flag, updateType) (operators/Operator.hpp:73-80; P1Operator.hpp:196-324).
 * src and dst must differ and have the same value type. */
int hyg_apply( hyg_op op, hyg_func src, hyg_func dst, int level, int flag, int update, void* stream );
/* VectorLaplaceOperator::apply (operators/VectorLaplaceOperator.hpp:36-66): the scalar operator applied to
 * each of n <= 3 component functions (same value type, distinct sources and destinations). For the
 * variable-coefficient P2 operator with stored stencils the cell interiors of all components are done in one
 * pass that reads every stored weight once; without stored stencils and n = 3, each cell-interior and cell-face
 * point's stencil is assembled once for the three components; otherwise the components are applied one after the
 * other. Every component's result equals its single-component apply bit for bit. */
int hyg_apply_vector( hyg_op op, int n, const hyg_func* srcs, const hyg_func* dsts, int level, int flag, int update, void* stream );
/* WeightedJacobiSmoothable::smooth_jac(dst, rhs, src, relax, level, flag) (solvers/Smoothables.hpp:32-43;
 * P1Operator.hpp:433-451): dst = src + relax * D^-1 (rhs - A src), fused into one pass. */
int hyg_smooth_jac( hyg_op op, hyg_func dst, hyg_func rhs, hyg_func src, double relax, int level, int flag,
                    void* stream );
/* SORSmoothable::smooth_sor / SORBackwardsSmoothable::smooth_sor_backwards (solvers/Smoothables.hpp:72-98;
 * P1Operator.hpp:343-423): one SOR sweep on dst in place, x_p += relax D^-1 (rhs - A x)_p. The reference
 * visits vertices, edges, faces and cells in that order, each lexicographically; here the points are visited
 * by colour (multi-colour Gauss-Seidel): the macro-vertices get a greedy colouring (vertex order, smallest
 * colour free among macro-edge neighbours; hyg_mesh_sor_colours), a lattice point with barycentric lattice
 * coordinates b_i over the vertices v_i of its cell gets the colour (sum_i c(v_i) b_i) mod ncolours, and
 * points of one colour, which the stencil never couples, are updated together; backwards = 1 takes the
 * colours in reverse order. Constant P1 operator only. */
int hyg_smooth_sor( hyg_op op, hyg_func dst, hyg_func rhs, double relax, int level, int flag, int backwards, void* stream );
/* SORSmoothable::smooth_sor / smooth_sor_backwards in the reference's own order (P1Operator.hpp:343-416): macro-
 * vertices, edges, faces, cells (backwards: cells, faces, edges, vertices), each primitive lexicographically
 * (VertexDoFMacroFace.hpp:751-810, generatedKernels/sor_3D_macrocell_P1.cpp), with the reference's ghost-layer
 * semantics: a point of primitive P reads a neighbour's updated value when the neighbour lies in P's closure or in a
 * primitive of higher dimension, and its value before the sweep otherwise. x_p = (1 - relax) x_p + relax (rhs_p -
 * sum_{q != p} a_pq x_q) / a_pp. Same results as the reference up to rounding; parallel over primitives, face
 * anti-diagonals and cell hyperplanes x + 2y + 3z (one launch per hyperplane, about 3 * 2^level launches), so it is
 * much slower than hyg_smooth_sor. Allocates a snapshot of the level on first use. Constant P1 operator only. */
int hyg_smooth_sor_ordered( hyg_op op, hyg_func dst, hyg_func rhs, double relax, int level, int flag, int backwards, void* stream );
/* The SOR colouring of a mesh: the number of colours and, if vertex_colours is not null, the colour of every
 * macro-vertex (nV entries). No counterpart in the reference (its SOR is lexicographic). */
int hyg_mesh_sor_colours( hyg_mesh mesh, int* ncolours, int* vertex_colours );
/* OperatorWithInverseDiagonal::getInverseDiagonalValues (Smoothables.hpp:99-106) written into `dst`. */
int hyg_inverse_diagonal( hyg_op op, hyg_func dst, int level, void* stream );
/* residual r = b - A x on `flag` (GeometricMultigridSolver.hpp:185-186 apply + assign), fused. */
int hyg_residual( hyg_op op, hyg_func r, hyg_func b, hyg_func x, int level, int flag, void* stream );

/* Stencil kernel configurations (no counterpart in the reference, whose generated kernels have one
 * fixed loop nest). The cell-interior kernel marches z-planes with R rows per wavefront and D planes of
 * loads in flight, one point per lane; configuration k has R = {1,2,4,1,2,4,8}[k], D = {2,2,2,4,4,4,2}[k]
 * (hyg_op_get_config reports R as `rows` and 1 as `points`). Modes: 0 apply,
 * 1 apply-add, 2 residual, 3 Jacobi sweep, 4 first Chebyshev step, 5 later Chebyshev steps.
 * For the P2 constant operator the configurations 0..7 are those of its cell-interior row kernel: z-segments of
 * {8, 16, 32, 64}[k / 2] planes with k % 2 + 1 planes of loads in flight (hyg_op_get_config reports the segment
 * length as `rows` and the planes in flight as `points`); default 4 at P2 level >= 7, 2 at level 6, 0 below.
 * hyg_op_tune times every configuration for every mode at `level` on the operator's mesh (`reps`
 * sweeps each, on scratch functions it allocates and frees) and keeps the fastest; chosen[0..5]
 * (may be NULL) receives the choices. The choice is shared by all operators of the mesh. */
int hyg_op_tune( hyg_op op, int level, int dtype, int reps, int* chosen, void* stream );
int hyg_op_set_config( hyg_op op, int level, int dtype, int mode, int config );
int hyg_op_get_config( hyg_op op, int level, int dtype, int mode, int* config, int* rows, int* points );

/* ---------------------------------------------------------------- grid transfer
 * P1toP1LinearRestriction::restrict(function, sourceLevel, flag) (gridtransferoperators/P1toP1LinearRestriction.cpp:168-345);
 * on a P2 function P2toP2QuadraticRestriction::restrict (P2toP2QuadraticRestriction.cpp:131-286) */
int hyg_restrict( hyg_func f, int source_level, int flag, void* stream );
/* P1toP1LinearProlongation::prolongate / prolongateAndAdd (gridtransferoperators/P1toP1LinearProlongation.cpp:194-398);
 * on a P2 function P2toP2QuadraticProlongation::prolongate / prolongateAndAdd (P2toP2QuadraticProlongation.cpp:217-424) */
int hyg_prolongate( hyg_func f, int source_level, int flag, int update, void* stream );

/* ---------------------------------------------------------------- solvers
 * ChebyshevSmoother::setupCoefficients(order, lower, upper) (solvers/ChebyshevSmoother.hpp:128-190), order 1..5. */
int hyg_cheb_coefficients( int order, double lower, double upper, double* out );
/* chebyshev::estimateRadius -> estimateSpectralRadiusWithPowerIteration on D^-1 A, flag All
 * (ChebyshevSmoother.hpp:255-265, numerictools/SpectrumEstimation.hpp:57-80). x holds the start vector. */
int hyg_power_iteration( hyg_op op, hyg_func x, hyg_func tmp, int iterations, int level, double* radius, void* stream );

/* GeometricMultigridSolver (solvers/GeometricMultigridSolver.hpp:42-235) with a WeightedJacobiSmoother
 * (smoother = 0, solvers/WeightedJacobiSmoother.hpp:42-62), ChebyshevSmoother (smoother = 1,
 * ChebyshevSmoother.hpp:48-97), SORSmoother (smoother = 2, solvers/SORSmoother.hpp:38-50, forward multi-colour
 * sweeps as hyg_smooth_sor) or SymmetricSORSmoother (smoother = 3, solvers/SymmetricSORSmoother.hpp:37-60, a
 * forward then a backward sweep); smoother = 4 and 5 are the same two in the reference's own visiting order
 * (hyg_smooth_sor_ordered, same iterates as the reference up to rounding); SOR for the constant P1 operator only,
 * and a CGSolver coarse solver
 * (solvers/CGSolver.hpp:66-134).
 * iparams = {min_level, max_level, pre, post, increment, smoother, cheb_order, cg_max_iter}
 * dparams = {relax, cheb_lower, cheb_upper, cg_tolerance}
 * The solver works on functions of the operator's value type (HYG_F64 or HYG_F32). */
int hyg_gmg_create( hyg_op op, int dtype, const int* iparams, const double* dparams, hyg_gmg* out );
int hyg_gmg_destroy( hyg_gmg g );
/* Solver::solve(A, x, b, level) once = one V-cycle (GeometricMultigridSolver.hpp:127-134). */
int hyg_gmg_vcycle( hyg_gmg g, hyg_func x, hyg_func b, int level, void* stream );
/* One V-cycle as a replayed hipGraph: the first call for (x, b, level, cg_iterations) captures the whole
 * cycle on `stream` (which must not be the default stream) and instantiates it; later calls only launch
 * it. The coarse CG runs `cg_iterations` iterations whose step lengths stay in device memory (no
 * host round trip); once its residual norm is below the solver's CG tolerance the remaining iterations
 * take no step, so with enough iterations the result equals hyg_gmg_vcycle. No counterpart in the
 * reference (its V-cycle is host-driven). */
int hyg_gmg_vcycle_graph( hyg_gmg g, hyg_func x, hyg_func b, int level, int cg_iterations, void* stream );
/* `cycles` V-cycles; res[0] = ||b - Ax|| over Inner|Neumann DoFs before, res[k] after cycle k (FP64 norms). */
int hyg_gmg_solve( hyg_gmg g, hyg_func x, hyg_func b, int level, int cycles, double* res, void* stream );
/* V-cycles until ||b - Ax|| <= rel_tol * ||b - Ax0|| or `max_cycles` cycles (the convergence loop around
 * Solver::solve in the reference's drivers, e.g. apps/MultigridStudies/MultigridStudies.cpp:575-660).
 * res[0..*cycles] as in hyg_gmg_solve; res must hold max_cycles + 1 values. */
int hyg_gmg_solve_tol( hyg_gmg g, hyg_func x, hyg_func b, int level, int max_cycles, double rel_tol, double* res, int* cycles,
                       void* stream );
/* Mixed-precision defect correction (tests/hyteg/mixedPrecision/mixedPrecisionIterativeRefinement.cpp:162-172):
 * FP64 x, b; each outer step r = b - A x (FP64), FP32 correction by `inner` FP32 V-cycles of g32, x += c (FP64).
 * Stops after `max_outer` steps or when ||r||/||r0|| <= rel_tol. res[k] = FP64 residual norm after step k
 * (res[0] initial); *steps = number of outer steps taken. */
int hyg_ir_solve( hyg_op op64, hyg_gmg g32, hyg_func x, hyg_func b, int level, int inner, int max_outer,
                  double rel_tol, double* res, int* steps, void* stream );

/* ---------------------------------------------------------------- batched independent problems
 * No counterpart in the reference, which runs one GeometricMultigridSolver::solve (GeometricMultigridSolver.hpp:
 * 136-235) per problem and one Operator::apply (p1functionspace/P1Operator.hpp:196-324) per call. A batch holds n
 * independent problems: problem i is solved by solvers[i] (its operator, mesh, levels and smoother scalars) on
 * x[i], b[i] at the solver's max level. Every grid operation of a cycle is one persistent launch over all problems
 * (a stencil operation is two: the sweep and the face/edge/vertex combine), each problem at its own level; per
 * problem the operations, their order and their arithmetic are those of hyg_gmg_vcycle. Requirements: P1 constant
 * or elementwise operators (the latter with hyg_op_store_local_matrices applied), the weighted Jacobi or the
 * Chebyshev smoother, the same value type, smoother order and pre/post/increment in every solver, and coarse
 * levels without inner DoFs (the batched cycle has no coarse solve). The handles must outlive the batch. */
int hyg_batch_create( int n, const hyg_gmg* solvers, const hyg_func* x, const hyg_func* b, hyg_batch* out );
int hyg_batch_destroy( hyg_batch bt );
/* Cell-interior kernel configuration of the constant-operator sweeps: -1 (the default) takes, for every batched
 * sweep, the configuration chosen for its largest problem's level, value type and mode (hyg_op_tune, see
 * hyg_op_set_config); 0..6 forces that configuration of hyg_op_set_config for every batched sweep. */
int hyg_batch_set_config( hyg_batch bt, int config );
/* One V-cycle of every problem; use_graph = 1 captures the batched cycle into a hipGraph on the first call (on
 * `stream`, which must not be the default stream) and replays it afterwards. */
int hyg_batch_vcycle( hyg_batch bt, int use_graph, void* stream );
/* dst[i] = A_i src[i] at problem i's top level on `flag` (Operator::apply, Replace), all problems in one launch. */
int hyg_batch_apply( hyg_batch bt, const hyg_func* src, const hyg_func* dst, int flag, void* stream );
/* out[i] = ||b_i - A_i x_i|| over Inner|Neumann DoFs at the top level (batched residual and dot). */
int hyg_batch_residual_norms( hyg_batch bt, double* out, void* stream );
/* Mixed-precision defect correction of every problem (as hyg_ir_solve): FP64 x64[i], b64[i]; each outer step
 * r = b - A x in FP64, `inner` batched FP32 V-cycles of this (FP32) batch on A c = r, x += c. Outer steps run until
 * every problem has ||r||/||r0|| <= rel_tol or after max_outer steps; use_graph = 1 replays the whole outer step
 * as one hipGraph. res[i * (max_outer + 1) + k] = FP64 residual norm of problem i after step k. */
int hyg_batch_ir_solve( hyg_batch bt, const hyg_func* x64, const hyg_func* b64, int inner, int max_outer, double rel_tol,
                        int use_graph, double* res, int* steps, void* stream );

#ifdef __cplusplus
}
#endif

#endif
