// gfx950 kernels of the variable-coefficient P1 operator -div(k grad u), assembled element by element
// on the fly (the reference's P1ElementwiseOperator<p1_div_k_grad_affine_q3>,
// elementwiseoperators/P1ElementwiseOperator.cpp:64-301, forms/form_hyteg_generated/p1/
// p1_div_k_grad_affine_q3.cpp:179-381).
//
// On an affine micro-tetrahedron the gradients are constant, so the element matrix is
// kb_e * K_t: K_t is the element matrix of k = 1 for the micro-cell type t of the macro-cell (held in
// ElemCell::mt) and kb_e = 6 sum_q w_q k(x_q) is the 5-point degree-3 quadrature of k over the
// element divided by its volume (w = -2/15 at the centroid, 3/40 at the four points with barycentric
// coordinates (1/2, 1/6, 1/6, 1/6)). Each apply
//   1. evaluates kb_e for every micro-cell (kKBar): per lattice point the phases of the Fourier modes of
//      k are formed once and the 30 quadrature points of the six cell types are reached by angle
//      addition with per-cell tables of cos/sin of the quadrature offsets (kKBarTables, once per level at
//      operator creation; read with scalar loads);
//   2. assembles at every point of every macro-cell the 15 weights of its stencil from the (up to) 24
//      incident elements of that cell and applies them (kElemApply); interior points get the epilogue,
//      points on faces, edges and vertices store the cell's one-sided sum;
//   3. adds the one-sided sums of face, edge and vertex points over their cells and writes every
//      replica (kFaceCombine, kLowCombine).
// The inverse diagonal (per point, since it depends on k) is built the same way at operator setup.

#include "elementwise.hpp"

namespace hyg {

// ------------------------------------------------------------------ coefficient tables
// Tables of one level, built once (one block per cell): per mode the phase at the cell's vertex 0, its steps
// along the three lattice axes and the amplitude (phase[(c nModes + m) 5 + ..]), and cos/sin of the phase offsets of
// the 30 quadrature points of the six types from their anchor (tab[(c nModes + m) 60 + 2 (5 t + q) + {0, 1}]).
__global__ __launch_bounds__( kBlock ) void kKBarTables( const ElemCell* __restrict__ cells, const double* __restrict__ modes, int nModes,
                                                         double* __restrict__ phase, double* __restrict__ tab, float* __restrict__ tabF )
{
   const int       c  = int( blockIdx.x );
   const ElemCell& ec = cells[c];
   double*         mp = phase + i64( c ) * nModes * 5;
   for ( int m = int( threadIdx.x ); m < nModes; m += kBlock )
   {
      const double kx = modes[5 * m], ky = modes[5 * m + 1], kz = modes[5 * m + 2];
      const double tp = 2.0 * M_PI;
      mp[5 * m]       = tp * ( kx * ec.v0[0] + ky * ec.v0[1] + kz * ec.v0[2] ) + modes[5 * m + 3];
      for ( int k = 0; k < 3; ++k )
         mp[5 * m + 1 + k] = tp * ( kx * ec.a[0][k] + ky * ec.a[1][k] + kz * ec.a[2][k] );
      mp[5 * m + 4] = modes[5 * m + 4];
   }
   __syncthreads();
   for ( int i = int( threadIdx.x ); i < 30 * nModes; i += kBlock )
   {
      const int m = i / 30, tq = i - 30 * m, t = tq / 5, q = tq - 5 * t;
      // lattice offset of quadrature point q of type t from the anchor
      double d[3] = { 0.0, 0.0, 0.0 };
      for ( int v = 0; v < 4; ++v )
      {
         const double lam = q == 0 ? 0.25 : ( v == q - 1 ? 0.5 : 1.0 / 6.0 );
         for ( int k = 0; k < 3; ++k )
            d[k] += lam * kMC[t][v][k];
      }
      const double ang = mp[5 * m + 1] * d[0] + mp[5 * m + 2] * d[1] + mp[5 * m + 3] * d[2];
      double       sn, cs;
      sincos( ang, &sn, &cs );
      const i64 o  = ( i64( c ) * nModes ) * 60 + 2 * i;
      tab[o]       = cs;
      tab[o + 1]   = sn;
      tabF[o]      = float( cs );
      tabF[o + 1]  = float( sn );
   }
}

// ------------------------------------------------------------------ launch
template < typename T >
static void launchKBar( const hyg_op_s& op, int L, int nC, T* kb, hipStream_t s )
{
   const int N  = 1 << L;
   const int tK = ceilDiv( S3i( N ), kBlock );
   if ( tK > 0 && S3i( N ) > 0 && op.kPolyDeg >= 0 )
      hipLaunchKernelGGL( ( kKBarPoly< T > ), dim3( nC * tK ), dim3( kBlock ), 0, s, L, tK, op.levels[L].elem, op.kPoly, op.kPolyDeg, kb );
   else if ( tK > 0 && S3i( N ) > 0 )
   {
      const hyg_op_s::Level& ol = op.levels[L];
      const T*               tb = std::is_same< T, double >::value ? (const T*) ol.kbTab : (const T*) ol.kbTabF;
      hipLaunchKernelGGL( ( kKBar< T > ), dim3( nC * tK ), dim3( kBlock ), 0, s, L, tK, (const double*) ol.kbPhase, tb, op.nModes, op.k0,
                          kb );
   }
}

// per-level tables of the Fourier coefficient (phases and quadrature-offset cos/sin), once at operator setup
void launchKBarTables( const MeshCtx& m, const hyg_op_s& op, int L, hipStream_t s )
{
   const hyg_op_s::Level& ol = op.levels[L];
   if ( op.nModes > 0 )
      hipLaunchKernelGGL( kKBarTables, dim3( m.mesh->nC() ), dim3( kBlock ), 0, s, ol.elem, op.modes, op.nModes, ol.kbPhase, ol.kbTab,
                          ol.kbTabF );
}

// coefficient integrals of every level in both value types, kept for the applies that follow
void launchElementwiseStore( const MeshCtx& m, hyg_op_s& op, hipStream_t s )
{
   for ( int L = op.minL; L <= op.maxL; ++L )
   {
      launchKBar< double >( op, L, m.mesh->nC(), op.levels[L].kb, s );
      launchKBar< float >( op, L, m.mesh->nC(), op.levels[L].kbF, s );
   }
   op.stored = true;
}

template < typename T, int MODE, bool DIAG >
static void elemT( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, const StencilCall& c, hipStream_t s )
{
   constexpr bool             dbl = std::is_same< T, double >::value;
   const hyg_op_s::Level&     ol  = op.levels[L];
   const int                  nC  = m.mesh->nC();
   const int                  N   = 1 << L, w = N + 1;
   T*                         kb  = dbl ? (T*) ol.kb : (T*) ol.kbF;
   T*                         part = dbl ? (T*) ol.facePart : (T*) ol.facePartF;
   T*                         low  = dbl ? (T*) ol.lowPart : (T*) ol.lowPartF;
   const T*                   dv   = DIAG ? nullptr : ( dbl ? (const T*) ol.dinv : (const T*) ol.dinvF );
   const T*                   x    = (const T*) c.x;
   const T*                   in1  = (const T*) c.in1;
   T*                         out0 = (T*) c.out0;
   T*                         out1 = (T*) c.out1;
   // 1. coefficient integrals of all micro-cells (kept from hyg_op_store_local_matrices when stored)
   if ( !op.stored )
      launchKBar< T >( op, L, nC, kb, s );
   // 2. per-cell stencils (interior results, one-sided sums of face, edge and vertex points); the surrogate operator
   //    does its interior points with the polynomial stencils
   const int tA = ceilDiv( S3i( w ), kBlock );
   if ( !DIAG && surrogateActive( op, L ) )
   {
      const int tB = ceilDiv( ol.surNBnd, kBlock );
      hipLaunchKernelGGL( ( kElemApply< T, MODE, DIAG, true > ), dim3( nC * tB ), dim3( kBlock ), 0, s, L, tB, g.stride, ol.elem, kb,
                          m.cellInfo, m.cellFaces, m.prims, m.entries, part, low, dv, c.flag, x, in1, out0, out1, T( c.a ), ol.surBnd,
                          ol.surNBnd );
      launchSurrogateInterior( m, g, op, L, std::is_same< T, double >::value ? 8 : 4, c, s );
   }
   else if ( op.blend )
      hipLaunchKernelGGL( ( kElemApply< T, MODE, DIAG, false, true > ), dim3( xcdGrid( nC * tA ) ), dim3( kBlock ), 0, s, L, tA, g.stride,
                          ol.elem, kb, m.cellInfo, m.cellFaces, m.prims, m.entries, part, low, dv, c.flag, x, in1, out0, out1, T( c.a ),
                          nullptr, 0, dbl ? (const T*) ol.bw : (const T*) ol.bwF, nC * tA );
   else
      hipLaunchKernelGGL( ( kElemApply< T, MODE, DIAG > ), dim3( xcdGrid( nC * tA ) ), dim3( kBlock ), 0, s, L, tA, g.stride, ol.elem, kb,
                          m.cellInfo, m.cellFaces, m.prims, m.entries, part, low, dv, c.flag, x, in1, out0, out1, T( c.a ), nullptr, 0,
                          nullptr, nC * tA );
   // 3. face, edge and vertex values
   const int nLow = g.firstFaceChunk, nFace = g.nChunks - g.firstFaceChunk;
   if ( nFace > 0 )
      hipLaunchKernelGGL( ( kFaceCombine< T, MODE > ), dim3( xcdGrid( ceilDiv( nFace, kBlock / 64 ) ) ), dim3( kBlock ), 0, s,
                          g.chunks + g.firstFaceChunk, nFace, m.prims, m.entries, m.mesh->nV() + m.mesh->nE(), part, (const T*) nullptr,
                          dv, L, g.stride, c.flag, x, in1, out0, out1, T( c.a ) );
   if ( nLow > 0 )
      hipLaunchKernelGGL( ( kLowCombine< T, MODE > ), dim3( ceilDiv( nLow, kBlock / 64 ) ), dim3( kBlock ), 0, s, g.chunks, nLow,
                          m.prims, m.entries, low, dv, L, g.stride, c.flag, x, in1, out0, out1, T( c.a ) );
}

template < typename T >
static void elemMode( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, const StencilCall& c, hipStream_t s )
{
   switch ( c.mode )
   {
   case MODE_APPLY:
      elemT< T, MODE_APPLY, false >( m, g, op, L, c, s );
      break;
   case MODE_APPLY_ADD:
      elemT< T, MODE_APPLY_ADD, false >( m, g, op, L, c, s );
      break;
   case MODE_RESIDUAL:
      elemT< T, MODE_RESIDUAL, false >( m, g, op, L, c, s );
      break;
   case MODE_JACOBI:
      elemT< T, MODE_JACOBI, false >( m, g, op, L, c, s );
      break;
   case MODE_CHEB0:
      elemT< T, MODE_CHEB0, false >( m, g, op, L, c, s );
      break;
   default:
      elemT< T, MODE_CHEBK, false >( m, g, op, L, c, s );
      break;
   }
}

void launchElementwise( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, int dtype, const StencilCall& c, hipStream_t s )
{
   if ( dtype == 8 )
      elemMode< double >( m, g, op, L, c, s );
   else
      elemMode< float >( m, g, op, L, c, s );
}

void launchElementwiseInverseDiagonal( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, hipStream_t s )
{
   // the coefficient tables of the level (called once per level when the operator is created)
   launchKBarTables( m, op, L, s );
   const hyg_op_s::Level& ol = op.levels[L];
   const i64              n  = i64( m.mesh->nC() ) * g.stride + kTailSlack;
   (void) hipMemsetAsync( ol.dinv, 0, size_t( n ) * sizeof( double ), s );
   StencilCall c;
   c.mode = MODE_APPLY;
   c.flag = 15;
   c.out0 = ol.dinv;
   elemT< double, MODE_APPLY, true >( m, g, op, L, c, s );
   hipLaunchKernelGGL( ( kInvertAll< double > ), dim3( 1024 ), dim3( kBlock ), 0, s, n, ol.dinv );
   if ( surrogateActive( op, L ) )
      launchSurrogateInverseDiagonal( m, g, op, L, s );
   launchConvert( n, 4, ol.dinvF, 8, ol.dinv, s );
}

} // namespace hyg
