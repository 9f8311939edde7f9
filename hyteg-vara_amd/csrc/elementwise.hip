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
//      addition with per-cell tables of cos/sin of the quadrature offsets;
//   2. assembles at every point of every macro-cell the 15 weights of its stencil from the (up to) 24
//      incident elements of that cell and applies them (kElemApply); interior points get the epilogue,
//      points on faces, edges and vertices store the cell's one-sided sum;
//   3. adds the one-sided sums of face, edge and vertex points over their cells and writes every
//      replica (kFaceCombine, kLowCombine).
// The inverse diagonal (per point, since it depends on k) is built the same way at operator setup.
#include "face_combine.hpp"

namespace hyg {

// stencil direction (C W E N S NW SE TC TW TS TSE BC BN BE BNW) from vertex k to vertex j of type t
constexpr int kMCDir[6][4][4] = { { { 0, 2, 3, 7 }, { 1, 0, 5, 8 }, { 4, 6, 0, 9 }, { 11, 13, 12, 0 } },
                                  { { 0, 3, 5, 7 }, { 4, 0, 1, 9 }, { 6, 2, 0, 10 }, { 11, 12, 14, 0 } },
                                  { { 0, 5, 7, 8 }, { 6, 0, 10, 9 }, { 11, 14, 0, 1 }, { 13, 12, 2, 0 } },
                                  { { 0, 7, 8, 9 }, { 11, 0, 1, 4 }, { 13, 2, 0, 6 }, { 12, 3, 5, 0 } },
                                  { { 0, 5, 1, 14 }, { 6, 0, 4, 11 }, { 2, 3, 0, 12 }, { 10, 7, 9, 0 } },
                                  { { 0, 2, 10, 7 }, { 1, 0, 9, 8 }, { 14, 12, 0, 5 }, { 11, 13, 6, 0 } } };
// offset of vertex k of type t from the anchor, encoded x + 2y + 4z
constexpr int kMCOff[6][4] = { { 0, 1, 2, 4 }, { 1, 3, 2, 5 }, { 1, 2, 5, 4 }, { 3, 7, 6, 5 }, { 5, 6, 4, 2 }, { 2, 3, 5, 6 } };
// 6 x the quadrature weights (centroid first), so that kb = 1 for k = 1
constexpr double kQW6[5] = { -0.8, 0.45, 0.45, 0.45, 0.45 };

__device__ __forceinline__ void sincosT( double a, double& s, double& c ) { sincos( a, &s, &c ); }
__device__ __forceinline__ void sincosT( float a, float& s, float& c ) { sincosf( a, &s, &c ); }
__device__ __forceinline__ double expT( double a ) { return exp( a ); }
__device__ __forceinline__ float  expT( float a ) { return expf( a ); }

// ------------------------------------------------------------------ 1. coefficient integrals
// One thread per anchor (x, y, z), x+y+z <= N-1, of one macro-cell; kb[c][t][anchor] for the six types.
// Dynamic LDS: nModes x 30 x (cos, sin) of the quadrature offsets, then nModes x (theta0, gx, gy, gz, a).
template < typename T >
__global__ __launch_bounds__( kBlock ) void kKBar( int L, int tilesPerCell, const ElemCell* __restrict__ cells,
                                                   const double* __restrict__ modes, int nModes, double k0, T* __restrict__ kb )
{
   extern __shared__ double ldsD[];
   const int c    = int( blockIdx.x ) / tilesPerCell;
   const int tile = int( blockIdx.x ) - c * tilesPerCell;
   const int N    = 1 << L;
   T*        tab  = reinterpret_cast< T* >( ldsD + 5 * nModes );
   double*   mp   = ldsD;
   const ElemCell& ec = cells[c];
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
      tab[2 * i]     = T( cs );
      tab[2 * i + 1] = T( sn );
   }
   __syncthreads();
   const int S = S3i( N );
   const int s = tile * kBlock + int( threadIdx.x );
   if ( s >= S )
      return;
   int x, y, z;
   tetXYZ( N, s, x, y, z );
   T g[30];
#pragma unroll
   for ( int k = 0; k < 30; ++k )
      g[k] = T( 0 );
   for ( int m = 0; m < nModes; ++m )
   {
      const T th = T( mp[5 * m] + mp[5 * m + 1] * x + mp[5 * m + 2] * y + mp[5 * m + 3] * z );
      T       sn, cs;
      sincosT( th, sn, cs );
      const T  am = T( mp[5 * m + 4] );
      const T  ac = am * cs, as = am * sn;
      const T* tb = tab + 60 * m;
#pragma unroll
      for ( int k = 0; k < 30; ++k )
         g[k] += ac * tb[2 * k] - as * tb[2 * k + 1];
   }
   T* out = kb + i64( c ) * 6 * S;
#pragma unroll
   for ( int t = 0; t < 6; ++t )
   {
      T v = T( 0 );
#pragma unroll
      for ( int q = 0; q < 5; ++q )
         v += T( kQW6[q] ) * expT( g[5 * t + q] );
      if ( x + y + z + kMCMax[t] <= N )
         out[t * S + s] = T( k0 ) * v;
   }
}

// ------------------------------------------------------------------ 2. stencil application per cell
// One thread per point of the closed macro-cell. DIAG: the diagonal entry instead of the product.
template < typename T, int MODE, bool DIAG >
__global__ __launch_bounds__( kBlock ) void kElemApply( int L, int tilesPerCell, i64 stride, const ElemCell* __restrict__ cells,
                                                        const T* __restrict__ kb, const CellInfo* __restrict__ cinfo,
                                                        const CellFaceSlot* __restrict__ cfs, const LowerPrim* __restrict__ prims,
                                                        const LowerEntry* __restrict__ entries, T* __restrict__ part,
                                                        T* __restrict__ lowPart, const T* __restrict__ dinv, int flag,
                                                        const T* __restrict__ x, const T* __restrict__ in1,
                                                        T* __restrict__ out0, T* __restrict__ out1, T a )
{
   const int c    = int( blockIdx.x ) / tilesPerCell;
   const int tile = int( blockIdx.x ) - c * tilesPerCell;
   const int N = 1 << L, w = N + 1, NP = N + 1;
   const int s = tile * kBlock + int( threadIdx.x );
   if ( s >= S3i( w ) )
      return;
   int px, py, pz;
   tetXYZ( w, s, px, py, pz );
   const int slot = classifySlot( N, px, py, pz );
   if ( slot == 0 && !( flag & 1 ) )
      return;
   const ElemCell& ec = cells[c];
   const int       SN = S3i( N );
   const T*        kc = kb + i64( c ) * 6 * SN;
   // anchors p - o for the eight offsets o in {0,1}^3 (encoded x + 2y + 4z)
   int  aIdx[8];
   bool aOk[8];
#pragma unroll
   for ( int o = 0; o < 8; ++o )
   {
      const int ax = px - ( o & 1 ), ay = py - ( ( o >> 1 ) & 1 ), az = pz - ( ( o >> 2 ) & 1 );
      aOk[o]       = ax >= 0 && ay >= 0 && az >= 0;
      aIdx[o]      = aOk[o] ? cidx( N, ax, ay, az ) : 0;
   }
   T W[NDIR];
#pragma unroll
   for ( int d = 0; d < NDIR; ++d )
      W[d] = T( 0 );
#pragma unroll
   for ( int t = 0; t < 6; ++t )
#pragma unroll
      for ( int k = 0; k < 4; ++k )
      {
         const int o   = kMCOff[t][k];
         const int as  = px + py + pz - ( ( o & 1 ) + ( ( o >> 1 ) & 1 ) + ( ( o >> 2 ) & 1 ) );
         const bool ok = aOk[o] && as + kMCMax[t] <= N;
         const T    kv = ok ? kc[t * SN + aIdx[o]] : T( 0 );
         if constexpr ( DIAG )
            W[0] += kv * T( ec.mt[t][5 * k] );
         else
#pragma unroll
            for ( int j = 0; j < 4; ++j )
               W[kMCDir[t][k][j]] += kv * T( ec.mt[t][4 * k + j] );
      }
   const i64 q = i64( c ) * stride + cidx( w, px, py, pz );
   T         val;
   if constexpr ( DIAG )
      val = W[0];
   else
   {
      const int Wz = w - pz;
      const int tc = T2i( Wz ) - py, bc = T2i( Wz + 1 ) - py, nn = Wz - py;
      const int off[NDIR] = { 0, -1, 1, nn, -( nn + 1 ), nn - 1, -nn, tc, tc - 1, tc - nn, tc - nn + 1, -bc, -bc + nn + 1, -bc + 1, -bc + nn };
      val = T( 0 );
#pragma unroll
      for ( int d = 0; d < NDIR; ++d )
         val += W[d] * x[q + ( W[d] != T( 0 ) ? off[d] : 0 )];
   }
   if ( slot == 0 )
   {
      if constexpr ( DIAG )
         out0[q] = val;
      else
      {
         T i1 = T( 0 ), r0, r1;
         if constexpr ( MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
            i1 = in1[q];
         if constexpr ( MODE == MODE_APPLY_ADD )
            i1 = out0[q];
         epilogue< T, MODE >( val, x[q], i1, i1, dinv[q], a, r0, r1 );
         out0[q] = r0;
         if constexpr ( MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
            out1[q] = r1;
      }
   }
   else if ( slot <= 4 )
   {
      const CellFaceSlot fs = cfs[4 * c + slot - 1];
      if ( fs.bc & flag )
         part[i64( fs.block ) * NP * NP + facePartIndex( slot - 1, N, px, py, pz )] = val;
   }
   else
   {
      const CellInfo& ci = cinfo[c];
      if ( prims[ci.prim[slot]].bc & flag )
      {
         const int e = ci.entry[slot];
         int       i = 0;
         if ( slot <= 10 )
         {
            const int b[4] = { N - px - py - pz, px, py, pz };
            i              = b[entries[e].l1];
         }
         lowPart[i64( e ) * NP + i] = val;
      }
   }
}

// ------------------------------------------------------------------ launch
template < typename T >
static void launchKBar( const hyg_op_s& op, int L, int nC, T* kb, hipStream_t s )
{
   const int N  = 1 << L;
   const int tK = ceilDiv( S3i( N ), kBlock );
   if ( tK > 0 && S3i( N ) > 0 )
   {
      const size_t lds = size_t( op.nModes ) * ( 5 * sizeof( double ) + 60 * sizeof( T ) );
      hipLaunchKernelGGL( ( kKBar< T > ), dim3( nC * tK ), dim3( kBlock ), lds, s, L, tK, op.levels[L].elem, op.modes, op.nModes, op.k0,
                          kb );
   }
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
   // 2. per-cell stencils (interior results, one-sided sums of face, edge and vertex points)
   const int tA = ceilDiv( S3i( w ), kBlock );
   hipLaunchKernelGGL( ( kElemApply< T, MODE, DIAG > ), dim3( nC * tA ), dim3( kBlock ), 0, s, L, tA, g.stride, ol.elem, kb,
                       m.cellInfo, m.cellFaces, m.prims, m.entries, part, low, dv, c.flag, x, in1, out0, out1, T( c.a ) );
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
   const hyg_op_s::Level& ol = op.levels[L];
   const i64              n  = i64( m.mesh->nC() ) * g.stride + kTailSlack;
   (void) hipMemsetAsync( ol.dinv, 0, size_t( n ) * sizeof( double ), s );
   StencilCall c;
   c.mode = MODE_APPLY;
   c.flag = 15;
   c.out0 = ol.dinv;
   elemT< double, MODE_APPLY, true >( m, g, op, L, c, s );
   hipLaunchKernelGGL( ( kInvertAll< double > ), dim3( 1024 ), dim3( kBlock ), 0, s, n, ol.dinv );
   launchConvert( n, 4, ol.dinvF, 8, ol.dinv, s );
}

} // namespace hyg
