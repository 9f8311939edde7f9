// gfx950 kernels of the P2 (quadratic Lagrange) path: the constant-coefficient -Laplace operator
// (P2ConstantOperator< p2_tet_diffusion >, p2functionspace/P2ConstantOperator.cpp:99-113), its inverse
// diagonal, and the quadratic grid transfers (P2toP2QuadraticProlongation.cpp:217-424,
// P2toP2QuadraticRestriction.cpp:131-286).
//
// A P2 function of level L holds its vertex DoFs and the DoFs at the midpoints of the level-L micro-edges;
// together they are the vertices of level L+1, so it is stored like a P1 function of "lattice" level L+1
// (same per-cell arrays, replicas of face/edge/vertex values in every incident cell). Every routine here
// takes lattice levels; the micro-cells of the P2 elements are those of level lattice-1, whose ten nodes
// are lattice points (P2Tables::nd).
//
// The operator is evaluated element by element per point: a point sums, over the micro-cells of its
// macro-cell that have it as a node, the row of the cell-type's 10x10 element matrix times the ten node
// values. Cell-interior points get the mode's epilogue; points on faces, edges and vertices store the
// cell's one-sided sum for kFaceCombine / kLowCombine, exactly as the P1 elementwise operator does.
//
// Variable coefficient (P2ElementwiseOperator< p2_div_k_grad_affine_q4 >, P2ElementwiseOperator.cpp:100-310):
// the element matrix of micro-cell e of type t is sum_q k_e,q M_t,q with the 11 quadrature-point matrices
// M_t,q = w_q |det J| grad phi(l_q) grad phi(l_q)^T of the cell's type t (affine map: the same for every
// micro-cell of that type) and the coefficient k_e,q at the quadrature points, evaluated once at operator
// setup (kP2KQ). A point assembles the rows it needs on the fly.
#include "p2_common.hpp"

#include <utility>

namespace hyg {

// ------------------------------------------------------------------ variable coefficient at the quadrature points
// k(x) = k0 exp(alpha (1 - |x|)) (1 + beta xi_e), xi_e uniform in [-1, 1) from SplitMix64(seed, element id) (kind 0),
// or 1 + x/2 + y^2/4 (kind 1)
struct P2Coef
{
   int      kind;
   double   k0, alpha, beta;
   uint64_t seed;
};

__device__ __forceinline__ double p2CoefAt( const P2Coef& k, double x, double y, double z, uint64_t eid )
{
   if ( k.kind == 1 )
      return 1.0 + 0.5 * x + 0.25 * y * y;
   uint64_t h = k.seed + ( eid + 1 ) * 0x9E3779B97F4A7C15ull;
   h          = ( h ^ ( h >> 30 ) ) * 0xBF58476D1CE4E5B9ull;
   h          = ( h ^ ( h >> 27 ) ) * 0x94D049BB133111EBull;
   h ^= h >> 31;
   const double xi = 2.0 * ( double( h >> 11 ) * 0x1.0p-53 ) - 1.0;
   return k.k0 * exp( k.alpha * ( 1.0 - sqrt( x * x + y * y + z * z ) ) ) * ( 1.0 + k.beta * xi );
}

// One thread per anchor of the micro-cell lattice (level L-1) of a macro-cell; the valid types of that anchor.
// Element id (6 c + t) S3(Nc + 1) + index(Nc + 1, anchor); kq[q kqStride + id], kqStride = 6 nC S3(Nc + 1), so that
// threads of consecutive anchors read consecutive values of one quadrature point.
template < typename T >
__global__ __launch_bounds__( kBlock ) void kP2KQ( int L, int tilesPerCell, const ElemCell* __restrict__ cells, P2Coef coef,
                                                   T* __restrict__ kq, i64 kqStride )
{
   const int c    = int( blockIdx.x ) / tilesPerCell;
   const int tile = int( blockIdx.x ) - c * tilesPerCell;
   const int Nc = 1 << ( L - 1 ), SA = S3i( Nc + 1 );
   const int s = tile * kBlock + int( threadIdx.x );
   if ( s >= SA )
      return;
   int ax, ay, az;
   tetXYZ( Nc + 1, s, ax, ay, az );
   const ElemCell& ec = cells[c];
   for ( int t = 0; t < 6; ++t )
   {
      if ( ax + ay + az + kMCMax[t] > Nc )
         continue;
      double X[4][3];
#pragma unroll
      for ( int v = 0; v < 4; ++v )
      {
         const double i = ax + kMC[t][v][0], j = ay + kMC[t][v][1], k = az + kMC[t][v][2];
#pragma unroll
         for ( int r = 0; r < 3; ++r )
            X[v][r] = ec.v0[r] + ec.a[r][0] * i + ec.a[r][1] * j + ec.a[r][2] * k;
      }
      const uint64_t eid = uint64_t( c * 6 + t ) * uint64_t( SA ) + uint64_t( s );
      for ( int q = 0; q < 11; ++q )
      {
         double x[3];
#pragma unroll
         for ( int r = 0; r < 3; ++r )
            x[r] = kXG4Lambda[q][0] * X[0][r] + kXG4Lambda[q][1] * X[1][r] + kXG4Lambda[q][2] * X[2][r] + kXG4Lambda[q][3] * X[3][r];
         kq[i64( q ) * kqStride + i64( eid )] = T( p2CoefAt( coef, x[0], x[1], x[2], eid ) );
      }
   }
}

// ------------------------------------------------------------------ operator application
// KQ: variable coefficient (el = 11 quadrature-point matrices per cell and type, kq per micro-cell)
// points on the cell edges and vertices of a cell lattice with N intervals per edge: the four vertices, then
// N - 1 points of each edge e0 = (v0, v1) ... e5 = (v2, v3)
constexpr int kP2EvLanes = 4; // lanes per point of the edge/vertex kernel
__host__ __device__ constexpr int p2EdgeVertexCount( int N )
{
   return 4 + 6 * ( N - 1 );
}
__device__ __forceinline__ bool edgeVertexXYZ( int N, int s, int& x, int& y, int& z )
{
   x = y = z = 0;
   if ( s < 4 )
   {
      x = s == 1 ? N : 0;
      y = s == 2 ? N : 0;
      z = s == 3 ? N : 0;
      return true;
   }
   s -= 4;
   const int e = s / ( N - 1 ), i = s - e * ( N - 1 ) + 1;
   switch ( e )
   {
   case 0: x = i; break;
   case 1: y = i; break;
   case 2: x = N - i, y = i; break;
   case 3: z = i; break;
   case 4: x = N - i, z = i; break;
   case 5: y = N - i, z = i; break;
   default: return false;
   }
   return true;
}

// SKIPIN: the cell-interior and cell-face points are done by the stencil kernels; the grid then covers only the
// points on cell edges and vertices, kP2EvLanes lanes per point (each sums every kP2EvLanes-th incident micro-cell)
template < typename T, int MODE, bool DIAG, bool KQ, bool SKIPIN = false >
__global__ __launch_bounds__( kBlock ) void kP2Apply( int L, int tilesPerCell, i64 stride, const T* __restrict__ el,
                                                      const T* __restrict__ kq, i64 kqStride,
                                                      const P2Tables* __restrict__ tb, const CellInfo* __restrict__ cinfo,
                                                      const CellFaceSlot* __restrict__ cfs, const LowerPrim* __restrict__ prims,
                                                      const LowerEntry* __restrict__ entries, T* __restrict__ part,
                                                      T* __restrict__ lowPart, const T* __restrict__ dinv, int flag,
                                                      const T* __restrict__ x, const T* __restrict__ in1, T* __restrict__ out0,
                                                      T* __restrict__ out1, T a )
{
   const int c    = int( blockIdx.x ) / tilesPerCell;
   const int tile = int( blockIdx.x ) - c * tilesPerCell;
   const int N = 1 << L, w = N + 1, NP = N + 1, Nc = N >> 1;
   const int s   = SKIPIN ? ( tile * kBlock + int( threadIdx.x ) ) / kP2EvLanes : tile * kBlock + int( threadIdx.x );
   const int sub = SKIPIN ? int( threadIdx.x ) % kP2EvLanes : 0;
   int       px, py, pz;
   if constexpr ( SKIPIN )
   {
      // the lanes of one point leave together
      if ( s >= p2EdgeVertexCount( N ) || !edgeVertexXYZ( N, s, px, py, pz ) )
         return;
   }
   else
   {
      if ( s >= S3i( w ) )
         return;
      tetXYZ( w, s, px, py, pz );
   }
   const int slot = classifySlot( N, px, py, pz );
   if ( slot == 0 && ( SKIPIN || !( flag & 1 ) ) )
      return;
   const int cls = ( px & 1 ) | ( ( py & 1 ) << 1 ) | ( ( pz & 1 ) << 2 );
   const T*  ec  = el + i64( c ) * ( KQ ? 6600 : 600 );
   const int SA  = S3i( Nc + 1 );
   const T*  xc  = x + i64( c ) * stride;
   T         val = T( 0 );
   const int ne  = tb->nEnt[cls];
   for ( int k = sub; k < ne; k += ( SKIPIN ? kP2EvLanes : 1 ) )
   {
      const int tl = tb->ent[cls][k];
      const int t = tl >> 4, l = tl & 15;
      // anchor of the micro-cell of type t that has this point as node l
      const int ax = ( px - tb->nd[t][l][0] ) >> 1, ay = ( py - tb->nd[t][l][1] ) >> 1, az = ( pz - tb->nd[t][l][2] ) >> 1;
      if ( ax < 0 || ay < 0 || az < 0 || ax + ay + az + kMCMax[t] > Nc )
         continue;
      if constexpr ( KQ )
      {
         const T* M  = ec + 1100 * t + 10 * l; // M[100 q + j]: row l of quadrature-point matrix q
         const T* kv = kq + i64( c * 6 + t ) * SA + cidx( Nc + 1, ax, ay, az );
         T        kk[11];
#pragma unroll
         for ( int q = 0; q < 11; ++q )
            kk[q] = kv[i64( q ) * kqStride];
         if constexpr ( DIAG )
         {
#pragma unroll
            for ( int q = 0; q < 11; ++q )
               val += kk[q] * M[100 * q + l];
         }
         else
         {
#pragma unroll
            for ( int j = 0; j < 10; ++j )
            {
               T r = T( 0 );
#pragma unroll
               for ( int q = 0; q < 11; ++q )
                  r += kk[q] * M[100 * q + j];
               val += r * xc[cidx( w, 2 * ax + tb->nd[t][j][0], 2 * ay + tb->nd[t][j][1], 2 * az + tb->nd[t][j][2] )];
            }
         }
      }
      else
      {
         const T* row = ec + 100 * t + 10 * l;
         if constexpr ( DIAG )
            val += row[l];
         else
         {
#pragma unroll
            for ( int j = 0; j < 10; ++j )
               val += row[j] * xc[cidx( w, 2 * ax + tb->nd[t][j][0], 2 * ay + tb->nd[t][j][1], 2 * az + tb->nd[t][j][2] )];
         }
      }
   }
   if constexpr ( SKIPIN )
   {
#pragma unroll
      for ( int m = 1; m < kP2EvLanes; m <<= 1 )
         val += __shfl_xor( val, m, kP2EvLanes );
      if ( sub != 0 )
         return;
   }
   const i64 q = i64( c ) * stride + cidx( w, px, py, pz );
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

// ------------------------------------------------------------------ constant operator: cell-face points
// Points inside a cell face (not on a cell edge) take the one-sided merged stencil of their face and parity
// class; the sum goes to the face's partial-sum block like the element kernel's. On each face the points split
// into four sub-lattices by the parities of their face coordinates (u, v) = (2 i + 2 - ku, 2 j + 2 - kv), u, v >= 1,
// u + v <= N - 1, and each sub-lattice has one parity class, so the offsets are known at compile time. The cell
// point of (u, v) is (u, v, 0), (u, 0, v), (0, u, v) and (u, N - u - v, v) on faces 0..3: consecutive threads (u)
// stay in one z-plane, on the diagonal face too.

// Stored stencils of the variable-coefficient operator: per cell, the merged stencils of the cell-interior
// points class by class ([e][point of the class sub-lattice], the class-kappa points p = 2 q + kappa of the cell
// form a tetrahedral lattice of width (N - |kappa|) / 2 + 1), then the one-sided stencils of the cell-face points
// face by face and sub-lattice by sub-lattice ([e][face sub-lattice index]).
__host__ __device__ constexpr int p2SubWidth( int N, int cls )
{
   return ( N - ( cls & 1 ) - ( ( cls >> 1 ) & 1 ) - ( ( cls >> 2 ) & 1 ) ) / 2 + 1;
}
__host__ __device__ inline i64 p2StoredIntBase( int N, int cls )
{
   constexpr int n[8] = { p2StencilOffsets( 0 ).n, p2StencilOffsets( 1 ).n, p2StencilOffsets( 2 ).n, p2StencilOffsets( 3 ).n,
                          p2StencilOffsets( 4 ).n, p2StencilOffsets( 5 ).n, p2StencilOffsets( 6 ).n, p2StencilOffsets( 7 ).n };
   i64           b    = 0;
   for ( int k = 0; k < cls; ++k )
      b += i64( n[k] ) * S3i( p2SubWidth( N, k ) );
   return b;
}
__host__ __device__ inline i64 p2StoredFaceBase( int N, int F, int sub )
{
   constexpr int nf[4][4] = { { p2StencilOffsets( p2FaceClass( 0, 0, 0 ), 0 ).n, p2StencilOffsets( p2FaceClass( 0, 1, 0 ), 0 ).n,
                                p2StencilOffsets( p2FaceClass( 0, 0, 1 ), 0 ).n, p2StencilOffsets( p2FaceClass( 0, 1, 1 ), 0 ).n },
                              { p2StencilOffsets( p2FaceClass( 1, 0, 0 ), 1 ).n, p2StencilOffsets( p2FaceClass( 1, 1, 0 ), 1 ).n,
                                p2StencilOffsets( p2FaceClass( 1, 0, 1 ), 1 ).n, p2StencilOffsets( p2FaceClass( 1, 1, 1 ), 1 ).n },
                              { p2StencilOffsets( p2FaceClass( 2, 0, 0 ), 2 ).n, p2StencilOffsets( p2FaceClass( 2, 1, 0 ), 2 ).n,
                                p2StencilOffsets( p2FaceClass( 2, 0, 1 ), 2 ).n, p2StencilOffsets( p2FaceClass( 2, 1, 1 ), 2 ).n },
                              { p2StencilOffsets( p2FaceClass( 3, 0, 0 ), 3 ).n, p2StencilOffsets( p2FaceClass( 3, 1, 0 ), 3 ).n,
                                p2StencilOffsets( p2FaceClass( 3, 0, 1 ), 3 ).n, p2StencilOffsets( p2FaceClass( 3, 1, 1 ), 3 ).n } };
   i64 b = p2StoredIntBase( N, 8 );
   for ( int k = 0; k < 4 * F + sub; ++k )
      b += i64( nf[k >> 2][k & 3] ) * p2FaceSubCount( N, k & 1, ( k >> 1 ) & 1 );
   return b;
}
__host__ __device__ inline i64 p2StoredCellSize( int N )
{
   return ( p2StoredFaceBase( N, 4, 0 ) + 63 ) / 64 * 64;
}

template < typename T, int F, int KU, int KV, bool STORED >
__device__ __forceinline__ void p2FacePoint( int c, int s, int N, i64 stride, const T* __restrict__ fst, const CellFaceSlot& fs,
                                             T* __restrict__ part, const T* __restrict__ x )
{
   // face coordinates (u, v) -> cell coordinates; parity class of the points
   constexpr int              CLS = p2FaceClass( F, KU, KV );
   constexpr P2StencilOffsets off = p2StencilOffsets( CLS, F );
   const int                  M   = ( N - 5 + KU + KV ) / 2;
   if ( s >= triCount( M + 1 ) )
      return;
   int i, j;
   planeXY( M + 1, s, i, j );
   const int u = 2 * i + 2 - KU, v = 2 * j + 2 - KV;
   const int px = F == 2 ? 0 : u, py = F == 0 ? v : ( F == 2 ? u : ( F == 3 ? N - u - v : 0 ) ), pz = F == 0 ? 0 : v;
   const int w = N + 1;
   typedef const T __attribute__( ( address_space( 4 ) ) ) CT;
   // uniform weights of the cell (constant coefficient), or this point's stored stencil
   const CT* wc = (const CT*) ( fst + ( i64( c * 4 + F ) * 8 + CLS ) * kP2StencilMax );
   const T*  ws = fst + i64( c ) * p2StoredCellSize( N ) + p2StoredFaceBase( N, F, KU | ( KV << 1 ) ) + s;
   const int cnt = triCount( M + 1 );
   const T*  xc  = x + i64( c ) * stride;
   int       rb[5][5];
#pragma unroll
   for ( int dz = -2; dz <= 2; ++dz )
   {
      const int W  = w - ( pz + dz );
      const int pb = S3i( w ) - S3i( W );
#pragma unroll
      for ( int dy = -2; dy <= 2; ++dy )
         rb[dz + 2][dy + 2] = pb + rowStart( W, py + dy );
   }
   T val = T( 0 );
#pragma unroll
   for ( int e = 0; e < off.n; ++e )
      val += ( STORED ? ws[i64( e ) * cnt] : wc[e] ) * xc[rb[off.d[e][2] + 2][off.d[e][1] + 2] + px + off.d[e][0]];
   part[i64( fs.block ) * w * w + facePartIndex( F, N, px, py, pz )] = val;
}

template < typename T, bool STORED >
__global__ __launch_bounds__( kBlock ) void kP2Face( int L, i64 stride, P2FaceTiles ft, const T* __restrict__ fst,
                                                     const CellFaceSlot* __restrict__ cfs, int flag, T* __restrict__ part,
                                                     const T* __restrict__ x, int nC )
{
   const int per = ft.start[4];
   // nC > 0: each XCD takes a contiguous range of cells (grid padded to a multiple of 8), keeping a cell's data in one
   // L2 (pays off for the stored stencils of many cells; the constant operator's launch keeps the plain order)
   const int b = nC > 0 ? remapXcd( int( blockIdx.x ), nC * 4 * per ) : int( blockIdx.x );
   if ( nC > 0 && b >= nC * 4 * per )
      return;
   const int c = b / ( 4 * per );
   int       r = b - c * 4 * per;
   const int f   = r / per;
   r -= f * per;
   int sub = 0;
   while ( sub < 3 && r >= ft.start[sub + 1] )
      ++sub;
   const int          s  = ( r - ft.start[sub] ) * kBlock + int( threadIdx.x );
   const CellFaceSlot fs = cfs[4 * c + f];
   if ( !( fs.bc & flag ) )
      return;
   const int N = 1 << L;
#define HYG_P2_FACE( F, SUB )                                                                                         \
   case 4 * F + SUB: p2FacePoint< T, F, ( SUB & 1 ), ( SUB >> 1 ), STORED >( c, s, N, stride, fst, fs, part, x ); break;
   switch ( 4 * f + sub )
   {
      HYG_P2_FACE( 0, 0 )
      HYG_P2_FACE( 0, 1 )
      HYG_P2_FACE( 0, 2 )
      HYG_P2_FACE( 0, 3 )
      HYG_P2_FACE( 1, 0 )
      HYG_P2_FACE( 1, 1 )
      HYG_P2_FACE( 1, 2 )
      HYG_P2_FACE( 1, 3 )
      HYG_P2_FACE( 2, 0 )
      HYG_P2_FACE( 2, 1 )
      HYG_P2_FACE( 2, 2 )
      HYG_P2_FACE( 2, 3 )
      HYG_P2_FACE( 3, 0 )
      HYG_P2_FACE( 3, 1 )
      HYG_P2_FACE( 3, 2 )
   default: p2FacePoint< T, 3, 1, 1, STORED >( c, s, N, stride, fst, fs, part, x ); break;
   }
#undef HYG_P2_FACE
}

// ------------------------------------------------------------------ constant operator: cell-interior points
constexpr int p2CentreIndex( const P2StencilOffsets& o )
{
   for ( int e = 0; e < o.n; ++e )
      if ( o.d[e][0] == 0 && o.d[e][1] == 0 && o.d[e][2] == 0 )
         return e;
   return -1;
}

// A cell-interior point's merged stencil depends on its parity class (x, y, z) mod 2. One wave computes the rows
// y0 (even) and y0+1 of one x segment: lane i holds the pair x0 + 2i (even x) and x0 + 2i + 1 (odd x),
// x0 = 124 k - 2, so each of the wave's four stencils per plane (even / odd x of either row) has compile-time
// offsets and weights that are the same for every lane (scalar loads). A value at x + dx reaches the lane that
// needs it through partial sums: the terms a lane's values add to its left and right neighbours are summed in
// place and shifted once per row (lanes 0 and 63 are halo only). The wave marches the planes z0 .. z1-1 (z0
// even); a register ring holds rows y0-2 .. y0+2 (every row either row's stencils reach: dy in [-2, 2] for the
// even row, [-1, 1] for the odd one) of the planes z-2 .. z+2 plus D planes in flight, so a plane step issues
// five pair loads for two output rows (a row-by-row kernel reads 9 to 19 stencil rows per output row). The
// parity of z alternates the step's classes between (y, z) = (even, even), (odd, even) and (even, odd),
// (odd, odd). The mode's epilogue (apply, residual, Jacobi, Chebyshev) is fused.
template < typename T >
struct P2ZPlane
{
   T v[5][2]; // rows y0-2 .. y0+2: (even x, odd x) of the lane
   T e[2][2]; // the epilogue input of rows y0 and y0+1
};

// the pair (even x, odd x) of row y0 + KY at plane z (parity KZ) from the planes A .. E = z-2 .. z+2
template < typename T, int KY, int KZ, int E0 >
__device__ __forceinline__ void p2ZTerms( const P2ZPlane< T >& A, const P2ZPlane< T >& B, const P2ZPlane< T >& C,
                                          const P2ZPlane< T >& Dn, const P2ZPlane< T >& E, const T* __restrict__ wc, T& a, T& r,
                                          T& l )
{
   constexpr int              cls = ( KY << 1 ) | ( KZ << 2 ) | E0;
   constexpr P2StencilOffsets o   = p2StencilOffsets( cls );
   typedef const T __attribute__( ( address_space( 4 ) ) ) CT;
   const CT* w = (const CT*) wc;
#pragma unroll
   for ( int e = 0; e < o.n; ++e )
   {
      const P2ZPlane< T >& Q  = o.d[e][2] == -2 ? A : o.d[e][2] == -1 ? B : o.d[e][2] == 0 ? C : o.d[e][2] == 1 ? Dn : E;
      const int            ry = KY + o.d[e][1] + 2;
      const T              vE = Q.v[ry][0], vO = Q.v[ry][1], wt = w[e];
      // even point x: offsets -2 / -1 come from lane-1 (shifted right), 0 / +1 are the lane's own pair, +2 comes
      // from lane+1 (shifted left); odd point x+1: -2 from lane-1, -1 / 0 own, +1 / +2 from lane+1
      if constexpr ( E0 == 0 )
         switch ( o.d[e][0] )
         {
         case -2: r += wt * vE; break;
         case -1: r += wt * vO; break;
         case 0: a += wt * vE; break;
         case 1: a += wt * vO; break;
         default: l += wt * vE; break;
         }
      else
         switch ( o.d[e][0] )
         {
         case -2: r += wt * vO; break;
         case -1: a += wt * vE; break;
         case 0: a += wt * vO; break;
         case 1: l += wt * vE; break;
         default: l += wt * vO; break;
         }
   }
}

// f(0), f(1), ... in order while they return true
template < typename F, int... I >
__device__ __forceinline__ void p2ZSteps( F& f, std::integer_sequence< int, I... > )
{
   (void) ( f( std::integral_constant< int, I >{} ) && ... );
}

template < typename T, int MODE, int D >
__global__ __launch_bounds__( kBlock ) void kP2Rows( int L, int nC, i64 stride, const ZSeg* __restrict__ segs, int nSegs,
                                                      const T* __restrict__ st, const T* __restrict__ x, const T* __restrict__ in1,
                                                      T* __restrict__ out0, T* __restrict__ out1, T a )
{
   constexpr bool kIn  = MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK || MODE == MODE_APPLY_ADD;
   constexpr bool kTwo = MODE == MODE_CHEB0 || MODE == MODE_CHEBK;
   constexpr bool kDiv = MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK;
   constexpr int  B    = int( sizeof( T ) );
   constexpr int  NR   = 5 + D; // ring slots: planes z-2 .. z+2 resident, D in flight after a step's loads
   const int      nbx  = ( nSegs + 3 ) / 4;
   const int      nb   = nbx * nC;
   const int      lb   = remapXcd( int( blockIdx.x ), nb );
   if ( lb >= nb )
      return;
   const int c  = lb / nbx;
   const int si = ( lb - c * nbx ) * 4 + __builtin_amdgcn_readfirstlane( int( threadIdx.x ) >> 6 );
   if ( si >= nSegs )
      return;
   const ZSeg     sg   = segs[si];
   const int      N    = 1 << L, w = N + 1, y0 = sg.y, z0 = sg.z0, z1 = sg.z1;
   const int      lane = int( threadIdx.x ) & 63;
   const int      xe   = kP2ZSeg * sg.x0 - 2 + 2 * lane;
   const i64      cb   = i64( c ) * stride;
   const unsigned nby  = unsigned( S3i( w ) ) * unsigned( B );
   const auto     rx   = makeRsrc( x + cb, nby );
   const auto     re   = makeRsrc( ( MODE == MODE_APPLY_ADD ? out0 : in1 ) + cb, nby );
   const auto     r0   = makeRsrc( out0 + cb, nby );
   const auto     r1   = makeRsrc( kTwo ? out1 + cb : out0 + cb, nby );
   const T*       wcl  = st + i64( c ) * 8 * kP2StencilMax;

   // the inverse diagonal of a cell-interior point is the inverse of its merged stencil's centre weight
   T dvc[8];
#pragma unroll
   for ( int k = 0; k < 8; ++k )
      dvc[k] = T( 1 );
   if constexpr ( kDiv )
   {
      constexpr int ce[8] = { p2CentreIndex( p2StencilOffsets( 0 ) ), p2CentreIndex( p2StencilOffsets( 1 ) ),
                              p2CentreIndex( p2StencilOffsets( 2 ) ), p2CentreIndex( p2StencilOffsets( 3 ) ),
                              p2CentreIndex( p2StencilOffsets( 4 ) ), p2CentreIndex( p2StencilOffsets( 5 ) ),
                              p2CentreIndex( p2StencilOffsets( 6 ) ), p2CentreIndex( p2StencilOffsets( 7 ) ) };
#pragma unroll
      for ( int k = 0; k < 8; ++k )
         dvc[k] = T( 1 ) / wcl[k * kP2StencilMax + ce[k]];
   }
   // plane p: rows y0-2 .. y0+2 (lanes outside a row, rows outside the plane and planes outside the cell read 0)
   int  pl        = z0 - 2;
   auto loadPlane = [&]( P2ZPlane< T >& Q ) __attribute__( ( always_inline ) ) {
      const int  W   = w - pl;
      const bool okp = pl >= 0 && pl <= N;
      const int  bz  = S3i( w ) - S3i( okp ? W : w );
#pragma unroll
      for ( int j = 0; j < 5; ++j )
      {
         const int      yy = y0 - 2 + j;
         const bool     ok = okp && yy >= 0 && xe >= 0 && xe < W - yy;
         const unsigned vo = ok ? unsigned( bz + rowStart( W, yy ) + xe ) * unsigned( B ) : kDropOffset;
         bufLoad2( rx, vo, 0, Q.v[j][0], Q.v[j][1] );
      }
      if constexpr ( kIn )
      {
#pragma unroll
         for ( int r = 0; r < 2; ++r )
         {
            const int      yy = y0 + r;
            const bool     ok = okp && yy >= 1 && pl >= 1 && xe >= 0 && xe <= N - 1 - yy - pl;
            const unsigned vo = ok ? unsigned( bz + rowStart( W, yy ) + xe ) * unsigned( B ) : kDropOffset;
            bufLoad2( re, vo, 0, Q.e[r][0], Q.e[r][1] );
         }
      }
      ++pl;
   };
   int  z    = z0;
   auto step = [&]( auto kz, const P2ZPlane< T >& A, const P2ZPlane< T >& Bp, const P2ZPlane< T >& C, const P2ZPlane< T >& Dn,
                    const P2ZPlane< T >& E, P2ZPlane< T >& F ) __attribute__( ( always_inline ) ) {
      constexpr int KZ = decltype( kz )::value;
      loadPlane( F );
      const int W  = w - z;
      const int bz = S3i( w ) - S3i( W );
      auto      rowOut = [&]( auto ky ) __attribute__( ( always_inline ) ) {
         constexpr int KY = decltype( ky )::value;
         constexpr int cE = ( KY << 1 ) | ( KZ << 2 );
         T             aE = T( 0 ), aO = T( 0 ), rE = T( 0 ), rO = T( 0 ), lE = T( 0 ), lO = T( 0 );
         // the weights stay in the scalar cache: an opaque copy of the pointer per step keeps the compiler from
         // holding all four classes of the march in registers
         const unsigned long long wv = reinterpret_cast< unsigned long long >( wcl );
         unsigned                 wlo = __builtin_amdgcn_readfirstlane( unsigned( wv ) );
         unsigned                 whi = __builtin_amdgcn_readfirstlane( unsigned( wv >> 32 ) );
         asm volatile( "" : "+s"( wlo ), "+s"( whi ) );
         const T* wb = reinterpret_cast< const T* >( ( ( unsigned long long ) whi << 32 ) | wlo );
         const T*      wE = wb + cE * kP2StencilMax;
         const T*      wO = wE + kP2StencilMax;
         p2ZTerms< T, KY, KZ, 0 >( A, Bp, C, Dn, E, wE, aE, rE, lE );
         p2ZTerms< T, KY, KZ, 1 >( A, Bp, C, Dn, E, wO, aO, rO, lO );
         const T    valE = aE + laneShift< kDppWaveShr1 >( rE ) + laneShift< kDppWaveShl1 >( lE );
         const T    valO = aO + laneShift< kDppWaveShr1 >( rO ) + laneShift< kDppWaveShl1 >( lO );
         const int  yy   = y0 + KY;
         const int  xmax = N - 1 - yy - z; // last interior x of the row
         const int  row  = bz + rowStart( W, yy );
         const bool okr  = lane >= 1 && lane <= 62 && yy >= 1 && z >= 1;
         T o0E, o1E, o0O, o1O;
         epilogue< T, MODE >( valE, C.v[KY + 2][0], C.e[KY][0], C.e[KY][0], dvc[cE], a, o0E, o1E );
         epilogue< T, MODE >( valO, C.v[KY + 2][1], C.e[KY][1], C.e[KY][1], dvc[cE | 1], a, o0O, o1O );
         const unsigned sE = okr && xe >= 1 && xe <= xmax ? unsigned( row + xe ) * unsigned( B ) : kDropOffset;
         const unsigned sO = okr && xe + 1 >= 1 && xe + 1 <= xmax ? unsigned( row + xe + 1 ) * unsigned( B ) : kDropOffset;
         bufStore( o0E, r0, sE, 0 );
         bufStore( o0O, r0, sO, 0 );
         if constexpr ( kTwo )
         {
            bufStore( o1E, r1, sE, 0 );
            bufStore( o1O, r1, sO, 0 );
         }
      };
      rowOut( std::integral_constant< int, 0 >{} );
      rowOut( std::integral_constant< int, 1 >{} );
      ++z;
   };
   P2ZPlane< T > ring[NR];
#pragma unroll
   for ( int u = 0; u < NR; ++u )
#pragma unroll
      for ( int r = 0; r < 2; ++r )
         ring[u].e[r][0] = ring[u].e[r][1] = T( 0 );
#pragma unroll
   for ( int u = 0; u < NR - 1; ++u )
      loadPlane( ring[u] ); // planes z0-2 .. z0+1+D
   // ring slot u % NR holds plane z0 - 2 + u; U steps return the ring to its start and keep the parity of z
   // (z0 is even, so step u has parity u & 1). The steps are expanded at compile time, so every ring slot is
   // a fixed set of registers.
   constexpr int U     = ( NR & 1 ) ? 2 * NR : NR;
   auto          stepU = [&]( auto uc ) __attribute__( ( always_inline ) ) {
      constexpr int u = decltype( uc )::value;
      step( std::integral_constant< int, ( u & 1 ) >{}, ring[u % NR], ring[( u + 1 ) % NR], ring[( u + 2 ) % NR], ring[( u + 3 ) % NR],
            ring[( u + 4 ) % NR], ring[( u + NR - 1 ) % NR] );
      return z < z1;
   };
   while ( z < z1 )
      p2ZSteps( stepU, std::make_integer_sequence< int, U >{} );
}

// ------------------------------------------------------------------ transfers: cell-interior points
// Interior fine points of one residue class r = P mod 4 share one interpolation stencil (coarse offsets and
// weights known at compile time); interior coarse points of one parity class share one restriction stencil.

// Interior fine points row by row: one wave per fine row (y, z), y, z >= 1, y + z <= Nf - 2; lane l holds the four
// consecutive points x = 4 q + r, r = 0..3, q = 64 j + l in pass j. All points of the wave share (y mod 4, z mod 4),
// so each of a lane's four points has its residue class's compile-time stencil (p2ProlongStencil, coarse offsets
// from (2 q, 2 (y div 4), 2 (z div 4))), and the wave reads and writes whole fine rows.
template < typename T, int RY, int RZ >
__device__ __forceinline__ void p2ProlongRow( int y, int z, int Nf, const T* __restrict__ cs, T* __restrict__ fs, int add )
{
   const int lane = int( threadIdx.x ) & 63;
   const int wc = ( Nf >> 1 ) + 1, wf = Nf + 1;
   const int qy = y >> 2, qz = z >> 2;
   const int xmax = Nf - 1 - y - z; // last interior x of the row
   const int rowF = S3i( wf ) - S3i( wf - z ) + rowStart( wf - z, y );
   for ( int x0 = 4 * lane; x0 <= xmax; x0 += 256 )
   {
      const int qx = x0 >> 2;
      T         v[4];
      auto      point = [&]( auto rx ) __attribute__( ( always_inline ) ) {
         constexpr int              RX = decltype( rx )::value;
         constexpr P2ProlongStencil ps = p2ProlongStencil( RX, RY, RZ );
         T                          u  = T( 0 );
         // only interior points: their stencils stay inside the coarse cell
         if ( x0 + RX >= 1 && x0 + RX <= xmax )
         {
#pragma unroll
            for ( int k = 0; k < ps.n; ++k )
               u += T( ps.w[k] ) * cs[cidx( wc, 2 * qx + ps.d[k][0], 2 * qy + ps.d[k][1], 2 * qz + ps.d[k][2] )];
         }
         v[RX] = u;
      };
      point( std::integral_constant< int, 0 >{} );
      point( std::integral_constant< int, 1 >{} );
      point( std::integral_constant< int, 2 >{} );
      point( std::integral_constant< int, 3 >{} );
#pragma unroll
      for ( int r = 0; r < 4; ++r )
      {
         const int x = x0 + r;
         if ( x >= 1 && x <= xmax )
            fs[rowF + x] = add ? fs[rowF + x] + v[r] : v[r];
      }
   }
}
template < typename T >
__global__ __launch_bounds__( kBlock ) void kP2ProlongRows( int Lf, int nC, i64 strideF, i64 strideC, int flag, T* __restrict__ fine,
                                                            const T* __restrict__ coarse, int add )
{
   // cell-interior points are prolongated whatever the flag (the reference's cell loop has no flag test,
   // P2toP2QuadraticProlongation.cpp:217-424)
   const int Nf = 1 << Lf, W = Nf - 3; // fine rows: a triangle with W rows
   if ( W <= 0 )
      return;
   const int rows = W * ( W + 1 ) / 2, nbx = ( rows + 3 ) / 4, nb = nbx * nC;
   const int lb   = remapXcd( int( blockIdx.x ), nb );
   if ( lb >= nb )
      return;
   const int c = lb / nbx;
   const int r = ( lb - c * nbx ) * 4 + __builtin_amdgcn_readfirstlane( int( threadIdx.x ) >> 6 );
   if ( r >= rows )
      return;
   int a, b;
   planeXY( W, r, a, b );
   const int y = a + 1, z = b + 1;
   const T*  cs = coarse + i64( c ) * strideC;
   T*        fs = fine + i64( c ) * strideF;
#define HYG_P2_PROW( RY, RZ )                                                                                          \
   case ( RY ) | ( ( RZ ) << 2 ): p2ProlongRow< T, RY, RZ >( y, z, Nf, cs, fs, add ); break;
   switch ( ( y & 3 ) | ( ( z & 3 ) << 2 ) )
   {
      HYG_P2_PROW( 0, 0 )
      HYG_P2_PROW( 1, 0 )
      HYG_P2_PROW( 2, 0 )
      HYG_P2_PROW( 3, 0 )
      HYG_P2_PROW( 0, 1 )
      HYG_P2_PROW( 1, 1 )
      HYG_P2_PROW( 2, 1 )
      HYG_P2_PROW( 3, 1 )
      HYG_P2_PROW( 0, 2 )
      HYG_P2_PROW( 1, 2 )
      HYG_P2_PROW( 2, 2 )
      HYG_P2_PROW( 3, 2 )
      HYG_P2_PROW( 0, 3 )
      HYG_P2_PROW( 1, 3 )
      HYG_P2_PROW( 2, 3 )
   default: p2ProlongRow< T, 3, 3 >( y, z, Nf, cs, fs, add ); break;
   }
#undef HYG_P2_PROW
}

template < typename T, int CLS >
__device__ __forceinline__ void p2RestrictPt( int c, int s, int Nc, i64 strideC, i64 strideF, T* __restrict__ coarse,
                                              const T* __restrict__ fine )
{
   constexpr P2RestrictStencil rs = p2RestrictStencil( CLS );
   constexpr int               kx = CLS & 1, ky = ( CLS >> 1 ) & 1, kz = ( CLS >> 2 ) & 1;
   const int                   M  = p2InteriorExtent( Nc, CLS );
   if ( M < 0 || s >= S3i( M + 1 ) )
      return;
   int qx, qy, qz;
   tetXYZ( M + 1, s, qx, qy, qz );
   const int nx = 2 * ( qx + 1 - kx ) + kx, ny = 2 * ( qy + 1 - ky ) + ky, nz = 2 * ( qz + 1 - kz ) + kz;
   const int wc = Nc + 1, wf = 2 * Nc + 1;
   // start of fine row 2 ny + dy in plane 2 nz + dz, |dy|, |dz| <= 3
   int rb[7][7];
#pragma unroll
   for ( int dz = -3; dz <= 3; ++dz )
   {
      const int W  = wf - ( 2 * nz + dz );
      const int pb = S3i( wf ) - S3i( W );
#pragma unroll
      for ( int dy = -3; dy <= 3; ++dy )
         rb[dz + 3][dy + 3] = pb + rowStart( W, 2 * ny + dy );
   }
   const T* fs = fine + i64( c ) * strideF;
   T        v  = T( 0 );
#pragma unroll
   for ( int k = 0; k < rs.n; ++k )
      v += T( rs.w[k] ) * fs[rb[rs.d[k][2] + 3][rs.d[k][1] + 3] + 2 * nx + rs.d[k][0]];
   coarse[i64( c ) * strideC + cidx( wc, nx, ny, nz )] = v;
}
template < typename T, int CLS >
__device__ __forceinline__ void p2RestrictDispatch( int cls, int c, int s, int Nc, i64 strideC, i64 strideF, T* __restrict__ coarse,
                                                    const T* __restrict__ fine )
{
   if ( cls == CLS )
      p2RestrictPt< T, CLS >( c, s, Nc, strideC, strideF, coarse, fine );
   else if constexpr ( CLS < 7 )
      p2RestrictDispatch< T, CLS + 1 >( cls, c, s, Nc, strideC, strideF, coarse, fine );
}
template < typename T >
__global__ __launch_bounds__( kBlock ) void kP2RestrictInt( int Lc, i64 strideC, i64 strideF, P2ClassTiles ct, T* __restrict__ coarse,
                                                            const T* __restrict__ fine )
{
   const int per = ct.start[8];
   const int c   = int( blockIdx.x ) / per;
   const int t   = int( blockIdx.x ) - c * per;
   int       cls = 0;
   while ( cls < 7 && t >= ct.start[cls + 1] )
      ++cls;
   p2RestrictDispatch< T, 0 >( cls, c, ( t - ct.start[cls] ) * kBlock + int( threadIdx.x ), 1 << Lc, strideC, strideF, coarse, fine );
}

// Stored stencils, cell-interior points one class per wave (the class sub-lattice, p2InteriorExtent): each
// thread gathers its point's neighbours and reads its own weights ([e][sub-lattice point], consecutive threads
// read consecutive weights), then the mode's epilogue; the inverse diagonal is the inverse centre weight.

template < typename T, int MODE, int CLS, int NC >
__device__ __forceinline__ void p2StoredIntPt( int c, int s, int N, i64 stride, const T* __restrict__ sw, const P2Comps< T >& cp, T a )
{
   constexpr int              kx = CLS & 1, ky = ( CLS >> 1 ) & 1, kz = ( CLS >> 2 ) & 1;
   constexpr P2StencilOffsets off = p2StencilOffsets( CLS );
   const int                  M   = p2InteriorExtent( N, CLS );
   if ( M < 0 || s >= S3i( M + 1 ) )
      return;
   int qx, qy, qz;
   tetXYZ( M + 1, s, qx, qy, qz );
   const int px = 2 * ( qx + 1 - kx ) + kx, py = 2 * ( qy + 1 - ky ) + ky, pz = 2 * ( qz + 1 - kz ) + kz;
   const int w = N + 1, SW = p2SubWidth( N, CLS ), nS = S3i( SW );
   const T*  ws = sw + i64( c ) * p2StoredCellSize( N ) + p2StoredIntBase( N, CLS ) + cidx( SW, px >> 1, py >> 1, pz >> 1 );
   int       rb[5][5];
#pragma unroll
   for ( int dz = -2; dz <= 2; ++dz )
   {
      const int W  = w - ( pz + dz );
      const int pb = S3i( w ) - S3i( W );
#pragma unroll
      for ( int dy = -2; dy <= 2; ++dy )
         rb[dz + 2][dy + 2] = pb + rowStart( W, py + dy );
   }
   // each weight is read once and applied to every component
   T val[NC];
#pragma unroll
   for ( int k = 0; k < NC; ++k )
      val[k] = T( 0 );
#pragma unroll
   for ( int e = 0; e < off.n; ++e )
   {
      const T   wt = ws[i64( e ) * nS];
      const i64 o  = i64( c ) * stride + rb[off.d[e][2] + 2][off.d[e][1] + 2] + px + off.d[e][0];
#pragma unroll
      for ( int k = 0; k < NC; ++k )
         val[k] += wt * cp.x[k][o];
   }
   const i64 q  = i64( c ) * stride + rb[2][2] + px;
   T         dv = T( 1 );
   if constexpr ( MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
      dv = T( 1 ) / ws[i64( p2CentreIndex( off ) ) * nS];
#pragma unroll
   for ( int k = 0; k < NC; ++k )
   {
      T i1 = T( 0 ), r0, r1, xv = T( 0 );
      if constexpr ( MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
         i1 = cp.in1[k][q];
      if constexpr ( MODE == MODE_APPLY_ADD )
         i1 = cp.out0[k][q];
      if constexpr ( MODE == MODE_JACOBI || MODE == MODE_CHEB0 )
         xv = cp.x[k][q];
      epilogue< T, MODE >( val[k], xv, i1, i1, dv, a, r0, r1 );
      cp.out0[k][q] = r0;
      if constexpr ( MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
         cp.out1[k][q] = r1;
   }
}
template < typename T, int MODE, int CLS, int NC >
__device__ __forceinline__ void p2StoredIntDispatch( int cls, int c, int s, int N, i64 stride, const T* sw, const P2Comps< T >& cp, T a )
{
   if ( cls == CLS )
      p2StoredIntPt< T, MODE, CLS, NC >( c, s, N, stride, sw, cp, a );
   else if constexpr ( CLS < 7 )
      p2StoredIntDispatch< T, MODE, CLS + 1, NC >( cls, c, s, N, stride, sw, cp, a );
}
template < typename T, int MODE, int NC >
__global__ __launch_bounds__( kBlock ) void kP2StoredInt( int L, i64 stride, P2ClassTiles ct, const T* __restrict__ sw, P2Comps< T > cp,
                                                          T a, int nC )
{
   const int per = ct.start[8];
   // each XCD takes a contiguous range of cells (grid padded to a multiple of 8), keeping a cell's source in one L2
   const int b = remapXcd( int( blockIdx.x ), nC * per );
   if ( b >= nC * per )
      return;
   const int c = b / per;
   const int t = b - c * per;
   int       cls = 0;
   while ( cls < 7 && t >= ct.start[cls + 1] )
      ++cls;
   p2StoredIntDispatch< T, MODE, 0, NC >( cls, c, ( t - ct.start[cls] ) * kBlock + int( threadIdx.x ), 1 << L, stride, sw, cp, a );
}

// ------------------------------------------------------------------ stored stencils (variable coefficient)
// Setup: the stencil of a point is the sum, over the micro-cells that have it as node l (on the cell's side of
// face F for a face point), of row l of sum_q k_q M_{t,q}, keyed by the offset to the other node.
// Stores the stencil of a point (p2StencilOffsets(CLS, F), entry e at out[e estride]): the sum, over the micro-cells
// that have the point as node l, of row l of sum_q k_q M_{t,q} (rows staged in LDS by p2StageRows), accumulated in
// place in the point's own output slots (one micro-cell per loop iteration, in the same order as the on-the-fly
// assembly), so no per-point stencil array is held.
template < int CLS, int F >
__device__ __forceinline__ void p2StorePoint( int c, int N, int px, int py, int pz, const double* __restrict__ kq, i64 kqStride,
                                              double* __restrict__ out, i64 estride, const double* __restrict__ Ms )
{
   constexpr P2StencilOffsets off = p2StencilOffsets( CLS, F );
#pragma unroll
   for ( int e = 0; e < off.n; ++e )
      out[i64( e ) * estride] = 0.0;
   const int          Nc = N >> 1, SA = S3i( Nc + 1 );
   const P2KQEntries& E  = kP2KQTab.c[CLS][F + 1];
#pragma unroll 1
   for ( int k = 0; k < E.n; ++k )
   {
      const P2KQEntry& e  = E.e[k];
      const int        ax = ( px - e.nd[0] ) >> 1, ay = ( py - e.nd[1] ) >> 1, az = ( pz - e.nd[2] ) >> 1;
      const double*    kv = kq + i64( c * 6 + e.t ) * SA + cidx( Nc + 1, ax, ay, az );
      double           kk[11];
#pragma unroll
      for ( int q = 0; q < 11; ++q )
         kk[q] = kv[i64( q ) * kqStride];
#pragma unroll
      for ( int jj = 0; jj < 10; ++jj )
      {
         const double* row = Ms + ( e.ord * 10 + jj ) * kP2RowPad;
         double        r   = 0.0;
#pragma unroll
         for ( int q = 0; q < 11; ++q )
            r += kk[q] * row[q];
         out[i64( e.slot[jj] ) * estride] += r;
      }
   }
}
template < int CLS >
__device__ __forceinline__ void p2StoreIntPt( int c, int s, int N, const double* __restrict__ el, const double* __restrict__ kq,
                                              i64 kqStride, double* __restrict__ sw, double* __restrict__ Ms )
{
   constexpr int kx = CLS & 1, ky = ( CLS >> 1 ) & 1, kz = ( CLS >> 2 ) & 1;
   p2StageRows< double, CLS >( el, c, Ms );
   const int M = p2InteriorExtent( N, CLS );
   if ( M < 0 || s >= S3i( M + 1 ) )
      return;
   int qx, qy, qz;
   tetXYZ( M + 1, s, qx, qy, qz );
   // lattice point and its position (p - CLS) / 2 in the class sub-lattice
   const int px = 2 * ( qx + 1 - kx ) + kx, py = 2 * ( qy + 1 - ky ) + ky, pz = 2 * ( qz + 1 - kz ) + kz;
   const int SW = p2SubWidth( N, CLS );
   double*   o  = sw + i64( c ) * p2StoredCellSize( N ) + p2StoredIntBase( N, CLS ) + cidx( SW, px >> 1, py >> 1, pz >> 1 );
   p2StorePoint< CLS, -1 >( c, N, px, py, pz, kq, kqStride, o, S3i( SW ), Ms );
}
template < int CLS >
__device__ __forceinline__ void p2StoreIntDispatch( int cls, int c, int s, int N, const double* el, const double* kq, i64 kqStride,
                                                    double* sw, double* Ms )
{
   if ( cls == CLS )
      p2StoreIntPt< CLS >( c, s, N, el, kq, kqStride, sw, Ms );
   else if constexpr ( CLS < 7 )
      p2StoreIntDispatch< CLS + 1 >( cls, c, s, N, el, kq, kqStride, sw, Ms );
}
__global__ __launch_bounds__( kBlock ) void kP2StoreInt( int L, P2ClassTiles ct, const double* __restrict__ el,
                                                         const double* __restrict__ kq, i64 kqStride, double* __restrict__ sw )
{
   __shared__ double Ms[kP2MaxPairs * 10 * kP2RowPad];
   const int         per = ct.start[8];
   const int         c   = int( blockIdx.x ) / per;
   const int         t   = int( blockIdx.x ) - c * per;
   int               cls = 0;
   while ( cls < 7 && t >= ct.start[cls + 1] )
      ++cls;
   p2StoreIntDispatch< 0 >( cls, c, ( t - ct.start[cls] ) * kBlock + int( threadIdx.x ), 1 << L, el, kq, kqStride, sw, Ms );
}
template < int F, int KU, int KV >
__device__ __forceinline__ void p2StoreFacePt( int c, int s, int N, const double* __restrict__ el, const double* __restrict__ kq,
                                               i64 kqStride, double* __restrict__ sw, double* __restrict__ Ms )
{
   p2StageRows< double, p2FaceClass( F, KU, KV ) >( el, c, Ms );
   const int M = ( N - 5 + KU + KV ) / 2;
   if ( s >= triCount( M + 1 ) )
      return;
   int i, j;
   planeXY( M + 1, s, i, j );
   const int u = 2 * i + 2 - KU, v = 2 * j + 2 - KV;
   const int px = F == 2 ? 0 : u, py = F == 0 ? v : ( F == 2 ? u : ( F == 3 ? N - u - v : 0 ) ), pz = F == 0 ? 0 : v;
   double*   o  = sw + i64( c ) * p2StoredCellSize( N ) + p2StoredFaceBase( N, F, KU | ( KV << 1 ) ) + s;
   p2StorePoint< p2FaceClass( F, KU, KV ), F >( c, N, px, py, pz, kq, kqStride, o, triCount( M + 1 ), Ms );
}
__global__ __launch_bounds__( kBlock ) void kP2StoreFace( int L, P2FaceTiles ft, const double* __restrict__ el,
                                                          const double* __restrict__ kq, i64 kqStride, double* __restrict__ sw )
{
   const int per = ft.start[4];
   const int c   = int( blockIdx.x ) / ( 4 * per );
   int       r   = int( blockIdx.x ) - c * 4 * per;
   const int f   = r / per;
   r -= f * per;
   int sub = 0;
   while ( sub < 3 && r >= ft.start[sub + 1] )
      ++sub;
   const int s = ( r - ft.start[sub] ) * kBlock + int( threadIdx.x );
   const int N = 1 << L;
   __shared__ double Ms[kP2MaxPairs * 10 * kP2RowPad];
#define HYG_P2_STORE_FACE( F, SUB ) \
   case 4 * F + SUB: p2StoreFacePt< F, ( SUB & 1 ), ( SUB >> 1 ) >( c, s, N, el, kq, kqStride, sw, Ms ); break;
   switch ( 4 * f + sub )
   {
      HYG_P2_STORE_FACE( 0, 0 )
      HYG_P2_STORE_FACE( 0, 1 )
      HYG_P2_STORE_FACE( 0, 2 )
      HYG_P2_STORE_FACE( 0, 3 )
      HYG_P2_STORE_FACE( 1, 0 )
      HYG_P2_STORE_FACE( 1, 1 )
      HYG_P2_STORE_FACE( 1, 2 )
      HYG_P2_STORE_FACE( 1, 3 )
      HYG_P2_STORE_FACE( 2, 0 )
      HYG_P2_STORE_FACE( 2, 1 )
      HYG_P2_STORE_FACE( 2, 2 )
      HYG_P2_STORE_FACE( 2, 3 )
      HYG_P2_STORE_FACE( 3, 0 )
      HYG_P2_STORE_FACE( 3, 1 )
      HYG_P2_STORE_FACE( 3, 2 )
   default: p2StoreFacePt< 3, 1, 1 >( c, s, N, el, kq, kqStride, sw, Ms ); break;
   }
#undef HYG_P2_STORE_FACE
}


// ------------------------------------------------------------------ prolongation
// Value at point p (fine lattice, N = 2^Lf) of the quadratic interpolant of the coarse function cs
// (lattice Lf-1, w = Nf/2 + 1) on a micro-cell of level Lf-2 of the macro-cell containing p.
template < typename T >
__device__ __forceinline__ T p2Interp( const P2Tables* __restrict__ tb, int Nf, const T* __restrict__ cs, int px, int py, int pz )
{
   const int wc = ( Nf >> 1 ) + 1;
   if ( ( ( px | py | pz ) & 1 ) == 0 )
      return cs[cidx( wc, px >> 1, py >> 1, pz >> 1 )];
   const int Ncc = Nf >> 2;
   const int r   = ( px & 3 ) + 4 * ( py & 3 ) + 16 * ( pz & 3 );
   const int qx = px >> 2, qy = py >> 2, qz = pz >> 2;
   const int nc  = tb->nCand[r];
   for ( int k = 0; k < nc; ++k )
   {
      const P2Tables::Cand& cd = tb->cand[r][k];
      const int             ax = qx - ( cd.o & 1 ), ay = qy - ( ( cd.o >> 1 ) & 1 ), az = qz - ( ( cd.o >> 2 ) & 1 );
      if ( ax < 0 || ay < 0 || az < 0 || ax + ay + az + kMCMax[cd.t] > Ncc )
         continue;
      T v = T( 0 );
#pragma unroll
      for ( int j = 0; j < 10; ++j )
         if ( cd.phi[j] != 0.0 )
            v += T( cd.phi[j] ) * cs[cidx( wc, 2 * ax + tb->nd[cd.t][j][0], 2 * ay + tb->nd[cd.t][j][1], 2 * az + tb->nd[cd.t][j][2] )];
      return v;
   }
   return T( 0 );
}

// points of macro-faces, -edges and -vertices: interpolated in the first incident cell, written to every replica
template < typename T >
__global__ __launch_bounds__( kBlock ) void kP2ProlongLower( const LowerChunk* __restrict__ chunks, int nChunks,
                                                             const LowerPrim* __restrict__ prims,
                                                             const LowerEntry* __restrict__ entries, const P2Tables* __restrict__ tb,
                                                             int Lf, i64 strideF, i64 strideC, int flag, T* __restrict__ fine,
                                                             const T* __restrict__ coarse, int add )
{
   const int ci = int( blockIdx.x ) * ( kBlock / 64 ) + int( threadIdx.x >> 6 );
   if ( ci >= nChunks )
      return;
   const int        lane = int( threadIdx.x & 63 );
   const LowerChunk ch   = chunks[ci];
   const LowerPrim  p    = prims[ch.prim];
   if ( !( p.bc & flag ) || lane >= ch.n )
      return;
   const int N = 1 << Lf, w = N + 1;
   const int i = ch.i0 + lane;
   int       qx, qy, qz;
   const LowerEntry e0 = entries[p.entry0];
   entryPos( e0, N, i, ch.j, qx, qy, qz );
   T v = p2Interp< T >( tb, N, coarse + i64( e0.cell ) * strideC, qx, qy, qz );
   if ( add )
      v += fine[i64( e0.cell ) * strideF + cidx( w, qx, qy, qz )];
   for ( int k = 0; k < p.nCells; ++k )
   {
      const LowerEntry e = entries[p.entry0 + k];
      entryPos( e, N, i, ch.j, qx, qy, qz );
      fine[i64( e.cell ) * strideF + cidx( w, qx, qy, qz )] = v;
   }
}

// ------------------------------------------------------------------ restriction (transpose of the prolongation)
// coarse points of macro-faces, -edges and -vertices: sum over the incident cells of the in-cell support,
// fine points on macro-primitives scaled by 1 / (#cells sharing them); written to every replica
template < typename T >
__global__ __launch_bounds__( kBlock ) void kP2RestrictLower( const LowerChunk* __restrict__ chunks, int nChunks,
                                                              const LowerPrim* __restrict__ prims,
                                                              const LowerEntry* __restrict__ entries, const CellInfo* __restrict__ cinfo,
                                                              const P2Tables* __restrict__ tb, int Lc, i64 strideC, i64 strideF,
                                                              int flag, T* __restrict__ coarse, const T* __restrict__ fine )
{
   const int ci = int( blockIdx.x ) * ( kBlock / 64 ) + int( threadIdx.x >> 6 );
   if ( ci >= nChunks )
      return;
   const int        lane = int( threadIdx.x & 63 );
   const LowerChunk ch   = chunks[ci];
   const LowerPrim  p    = prims[ch.prim];
   if ( !( p.bc & flag ) )
      return;
   const int Nc = 1 << Lc, wc = Nc + 1, Nf = 2 * Nc, wf = Nf + 1;
   // G lanes per coarse point (the largest power of two with G n <= 64 for the chunk's n points), each summing every
   // G-th support point: a macro-vertex point gathers from every incident cell, which one lane would serialise
   int G = 1;
   while ( G < 64 && 2 * G * ch.n <= 64 )
      G *= 2;
   for ( int base = 0; base < ch.n; base += 64 / G )
   {
      const int  pi  = base + lane / G, sub = lane % G;
      const bool ok  = pi < ch.n;
      const int  i   = ch.i0 + pi;
      T          v   = T( 0 );
      for ( int k = 0; ok && k < p.nCells; ++k )
      {
         const LowerEntry e = entries[p.entry0 + k];
         int              nx, ny, nz;
         entryPos( e, Nc, i, ch.j, nx, ny, nz );
         const int       cls = ( nx & 1 ) | ( ( ny & 1 ) << 1 ) | ( ( nz & 1 ) << 2 );
         const T*        fs  = fine + i64( e.cell ) * strideF;
         const CellInfo& cf  = cinfo[e.cell];
         const int       ns  = tb->nSup[cls];
         for ( int m = sub; m < ns; m += G )
         {
            const P2Tables::Sup& sp = tb->sup[cls][m];
            const int            fx = 2 * nx + sp.d[0], fy = 2 * ny + sp.d[1], fz = 2 * nz + sp.d[2];
            if ( !inCell( Nf, fx, fy, fz ) )
               continue;
            v += T( sp.w * cf.invNC[classifySlot( Nf, fx, fy, fz )] ) * fs[cidx( wf, fx, fy, fz )];
         }
      }
      for ( int o = 1; o < G; o <<= 1 )
         v += __shfl_xor( v, o );
      if ( !ok || sub != 0 )
         continue;
      for ( int k = 0; k < p.nCells; ++k )
      {
         const LowerEntry e = entries[p.entry0 + k];
         int              nx, ny, nz;
         entryPos( e, Nc, i, ch.j, nx, ny, nz );
         coarse[i64( e.cell ) * strideC + cidx( wc, nx, ny, nz )] = v;
      }
   }
}

// ------------------------------------------------------------------ launch
int p2Config( const LevelGeom& g, int dtype, int mode )
{
   const int c = g.p2Cfg[dtype == 8 ? 0 : 1][mode];
   return c >= 0 ? c : p2DefaultConfig( 1 << g.L );
}
void setP2Config( LevelGeom& g, int dtype, int mode, int cfg ) { g.p2Cfg[dtype == 8 ? 0 : 1][mode] = cfg; }

// element-wise coefficient values per quadrature point are kq[q kqStride + element]; per-component partial sums of
// the variable-coefficient operator lie partStride / lowStride apart
static i64 p2KqStride( const MeshCtx& m, int L ) { return i64( m.mesh->nC() ) * 6 * S3i( ( 1 << ( L - 1 ) ) + 1 ); }
static i64 p2PartStride( const MeshCtx& m, int L ) { return i64( m.mesh->nF() ) * 2 * i64( ( 1 << L ) + 1 ) * ( ( 1 << L ) + 1 ); }
static i64 p2LowStride( const MeshCtx& m, int L ) { return i64( m.entriesHost.size() ) * ( ( 1 << L ) + 1 ); }

template < typename T, int MODE >
static void p2Combines( const MeshCtx& m, const LevelGeom& g, int L, const StencilCall& c, T* part, T* low, const T* dv, hipStream_t s )
{
   const int nLow = g.firstFaceChunk, nFace = g.nChunks - g.firstFaceChunk;
   if ( nFace > 0 )
      hipLaunchKernelGGL( ( kFaceCombine< T, MODE > ), dim3( xcdGrid( ceilDiv( nFace, kBlock / 64 ) ) ), dim3( kBlock ), 0, s,
                          g.chunks + g.firstFaceChunk, nFace, m.prims, m.entries, m.mesh->nV() + m.mesh->nE(), part,
                          (const T*) nullptr, dv, L, g.stride, c.flag, (const T*) c.x, (const T*) c.in1, (T*) c.out0,
                          (T*) c.out1, T( c.a ) );
   if ( nLow > 0 )
      hipLaunchKernelGGL( ( kLowCombine< T, MODE > ), dim3( ceilDiv( nLow, kBlock / 64 ) ), dim3( kBlock ), 0, s, g.chunks, nLow,
                          m.prims, m.entries, low, dv, L, g.stride, c.flag, (const T*) c.x, (const T*) c.in1, (T*) c.out0,
                          (T*) c.out1, T( c.a ) );
}

// variable coefficient assembled on the fly for n = 1 or 3 components (one assembly per point for all of them)
template < typename T, int MODE >
static void p2KQOnTheFly( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, const StencilCall* cs, int n, hipStream_t s )
{
   constexpr bool         dbl  = std::is_same< T, double >::value;
   const hyg_op_s::Level& ol   = op.levels[L];
   const int              nC   = m.mesh->nC(), N = 1 << L;
   const T*               el   = dbl ? (const T*) ol.p2el : (const T*) ol.p2elF;
   const T*               kq   = dbl ? (const T*) ol.p2kq : (const T*) ol.p2kqF;
   T*                     part = dbl ? (T*) ol.facePart : (T*) ol.facePartF;
   T*                     low  = dbl ? (T*) ol.lowPart : (T*) ol.lowPartF;
   const T*               dv   = dbl ? (const T*) ol.dinv : (const T*) ol.dinvF;
   const i64              kqs = p2KqStride( m, L ), ps = p2PartStride( m, L ), ls = p2LowStride( m, L );
   const int              flag = cs[0].flag;
   P2Comps< T >           cp{};
   for ( int k = 0; k < n; ++k )
   {
      cp.x[k]    = (const T*) cs[k].x;
      cp.in1[k]  = (const T*) cs[k].in1;
      cp.out0[k] = (T*) cs[k].out0;
      cp.out1[k] = (T*) cs[k].out1;
   }
   launchP2KQPoints< T >( L, nC, g.stride, el, kq, kqs, dv, cp, n, MODE, T( cs[0].a ), m.cellFaces, flag, part, ps, s );
   const int tB = ceilDiv( i64( p2EdgeVertexCount( N ) ) * kP2EvLanes, kBlock );
   for ( int k = 0; k < n; ++k )
   {
      const StencilCall& c = cs[k];
      hipLaunchKernelGGL( ( kP2Apply< T, MODE, false, true, true > ), dim3( nC * tB ), dim3( kBlock ), 0, s, L, tB, g.stride, el, kq, kqs,
                          m.p2tab, m.cellInfo, m.cellFaces, m.prims, m.entries, part + k * ps, low + k * ls, dv, c.flag,
                          (const T*) c.x, (const T*) c.in1, (T*) c.out0, (T*) c.out1, T( c.a ) );
   }
   for ( int k = 0; k < n; ++k )
      p2Combines< T, MODE >( m, g, L, cs[k], part + k * ps, low + k * ls, dv, s );
}

template < typename T, int MODE, bool DIAG, bool KQ >
static void p2TK( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, const StencilCall& c, hipStream_t s,
                  bool skipInterior = false )
{
   if constexpr ( KQ && !DIAG )
   {
      if ( !op.levels[L].p2s )
      {
         p2KQOnTheFly< T, MODE >( m, g, op, L, &c, 1, s );
         return;
      }
   }
   constexpr bool         dbl  = std::is_same< T, double >::value;
   const hyg_op_s::Level& ol   = op.levels[L];
   const int              nC   = m.mesh->nC();
   const int              w    = ( 1 << L ) + 1;
   const T*               el   = dbl ? (const T*) ol.p2el : (const T*) ol.p2elF;
   const T*               kq   = dbl ? (const T*) ol.p2kq : (const T*) ol.p2kqF;
   const i64              kqs  = KQ ? p2KqStride( m, L ) : 0;
   T*                     part = dbl ? (T*) ol.facePart : (T*) ol.facePartF;
   T*                     low  = dbl ? (T*) ol.lowPart : (T*) ol.lowPartF;
   const T*               dv   = DIAG ? nullptr : ( dbl ? (const T*) ol.dinv : (const T*) ol.dinvF );
   const int              tA   = ceilDiv( S3i( w ), kBlock );
   if constexpr ( !KQ && !DIAG )
   {
      // constant coefficient: cell interiors by row pairs, cell faces by one-sided stencils, the points on cell
      // edges and vertices element by element
      const T*  stw  = dbl ? (const T*) ol.p2st : (const T*) ol.p2stF;
      const int cfg  = p2Config( g, dbl ? 8 : 4, MODE );
      const int ci   = cfg / 2, nSeg = g.nP2ZSegs[ci];
      const dim3 grid( xcdGrid( nC * ceilDiv( nSeg, kBlock / 64 ) ) );
      if ( ( c.flag & 1 ) && nSeg > 0 )
      {
         if ( cfg & 1 )
            hipLaunchKernelGGL( ( kP2Rows< T, MODE, 2 > ), grid, dim3( kBlock ), 0, s, L, nC, g.stride, g.p2ZSegs[ci], nSeg, stw,
                                (const T*) c.x, (const T*) c.in1, (T*) c.out0, (T*) c.out1, T( c.a ) );
         else
            hipLaunchKernelGGL( ( kP2Rows< T, MODE, 1 > ), grid, dim3( kBlock ), 0, s, L, nC, g.stride, g.p2ZSegs[ci], nSeg, stw,
                                (const T*) c.x, (const T*) c.in1, (T*) c.out0, (T*) c.out1, T( c.a ) );
      }
      P2FaceTiles ft;
      ft.start[0] = 0;
      for ( int k = 0; k < 4; ++k )
         ft.start[k + 1] = ft.start[k] + ceilDiv( p2FaceSubCount( 1 << L, k & 1, k >> 1 ), kBlock );
      const T* fsw = dbl ? (const T*) ol.p2fst : (const T*) ol.p2fstF;
      if ( ft.start[4] > 0 )
         hipLaunchKernelGGL( ( kP2Face< T, false > ), dim3( nC * 4 * ft.start[4] ), dim3( kBlock ), 0, s, L, g.stride, ft, fsw, m.cellFaces,
                             c.flag, part, (const T*) c.x, 0 );
      const int tB = ceilDiv( i64( p2EdgeVertexCount( 1 << L ) ) * kP2EvLanes, kBlock );
      hipLaunchKernelGGL( ( kP2Apply< T, MODE, DIAG, KQ, true > ), dim3( nC * tB ), dim3( kBlock ), 0, s, L, tB, g.stride, el, kq, kqs,
                          m.p2tab, m.cellInfo, m.cellFaces, m.prims, m.entries, part, low, dv, c.flag, (const T*) c.x,
                          (const T*) c.in1, (T*) c.out0, (T*) c.out1, T( c.a ) );
   }
   else if ( KQ && !DIAG && ol.p2s )
   {
      // variable coefficient with stored stencils: the same row and face kernels with per-point weights
      const T*           sw = dbl ? (const T*) ol.p2s : (const T*) ol.p2sF;
      const P2ClassTiles ct = p2InteriorTiles( 1 << L );
      if ( ct.start[8] > 0 && ( c.flag & 1 ) && !skipInterior )
      {
         const P2Comps< T > cp{ { (const T*) c.x }, { (const T*) c.in1 }, { (T*) c.out0 }, { (T*) c.out1 } };
         hipLaunchKernelGGL( ( kP2StoredInt< T, MODE, 1 > ), dim3( xcdGrid( nC * ct.start[8] ) ), dim3( kBlock ), 0, s, L, g.stride, ct, sw,
                             cp, T( c.a ), nC );
      }
      P2FaceTiles ft;
      ft.start[0] = 0;
      for ( int k = 0; k < 4; ++k )
         ft.start[k + 1] = ft.start[k] + ceilDiv( p2FaceSubCount( 1 << L, k & 1, k >> 1 ), kBlock );
      if ( ft.start[4] > 0 )
         hipLaunchKernelGGL( ( kP2Face< T, true > ), dim3( xcdGrid( nC * 4 * ft.start[4] ) ), dim3( kBlock ), 0, s, L, g.stride, ft, sw,
                             m.cellFaces, c.flag, part, (const T*) c.x, nC );
      const int tB = ceilDiv( i64( p2EdgeVertexCount( 1 << L ) ) * kP2EvLanes, kBlock );
      hipLaunchKernelGGL( ( kP2Apply< T, MODE, DIAG, KQ, true > ), dim3( nC * tB ), dim3( kBlock ), 0, s, L, tB, g.stride, el, kq, kqs,
                          m.p2tab, m.cellInfo, m.cellFaces, m.prims, m.entries, part, low, dv, c.flag, (const T*) c.x,
                          (const T*) c.in1, (T*) c.out0, (T*) c.out1, T( c.a ) );
   }
   else
   hipLaunchKernelGGL( ( kP2Apply< T, MODE, DIAG, KQ > ), dim3( nC * tA ), dim3( kBlock ), 0, s, L, tA, g.stride, el, kq, kqs, m.p2tab,
                       m.cellInfo, m.cellFaces, m.prims, m.entries, part, low, dv, c.flag, (const T*) c.x, (const T*) c.in1,
                       (T*) c.out0, (T*) c.out1, T( c.a ) );
   p2Combines< T, MODE >( m, g, L, c, part, low, dv, s );
}

template < typename T, int MODE, bool DIAG >
static void p2T( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, const StencilCall& c, hipStream_t s )
{
   if ( op.kind == 3 )
      p2TK< T, MODE, DIAG, true >( m, g, op, L, c, s );
   else
      p2TK< T, MODE, DIAG, false >( m, g, op, L, c, s );
}

template < typename T >
static void p2Mode( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, const StencilCall& c, hipStream_t s )
{
   switch ( c.mode )
   {
   case MODE_APPLY:
      p2T< T, MODE_APPLY, false >( m, g, op, L, c, s );
      break;
   case MODE_APPLY_ADD:
      p2T< T, MODE_APPLY_ADD, false >( m, g, op, L, c, s );
      break;
   case MODE_RESIDUAL:
      p2T< T, MODE_RESIDUAL, false >( m, g, op, L, c, s );
      break;
   case MODE_JACOBI:
      p2T< T, MODE_JACOBI, false >( m, g, op, L, c, s );
      break;
   case MODE_CHEB0:
      p2T< T, MODE_CHEB0, false >( m, g, op, L, c, s );
      break;
   default:
      p2T< T, MODE_CHEBK, false >( m, g, op, L, c, s );
      break;
   }
}

void launchP2( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, int dtype, const StencilCall& c, hipStream_t s )
{
   if ( dtype == 8 )
      p2Mode< double >( m, g, op, L, c, s );
   else
      p2Mode< float >( m, g, op, L, c, s );
}

// Apply (replace or add) to up to three functions at once. With stored stencils the cell interiors of all
// components go through one pass that reads every weight once; the face, edge and vertex points and the
// combine run per component. Other operators apply component by component.
template < typename T, int MODE >
static void p2VectorT( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, const StencilCall* cs, int n, hipStream_t s )
{
   constexpr bool         dbl = std::is_same< T, double >::value;
   const hyg_op_s::Level& ol  = op.levels[L];
   const P2ClassTiles     ct  = p2InteriorTiles( 1 << L );
   if ( ct.start[8] > 0 && ( cs[0].flag & 1 ) )
   {
      P2Comps< T > cp{};
      for ( int k = 0; k < n; ++k )
      {
         cp.x[k]    = (const T*) cs[k].x;
         cp.out0[k] = (T*) cs[k].out0;
      }
      const T*  sw = dbl ? (const T*) ol.p2s : (const T*) ol.p2sF;
      const int nC = m.mesh->nC();
      if ( n == 3 )
         hipLaunchKernelGGL( ( kP2StoredInt< T, MODE, 3 > ), dim3( xcdGrid( nC * ct.start[8] ) ), dim3( kBlock ), 0, s, L, g.stride, ct, sw, cp,
                             T( 0 ), nC );
      else if ( n == 2 )
         hipLaunchKernelGGL( ( kP2StoredInt< T, MODE, 2 > ), dim3( xcdGrid( nC * ct.start[8] ) ), dim3( kBlock ), 0, s, L, g.stride, ct, sw, cp,
                             T( 0 ), nC );
      else
         hipLaunchKernelGGL( ( kP2StoredInt< T, MODE, 1 > ), dim3( xcdGrid( nC * ct.start[8] ) ), dim3( kBlock ), 0, s, L, g.stride, ct, sw, cp,
                             T( 0 ), nC );
   }
   for ( int k = 0; k < n; ++k )
      p2TK< T, MODE, false, true >( m, g, op, L, cs[k], s, true );
}

void launchP2Vector( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, int dtype, const StencilCall* cs, int n, hipStream_t s )
{
   if ( op.kind == 3 && !op.levels[L].p2s && n == 3 && ( cs[0].mode == MODE_APPLY || cs[0].mode == MODE_APPLY_ADD ) )
   {
      // one on-the-fly assembly per point for the three components
      if ( dtype == 8 )
         cs[0].mode == MODE_APPLY ? p2KQOnTheFly< double, MODE_APPLY >( m, g, op, L, cs, n, s )
                                  : p2KQOnTheFly< double, MODE_APPLY_ADD >( m, g, op, L, cs, n, s );
      else
         cs[0].mode == MODE_APPLY ? p2KQOnTheFly< float, MODE_APPLY >( m, g, op, L, cs, n, s )
                                  : p2KQOnTheFly< float, MODE_APPLY_ADD >( m, g, op, L, cs, n, s );
      return;
   }
   const bool fused = op.kind == 3 && op.levels[L].p2s && n >= 1 && n <= 3 &&
                      ( cs[0].mode == MODE_APPLY || cs[0].mode == MODE_APPLY_ADD );
   if ( !fused )
   {
      for ( int k = 0; k < n; ++k )
         launchP2( m, g, op, L, dtype, cs[k], s );
      return;
   }
   if ( dtype == 8 )
      cs[0].mode == MODE_APPLY ? p2VectorT< double, MODE_APPLY >( m, g, op, L, cs, n, s ) : p2VectorT< double, MODE_APPLY_ADD >( m, g, op, L, cs, n, s );
   else
      cs[0].mode == MODE_APPLY ? p2VectorT< float, MODE_APPLY >( m, g, op, L, cs, n, s ) : p2VectorT< float, MODE_APPLY_ADD >( m, g, op, L, cs, n, s );
}

// coefficient values of every micro-cell at its 11 quadrature points (variable-coefficient P2 operator)
void launchP2Coefficients( const MeshCtx& m, const hyg_op_s& op, int L, hipStream_t s )
{
   const hyg_op_s::Level& ol = op.levels[L];
   const int              nC = m.mesh->nC();
   const int              tA = ceilDiv( S3i( ( 1 << ( L - 1 ) ) + 1 ), kBlock );
   const P2Coef           k{ op.p2k.kind, op.p2k.k0, op.p2k.alpha, op.p2k.beta, op.p2k.seed };
   hipLaunchKernelGGL( ( kP2KQ< double > ), dim3( nC * tA ), dim3( kBlock ), 0, s, L, tA, ol.elem, k, ol.p2kq, p2KqStride( m, L ) );
   hipLaunchKernelGGL( ( kP2KQ< float > ), dim3( nC * tA ), dim3( kBlock ), 0, s, L, tA, ol.elem, k, ol.p2kqF, p2KqStride( m, L ) );
}

// stored stencils of the variable-coefficient P2 operator at lattice level L (FP64, and an FP32 copy)
void launchP2StoreStencils( const MeshCtx& m, hyg_op_s& op, int L, hipStream_t s )
{
   hyg_op_s::Level& ol = op.levels[L];
   const int        nC = m.mesh->nC(), N = 1 << L;
   const i64        n  = i64( nC ) * p2StoredCellSize( N );
   if ( !ol.p2s )
   {
      if ( hipMalloc( &ol.p2s, size_t( n ) * sizeof( double ) ) != hipSuccess ||
           hipMalloc( &ol.p2sF, size_t( n ) * sizeof( float ) ) != hipSuccess )
         throw std::runtime_error( "hipMalloc failed for the stored P2 stencils" );
   }
   // positions of a class sub-lattice that are no cell-interior points keep weight 0
   (void) hipMemsetAsync( ol.p2s, 0, size_t( n ) * sizeof( double ), s );
   P2ClassTiles ct;
   ct.start[0] = 0;
   for ( int k = 0; k < 8; ++k )
   {
      const int M     = p2InteriorExtent( N, k );
      ct.start[k + 1] = ct.start[k] + ( M >= 0 ? ceilDiv( S3i( M + 1 ), kBlock ) : 0 );
   }
   if ( ct.start[8] > 0 )
      hipLaunchKernelGGL( kP2StoreInt, dim3( nC * ct.start[8] ), dim3( kBlock ), 0, s, L, ct, ol.p2el, ol.p2kq, p2KqStride( m, L ), ol.p2s );
   P2FaceTiles ft;
   ft.start[0] = 0;
   for ( int k = 0; k < 4; ++k )
      ft.start[k + 1] = ft.start[k] + ceilDiv( p2FaceSubCount( N, k & 1, k >> 1 ), kBlock );
   if ( ft.start[4] > 0 )
      hipLaunchKernelGGL( kP2StoreFace, dim3( nC * 4 * ft.start[4] ), dim3( kBlock ), 0, s, L, ft, ol.p2el, ol.p2kq, p2KqStride( m, L ), ol.p2s );
   launchConvert( n, 4, ol.p2sF, 8, ol.p2s, s );
}

// inverse diagonal of a P2 operator at lattice level L into its per-point arrays (FP64 and FP32 copy)
void launchP2InverseDiagonal( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, hipStream_t s )
{
   const hyg_op_s::Level& ol = op.levels[L];
   const i64              n  = i64( m.mesh->nC() ) * g.stride + kTailSlack;
   (void) hipMemsetAsync( ol.dinv, 0, size_t( n ) * sizeof( double ), s );
   StencilCall c;
   c.mode = MODE_APPLY;
   c.flag = 15;
   c.out0 = ol.dinv;
   p2T< double, MODE_APPLY, true >( m, g, op, L, c, s );
   hipLaunchKernelGGL( ( kInvertAll< double > ), dim3( 1024 ), dim3( kBlock ), 0, s, n, ol.dinv );
   launchConvert( n, 4, ol.dinvF, 8, ol.dinv, s );
}

template < typename T >
static void p2ProlongT( const MeshCtx& m, const LevelGeom& gf, const P2Tables* tb, int Lf, void* fine, const void* coarse, int flag,
                        bool add, hipStream_t st )
{
   const int    nC      = m.mesh->nC();
   const i64 strideC = cellStride( Lf - 1 );
   const int W       = ( 1 << Lf ) - 3;
   if ( W > 0 )
      hipLaunchKernelGGL( ( kP2ProlongRows< T > ), dim3( xcdGrid( nC * ceilDiv( i64( W ) * ( W + 1 ) / 2, kBlock / 64 ) ) ), dim3( kBlock ),
                          0, st, Lf, nC, gf.stride, strideC, flag, (T*) fine, (const T*) coarse, add ? 1 : 0 );
   if ( gf.nChunks > 0 )
      hipLaunchKernelGGL( ( kP2ProlongLower< T > ), dim3( ceilDiv( gf.nChunks, kBlock / 64 ) ), dim3( kBlock ), 0, st, gf.chunks,
                          gf.nChunks, m.prims, m.entries, tb, Lf, gf.stride, strideC, flag, (T*) fine, (const T*) coarse,
                          add ? 1 : 0 );
}

void launchP2Prolongate( const MeshCtx& m, const LevelGeom& gf, const P2Tables* tb, int Lf, int dtype, void* fine,
                         const void* coarse, int flag, bool add, hipStream_t st )
{
   if ( dtype == 8 )
      p2ProlongT< double >( m, gf, tb, Lf, fine, coarse, flag, add, st );
   else
      p2ProlongT< float >( m, gf, tb, Lf, fine, coarse, flag, add, st );
}

template < typename T >
static void p2RestrictT( const MeshCtx& m, const LevelGeom& gc, const P2Tables* tb, int Lc, void* coarse, const void* fine, int flag,
                         hipStream_t st )
{
   const int    nC      = m.mesh->nC();
   const i64    strideF = cellStride( Lc + 1 );
   P2ClassTiles ct;
   ct.start[0] = 0;
   for ( int k = 0; k < 8; ++k )
   {
      const int M     = p2InteriorExtent( 1 << Lc, k );
      ct.start[k + 1] = ct.start[k] + ( M >= 0 ? ceilDiv( S3i( M + 1 ), kBlock ) : 0 );
   }
   // the cell interiors take no flag, as in P2toP2QuadraticRestriction.cpp:146 (every cell, whatever the flag)
   if ( ct.start[8] > 0 )
      hipLaunchKernelGGL( ( kP2RestrictInt< T > ), dim3( nC * ct.start[8] ), dim3( kBlock ), 0, st, Lc, gc.stride, strideF, ct, (T*) coarse,
                          (const T*) fine );
   if ( gc.nChunks > 0 )
      hipLaunchKernelGGL( ( kP2RestrictLower< T > ), dim3( ceilDiv( gc.nChunks, kBlock / 64 ) ), dim3( kBlock ), 0, st, gc.chunks,
                          gc.nChunks, m.prims, m.entries, m.cellInfo, tb, Lc, gc.stride, strideF, flag, (T*) coarse,
                          (const T*) fine );
}

void launchP2Restrict( const MeshCtx& m, const LevelGeom& gc, const P2Tables* tb, int Lc, int dtype, void* coarse, const void* fine,
                       int flag, hipStream_t st )
{
   if ( dtype == 8 )
      p2RestrictT< double >( m, gc, tb, Lc, coarse, fine, flag, st );
   else
      p2RestrictT< float >( m, gc, tb, Lc, coarse, fine, flag, st );
}

} // namespace hyg
