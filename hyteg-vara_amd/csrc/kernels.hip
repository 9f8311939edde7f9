// gfx950 kernels of the P1 matrix-free multigrid path.
//
// Work decomposition. Cell arrays are processed in plane tiles: 256 consecutive values of one
// z-plane (contiguous in memory), one thread per value, so loads and stores of a wave are
// contiguous. The tile list is z-major within a cell and a block's logical index is remapped
// so that each of the 8 XCDs walks a contiguous range of (cell, plane) tiles: the planes
// z-1, z, z+1 read by a stencil tile were just streamed through the same XCD's L2.
// Macro-face, -edge and -vertex points are processed in chunks of up to 64 points of one
// primitive (one wavefront per chunk); their results are written to every incident cell.
#include "kernels.hpp"

#include <algorithm>
#include <type_traits>

namespace hyg {

constexpr int kBlock = 256;

inline int ceilDiv( i64 a, i64 b ) { return int( ( a + b - 1 ) / b ); }
inline int xcdGrid( int nb ) { return ( ( nb + 7 ) / 8 ) * 8; }

// hardware dispatches consecutive workgroups round-robin over the 8 XCDs
__device__ __forceinline__ int remapXcd( int b, int nb )
{
   const int per = ( nb + 7 ) >> 3;
   return ( b & 7 ) * per + ( b >> 3 );
}

// (x, y) of the s-th point of a plane whose row 0 holds W points (row y holds W - y points)
__device__ __forceinline__ void planeXY( int W, int s, int& x, int& y )
{
   const float b  = 2.0f * float( W ) + 1.0f;
   int         yy = int( ( b - sqrtf( b * b - 8.0f * float( s ) ) ) * 0.5f );
   yy             = yy < 0 ? 0 : yy;
   if ( rowStart( W, yy + 1 ) <= s )
      ++yy;
   if ( rowStart( W, yy ) > s )
      --yy;
   y = yy;
   x = s - rowStart( W, yy );
}

__device__ __forceinline__ int planeStart( int w, int z ) { return int( S3( w ) - S3( w - z ) ); }

// cell coordinates of point (i, j) of a lower-dimensional primitive in an incident cell
__device__ __forceinline__ void entryPos( const LowerEntry& e, int N, int i, int j, int& qx, int& qy, int& qz )
{
   const int u0x = e.l0 == 1, u0y = e.l0 == 2, u0z = e.l0 == 3;
   const int u1x = e.l1 == 1, u1y = e.l1 == 2, u1z = e.l1 == 3;
   const int u2x = e.l2 == 1, u2y = e.l2 == 2, u2z = e.l2 == 3;
   qx = N * u0x + i * ( u1x - u0x ) + j * ( u2x - u0x );
   qy = N * u0y + i * ( u1y - u0y ) + j * ( u2y - u0y );
   qz = N * u0z + i * ( u1z - u0z ) + j * ( u2z - u0z );
}

__device__ __forceinline__ int slotType( const CellInfo& ci, int slot ) { return int( ( ci.bcword >> ( 4 * slot ) ) & 15u ); }

// ------------------------------------------------------------------ stencil epilogue
template < typename T, int MODE >
__device__ __forceinline__ void epilogue( T Ax, T xc, T in1v, T out0v, T dinv, T a, T& o0, T& o1 )
{
   if constexpr ( MODE == MODE_APPLY )
      o0 = Ax;
   else if constexpr ( MODE == MODE_APPLY_ADD )
      o0 = out0v + Ax;
   else if constexpr ( MODE == MODE_RESIDUAL )
      o0 = in1v - Ax;
   else if constexpr ( MODE == MODE_JACOBI )
      o0 = xc + a * dinv * ( in1v - Ax );
   else if constexpr ( MODE == MODE_CHEB0 )
   {
      const T t = dinv * ( in1v - Ax );
      o0        = t;
      o1        = xc + a * t;
   }
   else
   {
      const T t = dinv * Ax;
      o0        = t;
      o1        = in1v + a * t;
   }
}

// ------------------------------------------------------------------ cell-interior 15-point stencil
template < typename T, int MODE >
__global__ __launch_bounds__( kBlock ) void kCellStencil( const PlaneTile* __restrict__ tiles, int tile0, int nTiles, int nC,
                                                          int L, i64 stride, const T* __restrict__ cellSt,
                                                          const T* __restrict__ x, const T* __restrict__ in1,
                                                          T* __restrict__ out0, T* __restrict__ out1, T a )
{
   const int nb = nTiles * nC;
   const int lb = remapXcd( blockIdx.x, nb );
   if ( lb >= nb )
      return;
   const int       c    = lb / nTiles;
   const PlaneTile tile = tiles[tile0 + ( lb - c * nTiles )];
   const int       N = 1 << L, w = N + 1, z = tile.z, W = w - z;
   const int       s = tile.s0 + int( threadIdx.x );
   if ( s >= ( W * ( W + 1 ) ) / 2 )
      return;
   int px, py;
   planeXY( W, s, px, py );
   if ( px < 1 || py < 1 || px + py + z > N - 1 )
      return;
   const i64 base = i64( c ) * stride;
   const T*  xs   = x + base;
   const T*  st   = cellSt + 16 * c;

   const int pm = planeStart( w, z - 1 ), p0 = planeStart( w, z ), pp = planeStart( w, z + 1 );
   const int rB0 = pm + rowStart( W + 1, py ) + px;     // (y,   z-1)
   const int rB1 = pm + rowStart( W + 1, py + 1 ) + px; // (y+1, z-1)
   const int rS  = p0 + rowStart( W, py - 1 ) + px;     // (y-1, z)
   const int rC  = p0 + s;                              // (y,   z)
   const int rN  = p0 + rowStart( W, py + 1 ) + px;     // (y+1, z)
   const int rTS = pp + rowStart( W - 1, py - 1 ) + px; // (y-1, z+1)
   const int rT  = pp + rowStart( W - 1, py ) + px;     // (y,   z+1)

   const T xc = xs[rC];
   T       Ax = st[0] * xc;
   Ax += st[1] * xs[rC - 1];
   Ax += st[2] * xs[rC + 1];
   Ax += st[3] * xs[rN];
   Ax += st[4] * xs[rS];
   Ax += st[5] * xs[rN - 1];
   Ax += st[6] * xs[rS + 1];
   Ax += st[7] * xs[rT];
   Ax += st[8] * xs[rT - 1];
   Ax += st[9] * xs[rTS];
   Ax += st[10] * xs[rTS + 1];
   Ax += st[11] * xs[rB0];
   Ax += st[12] * xs[rB1];
   Ax += st[13] * xs[rB0 + 1];
   Ax += st[14] * xs[rB1 - 1];

   const i64 o  = base + rC;
   T         i1 = T( 0 ), ov = T( 0 ), r0, r1;
   if constexpr ( MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
      i1 = in1[o];
   if constexpr ( MODE == MODE_APPLY_ADD )
      ov = out0[o];
   epilogue< T, MODE >( Ax, xc, i1, ov, st[15], a, r0, r1 );
   out0[o] = r0;
   if constexpr ( MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
      out1[o] = r1;
}

// ------------------------------------------------------------------ macro-face / -edge / -vertex points
template < typename T, int MODE >
__global__ __launch_bounds__( kBlock ) void kLowerStencil( const LowerChunk* __restrict__ chunks, int nChunks,
                                                           const LowerPrim* __restrict__ prims,
                                                           const LowerEntry* __restrict__ entries, const T* __restrict__ entryW,
                                                           const T* __restrict__ primInv, int L, i64 stride, int flag,
                                                           const T* __restrict__ x, const T* __restrict__ in1,
                                                           T* __restrict__ out0, T* __restrict__ out1, T a )
{
   const int ci = int( blockIdx.x ) * ( kBlock / 64 ) + int( threadIdx.x >> 6 );
   if ( ci >= nChunks )
      return;
   const int        lane = int( threadIdx.x & 63 );
   const LowerChunk ch   = chunks[ci];
   if ( lane >= ch.n )
      return;
   const LowerPrim p = prims[ch.prim];
   if ( !( p.bc & flag ) )
      return;
   const int N = 1 << L, w = N + 1;
   const int i = ch.i0 + lane, j = ch.j;
   T         Ax = T( 0 );
   for ( int k = 0; k < p.nCells; ++k )
   {
      const LowerEntry e = entries[p.entry0 + k];
      int              qx, qy, qz;
      entryPos( e, N, i, j, qx, qy, qz );
      const T* xs = x + i64( e.cell ) * stride;
      const T* wv = entryW + i64( p.entry0 + k ) * 16;
#pragma unroll
      for ( int d = 0; d < NDIR; ++d )
      {
         const T wd = wv[d];
         const int ax = qx + dirX( d ), ay = qy + dirY( d ), az = qz + dirZ( d );
         if ( wd != T( 0 ) && inCell( N, ax, ay, az ) )
            Ax += wd * xs[cidx( w, ax, ay, az )];
      }
   }
   // pointwise terms from the first replica (all replicas hold the same value)
   const LowerEntry e0 = entries[p.entry0];
   int              qx, qy, qz;
   entryPos( e0, N, i, j, qx, qy, qz );
   const i64 o0i = i64( e0.cell ) * stride + cidx( w, qx, qy, qz );
   T         xc = T( 0 ), i1 = T( 0 ), ov = T( 0 ), r0, r1;
   if constexpr ( MODE == MODE_JACOBI || MODE == MODE_CHEB0 )
      xc = x[o0i];
   if constexpr ( MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
      i1 = in1[o0i];
   if constexpr ( MODE == MODE_APPLY_ADD )
      ov = out0[o0i];
   epilogue< T, MODE >( Ax, xc, i1, ov, primInv[ch.prim], a, r0, r1 );
   for ( int k = 0; k < p.nCells; ++k )
   {
      const LowerEntry e = entries[p.entry0 + k];
      entryPos( e, N, i, j, qx, qy, qz );
      const i64 o = i64( e.cell ) * stride + cidx( w, qx, qy, qz );
      out0[o]     = r0;
      if constexpr ( MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
         out1[o] = r1;
   }
}

template < typename T, int MODE >
static void stencilT( const MeshCtx& m, const LevelGeom& g, const hyg_op_s::Level& ol, int L, const StencilCall& c, hipStream_t s )
{
   const int nC = m.mesh->nC();
   const T*  cst = std::is_same< T, double >::value ? (const T*) ol.cellSt : (const T*) ol.cellStF;
   const T*  ew  = std::is_same< T, double >::value ? (const T*) ol.entryW : (const T*) ol.entryWF;
   const T*  pi  = std::is_same< T, double >::value ? (const T*) ol.primInv : (const T*) ol.primInvF;
   if ( ( c.flag & 1 ) && g.nInteriorTiles > 0 )
   {
      const int nb = g.nInteriorTiles * nC;
      hipLaunchKernelGGL( ( kCellStencil< T, MODE > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, s, g.tiles, g.interiorTile0,
                          g.nInteriorTiles, nC, L, g.stride, cst, (const T*) c.x, (const T*) c.in1, (T*) c.out0, (T*) c.out1,
                          T( c.a ) );
   }
   if ( g.nChunks > 0 )
      hipLaunchKernelGGL( ( kLowerStencil< T, MODE > ), dim3( ceilDiv( g.nChunks, kBlock / 64 ) ), dim3( kBlock ), 0, s, g.chunks,
                          g.nChunks, m.prims, m.entries, ew, pi, L, g.stride, c.flag, (const T*) c.x, (const T*) c.in1,
                          (T*) c.out0, (T*) c.out1, T( c.a ) );
}

template < typename T >
static void stencilMode( const MeshCtx& m, const LevelGeom& g, const hyg_op_s::Level& ol, int L, const StencilCall& c, hipStream_t s )
{
   switch ( c.mode )
   {
   case MODE_APPLY:
      stencilT< T, MODE_APPLY >( m, g, ol, L, c, s );
      break;
   case MODE_APPLY_ADD:
      stencilT< T, MODE_APPLY_ADD >( m, g, ol, L, c, s );
      break;
   case MODE_RESIDUAL:
      stencilT< T, MODE_RESIDUAL >( m, g, ol, L, c, s );
      break;
   case MODE_JACOBI:
      stencilT< T, MODE_JACOBI >( m, g, ol, L, c, s );
      break;
   case MODE_CHEB0:
      stencilT< T, MODE_CHEB0 >( m, g, ol, L, c, s );
      break;
   default:
      stencilT< T, MODE_CHEBK >( m, g, ol, L, c, s );
      break;
   }
}

void launchStencil( const MeshCtx& m, const LevelGeom& g, const hyg_op_s::Level& ol, int L, int dtype, const StencilCall& c,
                    hipStream_t s )
{
   if ( dtype == 8 )
      stencilMode< double >( m, g, ol, L, c, s );
   else
      stencilMode< float >( m, g, ol, L, c, s );
}

// ------------------------------------------------------------------ pointwise kernels over all cell slots
// Resolves block -> (cell, tile, point); returns false for threads without a point.
struct SlotPos
{
   int c, x, y, z, slot;
   i64 o;
};

__device__ __forceinline__ bool slotPos( const PlaneTile* tiles, int nTiles, int nC, int L, i64 stride, int lb, SlotPos& sp )
{
   const int nb = nTiles * nC;
   if ( lb >= nb )
      return false;
   const int       c    = lb / nTiles;
   const PlaneTile tile = tiles[lb - c * nTiles];
   const int       N = 1 << L, w = N + 1, W = w - tile.z;
   const int       s = tile.s0 + int( threadIdx.x );
   if ( s >= ( W * ( W + 1 ) ) / 2 )
      return false;
   planeXY( W, s, sp.x, sp.y );
   sp.z    = tile.z;
   sp.c    = c;
   sp.slot = classifySlot( N, sp.x, sp.y, sp.z );
   sp.o    = i64( c ) * stride + planeStart( w, tile.z ) + s;
   return true;
}

template < typename T >
__global__ __launch_bounds__( kBlock ) void kLincomb( const PlaneTile* __restrict__ tiles, int nTiles, int nC, int L, i64 stride,
                                                      const CellInfo* __restrict__ cinfo, int flag, T* __restrict__ dst, int nsrc,
                                                      const T* __restrict__ s0, const T* __restrict__ s1,
                                                      const T* __restrict__ s2, T a0, T a1, T a2, int accumulate, T value )
{
   SlotPos sp;
   if ( !slotPos( tiles, nTiles, nC, L, stride, remapXcd( blockIdx.x, nTiles * nC ), sp ) )
      return;
   if ( !( slotType( cinfo[sp.c], sp.slot ) & flag ) )
      return;
   T r = accumulate ? dst[sp.o] : T( 0 );
   if ( nsrc == 0 )
      r = value;
   if ( nsrc > 0 )
      r += a0 * s0[sp.o];
   if ( nsrc > 1 )
      r += a1 * s1[sp.o];
   if ( nsrc > 2 )
      r += a2 * s2[sp.o];
   dst[sp.o] = r;
}

template < typename T >
__global__ __launch_bounds__( kBlock ) void kMult( const PlaneTile* __restrict__ tiles, int nTiles, int nC, int L, i64 stride,
                                                   const CellInfo* __restrict__ cinfo, int flag, T* dst, const T* a, const T* b )
{
   SlotPos sp;
   if ( !slotPos( tiles, nTiles, nC, L, stride, remapXcd( blockIdx.x, nTiles * nC ), sp ) )
      return;
   if ( !( slotType( cinfo[sp.c], sp.slot ) & flag ) )
      return;
   dst[sp.o] = a[sp.o] * b[sp.o];
}

__device__ __forceinline__ double waveSum( double v )
{
#pragma unroll
   for ( int off = 32; off > 0; off >>= 1 )
      v += __shfl_xor( v, off, 64 );
   return v;
}

__device__ __forceinline__ double blockSum( double v )
{
   __shared__ double red[kBlock / 64];
   v = waveSum( v );
   if ( ( threadIdx.x & 63 ) == 0 )
      red[threadIdx.x >> 6] = v;
   __syncthreads();
   double t = 0.0;
   if ( threadIdx.x == 0 )
      for ( int k = 0; k < kBlock / 64; ++k )
         t += red[k];
   return t;
}

template < typename T >
__global__ __launch_bounds__( kBlock ) void kDot( const PlaneTile* __restrict__ tiles, int nTiles, int nC, int L, i64 stride,
                                                  const CellInfo* __restrict__ cinfo, int flag, const T* __restrict__ a,
                                                  const T* __restrict__ b, double* __restrict__ partials )
{
   SlotPos sp;
   double  v = 0.0;
   if ( slotPos( tiles, nTiles, nC, L, stride, int( blockIdx.x ), sp ) )
   {
      const CellInfo& ci = cinfo[sp.c];
      if ( ( slotType( ci, sp.slot ) & flag ) && ( ( ci.ownmask >> sp.slot ) & 1u ) )
         v = double( a[sp.o] ) * double( b[sp.o] );
   }
   v = blockSum( v );
   if ( threadIdx.x == 0 )
      partials[blockIdx.x] = v;
}

__global__ __launch_bounds__( kBlock ) void kDotFinal( const double* __restrict__ partials, int n, double* __restrict__ result )
{
   double v = 0.0;
   for ( int k = int( threadIdx.x ); k < n; k += kBlock )
      v += partials[k];
   v = blockSum( v );
   if ( threadIdx.x == 0 )
      result[0] = v;
}

template < typename D, typename S >
__global__ __launch_bounds__( kBlock ) void kConvert( i64 n, D* __restrict__ dst, const S* __restrict__ src )
{
   for ( i64 k = i64( blockIdx.x ) * kBlock + threadIdx.x; k < n; k += i64( gridDim.x ) * kBlock )
      dst[k] = D( src[k] );
}

template < typename T >
__global__ __launch_bounds__( kBlock ) void kInvDiag( const PlaneTile* __restrict__ tiles, int nTiles, int nC, int L, i64 stride,
                                                      const CellInfo* __restrict__ cinfo, const T* __restrict__ cellSt,
                                                      const T* __restrict__ primInv, T* __restrict__ dst )
{
   SlotPos sp;
   if ( !slotPos( tiles, nTiles, nC, L, stride, remapXcd( blockIdx.x, nTiles * nC ), sp ) )
      return;
   dst[sp.o] = sp.slot == 0 ? cellSt[16 * sp.c + 15] : primInv[cinfo[sp.c].prim[sp.slot]];
}

// ------------------------------------------------------------------ grid transfer
// Prolongation: every fine point is a coarse point (weight 1) or the midpoint of a coarse micro-edge
// (weights 1/2); both parents lie in the same macro-cell, so each replica is computed in its own cell.
template < typename T >
__global__ __launch_bounds__( kBlock ) void kProlongate( const PlaneTile* __restrict__ tiles, int nTiles, int nC, int Lf,
                                                         i64 strideF, i64 strideC, const CellInfo* __restrict__ cinfo, int flag,
                                                         T* __restrict__ fine, const T* __restrict__ coarse, int add )
{
   SlotPos sp;
   if ( !slotPos( tiles, nTiles, nC, Lf, strideF, remapXcd( blockIdx.x, nTiles * nC ), sp ) )
      return;
   if ( !( slotType( cinfo[sp.c], sp.slot ) & flag ) )
      return;
   const int wc = ( 1 << ( Lf - 1 ) ) + 1;
   const T*  cs = coarse + i64( sp.c ) * strideC;
   const int px = sp.x & 1, py = sp.y & 1, pz = sp.z & 1;
   T         v;
   if ( !( px | py | pz ) )
      v = cs[cidx( wc, sp.x >> 1, sp.y >> 1, sp.z >> 1 )];
   else
   {
      // micro-edge direction of each parity class px + 2 py + 4 pz (the refinement's edges are
      // E, N, SE, TC, BE, BN, TSE up to sign): (1,0,0) (0,1,0) (1,-1,0) (0,0,1) (1,0,-1) (0,1,-1) (1,-1,1)
      constexpr int EX[8] = { 0, 1, 0, 1, 0, 1, 0, 1 };
      constexpr int EY[8] = { 0, 0, 1, -1, 0, 0, 1, -1 };
      constexpr int EZ[8] = { 0, 0, 0, 0, 1, -1, -1, 1 };
      const int     pc    = px | ( py << 1 ) | ( pz << 2 );
      const int     ex = EX[pc], ey = EY[pc], ez = EZ[pc];
      const int     ax = ( sp.x - ex ) >> 1, ay = ( sp.y - ey ) >> 1, az = ( sp.z - ez ) >> 1;
      const int     bx = ( sp.x + ex ) >> 1, by = ( sp.y + ey ) >> 1, bz = ( sp.z + ez ) >> 1;
      v = T( 0.5 ) * ( cs[cidx( wc, ax, ay, az )] + cs[cidx( wc, bx, by, bz )] );
   }
   fine[sp.o] = add ? fine[sp.o] + v : v;
}

// Restriction, coarse points inside a cell: all 14 fine neighbours of 2q are cell-interior.
template < typename T >
__global__ __launch_bounds__( kBlock ) void kRestrictCell( const PlaneTile* __restrict__ tiles, int tile0, int nTiles, int nC,
                                                           int Lc, i64 strideC, i64 strideF, T* __restrict__ coarse,
                                                           const T* __restrict__ fine )
{
   const int nb = nTiles * nC;
   const int lb = remapXcd( blockIdx.x, nb );
   if ( lb >= nb )
      return;
   const int       c    = lb / nTiles;
   const PlaneTile tile = tiles[tile0 + ( lb - c * nTiles )];
   const int       Nc = 1 << Lc, wc = Nc + 1, z = tile.z, W = wc - z;
   const int       s = tile.s0 + int( threadIdx.x );
   if ( s >= ( W * ( W + 1 ) ) / 2 )
      return;
   int x, y;
   planeXY( W, s, x, y );
   if ( x < 1 || y < 1 || x + y + z > Nc - 1 )
      return;
   const int wf = 2 * Nc + 1;
   const T*  fs = fine + i64( c ) * strideF;
   const int fx = 2 * x, fy = 2 * y, fz = 2 * z;
   T         v  = fs[cidx( wf, fx, fy, fz )];
   T         h  = T( 0 );
#pragma unroll
   for ( int d = 1; d < NDIR; ++d )
      h += fs[cidx( wf, fx + dirX( d ), fy + dirY( d ), fz + dirZ( d ) )];
   coarse[i64( c ) * strideC + planeStart( wc, z ) + s] = v + T( 0.5 ) * h;
}

// Restriction, coarse points on macro-faces/-edges/-vertices: sum over the incident cells of the
// in-cell fine neighbours, each scaled by 1 / (#cells sharing the fine point's primitive).
template < typename T >
__global__ __launch_bounds__( kBlock ) void kRestrictLower( const LowerChunk* __restrict__ chunks, int nChunks,
                                                            const LowerPrim* __restrict__ prims,
                                                            const LowerEntry* __restrict__ entries,
                                                            const CellInfo* __restrict__ cinfo, int Lc, i64 strideC,
                                                            i64 strideF, int flag, T* __restrict__ coarse,
                                                            const T* __restrict__ fine )
{
   const int ci = int( blockIdx.x ) * ( kBlock / 64 ) + int( threadIdx.x >> 6 );
   if ( ci >= nChunks )
      return;
   const int        lane = int( threadIdx.x & 63 );
   const LowerChunk ch   = chunks[ci];
   if ( lane >= ch.n )
      return;
   const LowerPrim p = prims[ch.prim];
   if ( !( p.bc & flag ) )
      return;
   const int Nc = 1 << Lc, wc = Nc + 1, Nf = 2 * Nc, wf = Nf + 1;
   const int i = ch.i0 + lane, j = ch.j;
   T         v = T( 0 );
   for ( int k = 0; k < p.nCells; ++k )
   {
      const LowerEntry e = entries[p.entry0 + k];
      int              qx, qy, qz;
      entryPos( e, Nc, i, j, qx, qy, qz );
      const T*        fs = fine + i64( e.cell ) * strideF;
      const CellInfo& cf = cinfo[e.cell];
#pragma unroll
      for ( int d = 0; d < NDIR; ++d )
      {
         const int ax = 2 * qx + dirX( d ), ay = 2 * qy + dirY( d ), az = 2 * qz + dirZ( d );
         if ( !inCell( Nf, ax, ay, az ) )
            continue;
         const T wgt = ( d == 0 ? T( 1 ) : T( 0.5 ) ) * T( cf.invNC[classifySlot( Nf, ax, ay, az )] );
         v += wgt * fs[cidx( wf, ax, ay, az )];
      }
   }
   for ( int k = 0; k < p.nCells; ++k )
   {
      const LowerEntry e = entries[p.entry0 + k];
      int              qx, qy, qz;
      entryPos( e, Nc, i, j, qx, qy, qz );
      coarse[i64( e.cell ) * strideC + cidx( wc, qx, qy, qz )] = v;
   }
}

// ------------------------------------------------------------------ launchers
template < typename T >
static void lincombT( const MeshCtx& m, const LevelGeom& g, int L, void* dst, int nsrc, const double* s, const void* const* src,
                      bool accumulate, double value, int flag, hipStream_t st )
{
   const int nC = m.mesh->nC();
   const int nb = g.nTiles * nC;
   hipLaunchKernelGGL( ( kLincomb< T > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride,
                       m.cellInfo, flag, (T*) dst, nsrc, nsrc > 0 ? (const T*) src[0] : nullptr,
                       nsrc > 1 ? (const T*) src[1] : nullptr, nsrc > 2 ? (const T*) src[2] : nullptr, T( nsrc > 0 ? s[0] : 0 ),
                       T( nsrc > 1 ? s[1] : 0 ), T( nsrc > 2 ? s[2] : 0 ), accumulate ? 1 : 0, T( value ) );
}

void launchLincomb( const MeshCtx& m, const LevelGeom& g, int L, int dtype, void* dst, int nsrc, const double* s,
                    const void* const* src, bool accumulate, double value, int flag, hipStream_t st )
{
   if ( dtype == 8 )
      lincombT< double >( m, g, L, dst, nsrc, s, src, accumulate, value, flag, st );
   else
      lincombT< float >( m, g, L, dst, nsrc, s, src, accumulate, value, flag, st );
}

void launchMult( const MeshCtx& m, const LevelGeom& g, int L, int dtype, void* dst, const void* a, const void* b, int flag,
                 hipStream_t st )
{
   const int nC = m.mesh->nC();
   const int nb = g.nTiles * nC;
   if ( dtype == 8 )
      hipLaunchKernelGGL( ( kMult< double > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride,
                          m.cellInfo, flag, (double*) dst, (const double*) a, (const double*) b );
   else
      hipLaunchKernelGGL( ( kMult< float > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride,
                          m.cellInfo, flag, (float*) dst, (const float*) a, (const float*) b );
}

int dotPartialsNeeded( const MeshCtx& m, const LevelGeom& g ) { return g.nTiles * m.mesh->nC(); }

void launchDot( const MeshCtx& m, const LevelGeom& g, int L, int dtype, const void* a, const void* b, int flag, double* partials,
                double* result, hipStream_t st )
{
   const int nC = m.mesh->nC();
   const int nb = g.nTiles * nC;
   if ( dtype == 8 )
      hipLaunchKernelGGL( ( kDot< double > ), dim3( nb ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride, m.cellInfo,
                          flag, (const double*) a, (const double*) b, partials );
   else
      hipLaunchKernelGGL( ( kDot< float > ), dim3( nb ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride, m.cellInfo,
                          flag, (const float*) a, (const float*) b, partials );
   hipLaunchKernelGGL( kDotFinal, dim3( 1 ), dim3( kBlock ), 0, st, partials, nb, result );
}

void launchConvert( i64 n, int dstType, void* dst, int srcType, const void* src, hipStream_t st )
{
   const int grid = int( std::min< i64 >( ( n + kBlock - 1 ) / kBlock, 65536 ) );
   if ( dstType == 8 && srcType == 4 )
      hipLaunchKernelGGL( ( kConvert< double, float > ), dim3( grid ), dim3( kBlock ), 0, st, n, (double*) dst, (const float*) src );
   else if ( dstType == 4 && srcType == 8 )
      hipLaunchKernelGGL( ( kConvert< float, double > ), dim3( grid ), dim3( kBlock ), 0, st, n, (float*) dst, (const double*) src );
   else if ( dstType == 8 )
      hipLaunchKernelGGL( ( kConvert< double, double > ), dim3( grid ), dim3( kBlock ), 0, st, n, (double*) dst, (const double*) src );
   else
      hipLaunchKernelGGL( ( kConvert< float, float > ), dim3( grid ), dim3( kBlock ), 0, st, n, (float*) dst, (const float*) src );
}

void launchCopy( i64 n, int dtype, void* dst, const void* src, hipStream_t st ) { launchConvert( n, dtype, dst, dtype, src, st ); }

void launchInvDiag( const MeshCtx& m, const LevelGeom& g, const hyg_op_s::Level& ol, int L, int dtype, void* dst, hipStream_t st )
{
   const int nC = m.mesh->nC();
   const int nb = g.nTiles * nC;
   if ( dtype == 8 )
      hipLaunchKernelGGL( ( kInvDiag< double > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride,
                          m.cellInfo, ol.cellSt, ol.primInv, (double*) dst );
   else
      hipLaunchKernelGGL( ( kInvDiag< float > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride,
                          m.cellInfo, ol.cellStF, ol.primInvF, (float*) dst );
}

template < typename T >
static void restrictT( const MeshCtx& m, const LevelGeom& gc, int Lc, void* coarse, const void* fine, int flag, hipStream_t st )
{
   const int nC      = m.mesh->nC();
   const i64 strideF = cellStride( Lc + 1 );
   if ( ( flag & 1 ) && gc.nInteriorTiles > 0 )
   {
      const int nb = gc.nInteriorTiles * nC;
      hipLaunchKernelGGL( ( kRestrictCell< T > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, gc.tiles, gc.interiorTile0,
                          gc.nInteriorTiles, nC, Lc, gc.stride, strideF, (T*) coarse, (const T*) fine );
   }
   if ( gc.nChunks > 0 )
      hipLaunchKernelGGL( ( kRestrictLower< T > ), dim3( ceilDiv( gc.nChunks, kBlock / 64 ) ), dim3( kBlock ), 0, st, gc.chunks,
                          gc.nChunks, m.prims, m.entries, m.cellInfo, Lc, gc.stride, strideF, flag, (T*) coarse,
                          (const T*) fine );
}

void launchRestrict( const MeshCtx& m, const LevelGeom& gc, int Lc, int dtype, void* coarse, const void* fine, int flag,
                     hipStream_t st )
{
   if ( dtype == 8 )
      restrictT< double >( m, gc, Lc, coarse, fine, flag, st );
   else
      restrictT< float >( m, gc, Lc, coarse, fine, flag, st );
}

void launchProlongate( const MeshCtx& m, const LevelGeom& gf, int Lf, int dtype, void* fine, const void* coarse, int flag,
                       bool add, hipStream_t st )
{
   const int nC      = m.mesh->nC();
   const int nb      = gf.nTiles * nC;
   const i64 strideC = cellStride( Lf - 1 );
   if ( dtype == 8 )
      hipLaunchKernelGGL( ( kProlongate< double > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, gf.tiles, gf.nTiles, nC, Lf,
                          gf.stride, strideC, m.cellInfo, flag, (double*) fine, (const double*) coarse, add ? 1 : 0 );
   else
      hipLaunchKernelGGL( ( kProlongate< float > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, gf.tiles, gf.nTiles, nC, Lf,
                          gf.stride, strideC, m.cellInfo, flag, (float*) fine, (const float*) coarse, add ? 1 : 0 );
}

} // namespace hyg
