// gfx950 kernels of the 15-point P1 stencil operations (apply, residual, Jacobi, Chebyshev steps).
//
// Cell interior: z-marching wavefronts. A wavefront owns the row segment (y, x0-1 .. x0+62) of one
// macro-cell and marches up the planes z0 .. z1-1. It keeps rows y-1, y, y+1 of planes z-1, z, z+1
// in registers and has the rows of planes z+2 and z+3 in flight, so each plane step issues three
// coalesced row loads (plus one row of the epilogue input, also prefetched) and no barrier. Row and
// plane offsets are wave-uniform and advanced incrementally in scalar registers, so every load and
// store is `uniform base + lane`. The x-1 / x+1 neighbours come from the adjacent lanes: the terms
// needing lane-1 (W, NW, TW, BNW) and lane+1 (E, SE, TSE, BE) are summed per lane first and each sum
// is moved with one DPP shift (the coefficients are uniform over a macro-cell). Lanes 1..62 produce
// results. Windows may run past a row end; such lanes read other points of the array (function
// arrays carry kTailSlack values after the last cell) and their results are discarded. A valid
// output only reads points of its own cell. Loads and stores go through per-cell buffer
// descriptors with a uniform byte offset, so a step has no per-lane address arithmetic, and the
// stores of discarded lanes are dropped by the descriptor's range check instead of a branch.
//
// Macro-faces, -edges and -vertices. One wavefront per chunk of up to 64 points of a primitive;
// each point sums the one-sided stencils of its incident cells and writes the result to every
// cell's replica of the point. Chunks of one primitive are consecutive and stay on one XCD, so the
// cache lines around neighbouring chunk rows are shared through that XCD's L2.
#include "device_common.hpp"

namespace hyg {

// ------------------------------------------------------------------ cell interior, z-marching wavefronts
// R output rows per wavefront: one plane of the register ring holds rows y0-1 .. y0+R of x and
// rows y0 .. y0+R-1 of the epilogue input.
template < typename T, int R >
struct PlaneRows
{
   T r[R + 2];
   T e[R];
};

template < typename T, int MODE, int R >
__global__ __launch_bounds__( kBlock ) void kCellZ( const ZSeg* __restrict__ segs, int nSegs, int nC, int L, i64 stride,
                                                    const T* __restrict__ cellSt, const T* __restrict__ x,
                                                    const T* __restrict__ in1, T* __restrict__ out0, T* __restrict__ out1, T a )
{
   constexpr bool kIn1 = MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK;
   constexpr bool kE   = kIn1 || MODE == MODE_APPLY_ADD;
   constexpr bool kTwo = MODE == MODE_CHEB0 || MODE == MODE_CHEBK;
   constexpr int  B    = int( sizeof( T ) );
   const int      nbx  = ( nSegs + 3 ) / 4; // blocks per cell
   const int      nb   = nbx * nC;
   const int      lb   = remapXcd( int( blockIdx.x ), nb );
   if ( lb >= nb )
      return;
   const int c  = lb / nbx;
   const int si = ( lb - c * nbx ) * 4 + __builtin_amdgcn_readfirstlane( int( threadIdx.x ) >> 6 );
   if ( si >= nSegs )
      return;
   const ZSeg     sg    = segs[si];
   const int      N     = 1 << L, y0 = sg.y, z0 = sg.z0, z1 = sg.z1;
   const unsigned lane  = threadIdx.x & 63u;
   const unsigned vo    = lane * unsigned( B );
   const i64      cb    = i64( c ) * stride;
   const unsigned bytes = unsigned( ( stride + kTailSlack ) * B ); // cell array plus the read-past tail
   const auto     rx    = makeRsrc( x + cb, bytes );
   const auto     re    = makeRsrc( kIn1 ? in1 + cb : out0 + cb, bytes );
   const auto     r0    = makeRsrc( out0 + cb, bytes );
   const auto     r1    = makeRsrc( kTwo ? out1 + cb : out0 + cb, bytes );
   const T*       st    = cellSt + 16 * c;
   T              w[16];
#pragma unroll
   for ( int d = 0; d < 16; ++d )
      w[d] = st[d];

   // load front: offset of (x0-1, y0-1) in plane pl, whose row 0 has Wl points. Every step issues the
   // same loads (the front stops at the last plane needed, min(z1, N), and repeats it), so the
   // wait before a step's arithmetic only covers loads issued two steps earlier.
   const int plMax = min( z1, N );
   int       pl    = z0 - 1;
   int       Wl    = N + 1 - pl;
   int       off   = S3i( N + 1 ) - S3i( Wl ) + rowStart( Wl, y0 - 1 ) + sg.x0 - 1;
   // lanes past the end of a row take the out-of-range offset: they read 0 and fetch nothing
   const int laneEnd = N + 2 - sg.x0 - y0; // last lane of row y0-1 in plane 0
   auto      loadPlane = [&]( PlaneRows< T, R >& P ) {
      int o = off;
#pragma unroll
      for ( int k = 0; k < R + 2; ++k )
      {
         const int le = laneEnd - pl - k; // last lane holding a point of row y0-1+k in plane pl
         P.r[k]       = bufLoad< T >( rx, int( lane ) <= le ? vo : kDropOffset, o * B );
         if constexpr ( kE )
         {
            if ( k >= 1 && k <= R )
               P.e[k - 1] = bufLoad< T >( re, int( lane ) < le ? vo : kDropOffset, o * B );
         }
         o += Wl - ( y0 - 1 + k ); // start of the next row
      }
      if ( pl < plMax )
      {
         off += T2i( Wl ) - ( y0 - 1 );
         --Wl;
         ++pl;
      }
   };
   // output plane z: offset oz of (x0-1, y0-1), row 0 of the plane has Wz points
   int       z = z0, oz = 0, Wz = N + 1 - z0;
   const int lastLane = N - y0 - sg.x0; // lane of x = N-1-y-z is lastLane - (y - y0) - z
   // A, B, C: planes z-1, z, z+1 (resident); plane z+2 is in flight; E receives plane z+3
   auto step = [&]( const PlaneRows< T, R >& A, const PlaneRows< T, R >& Bp, const PlaneRows< T, R >& C, PlaneRows< T, R >& E ) {
      loadPlane( E );
      int o = oz + ( Wz - ( y0 - 1 ) ); // row y0 of plane z
#pragma unroll
      for ( int r = 0; r < R; ++r )
      {
         // plane z: C = row r+1, N = row r+2, S = row r; plane z+1: TC = row r+1, TS = row r;
         // plane z-1: BC = row r+1, BN = row r+2
         T Ax = w[0] * Bp.r[r + 1];
         Ax += w[3] * Bp.r[r + 2];
         Ax += w[4] * Bp.r[r];
         Ax += w[7] * C.r[r + 1];
         Ax += w[9] * C.r[r];
         Ax += w[11] * A.r[r + 1];
         Ax += w[12] * A.r[r + 2];
         // terms read at x-1 (W, NW, TW, BNW) and at x+1 (E, SE, TSE, BE)
         T sr = w[1] * Bp.r[r + 1];
         sr += w[5] * Bp.r[r + 2];
         sr += w[8] * C.r[r + 1];
         sr += w[14] * A.r[r + 2];
         T sl = w[2] * Bp.r[r + 1];
         sl += w[6] * Bp.r[r];
         sl += w[10] * C.r[r];
         sl += w[13] * A.r[r + 1];
         Ax += laneMove< kDppWaveShr1 >( sr );
         Ax += laneMove< kDppWaveShl1 >( sl );
         T o0, o1;
         epilogue< T, MODE >( Ax, Bp.r[r + 1], Bp.e[r], Bp.e[r], w[15], a, o0, o1 );
         const unsigned so = ( lane - 1u < 62u && int( lane ) <= lastLane - r - z ) ? vo : kDropOffset;
         bufStore( o0, r0, so, o * B );
         if constexpr ( kTwo )
            bufStore( o1, r1, so, o * B );
         o += Wz - ( y0 + r );
      }
      oz += T2i( Wz ) - ( y0 - 1 );
      --Wz;
      ++z;
   };
   PlaneRows< T, R > R0, R1, R2, R3, R4;
#pragma unroll
   for ( int k = 0; k < R; ++k )
      R0.e[k] = R1.e[k] = R2.e[k] = R3.e[k] = R4.e[k] = T( 0 );
   loadPlane( R0 ); // plane z0-1
   oz = off;        // plane z0 (z0 < z1 <= plMax, so the front advanced)
   loadPlane( R1 );
   loadPlane( R2 );
   loadPlane( R3 ); // plane z0+2; the front is now at z0+3
   // five-step register ring: no register copies; every path through a step issues the same
   // memory operations, so loads of two planes stay in flight at each step
   step( R0, R1, R2, R4 );
   while ( z < z1 )
   {
      step( R1, R2, R3, R0 );
      if ( z >= z1 )
         break;
      step( R2, R3, R4, R1 );
      if ( z >= z1 )
         break;
      step( R3, R4, R0, R2 );
      if ( z >= z1 )
         break;
      step( R4, R0, R1, R3 );
      if ( z >= z1 )
         break;
      step( R0, R1, R2, R4 );
   }
}

// ------------------------------------------------------------------ macro-face / -edge / -vertex points
// Each lane gathers the one-sided stencil of every incident cell. Directions leaving a cell carry a
// zero weight in that cell's one-sided stencil; they read the point itself instead, so all loads of a
// lane are unconditional and independent and issue back to back.
template < typename T, int MODE >
__global__ __launch_bounds__( kBlock ) void kLower( const LowerChunk* __restrict__ chunks, int nChunks,
                                                    const LowerPrim* __restrict__ prims, const LowerEntry* __restrict__ entries,
                                                    const T* __restrict__ entryW, const T* __restrict__ primInv, int L,
                                                    i64 stride, int flag, const T* __restrict__ x, const T* __restrict__ in1,
                                                    T* __restrict__ out0, T* __restrict__ out1, T a )
{
   const int nb = ( nChunks + 3 ) / 4;
   const int ci = remapXcd( int( blockIdx.x ), nb ) * ( kBlock / 64 ) + __builtin_amdgcn_readfirstlane( int( threadIdx.x >> 6 ) );
   if ( ci >= nChunks )
      return;
   const int        lane = int( threadIdx.x & 63 );
   const LowerChunk ch   = chunks[ci];
   const LowerPrim  p    = prims[ch.prim];
   if ( !( p.bc & flag ) )
      return;
   const bool on = lane < ch.n;
   const int  N = 1 << L, w = N + 1;
   const int  i = ch.i0 + ( on ? lane : 0 ), j = ch.j;
   T          Ax = T( 0 );
   for ( int k = 0; k < p.nCells; ++k )
   {
      const LowerEntry e = entries[p.entry0 + k];
      int              qx, qy, qz;
      entryPos( e, N, i, j, qx, qy, qz );
      const T*  xs = x + i64( e.cell ) * stride;
      const T*  wv = entryW + i64( p.entry0 + k ) * 16;
      const int q  = cidx( w, qx, qy, qz );
      T         v[NDIR];
#pragma unroll
      for ( int d = 0; d < NDIR; ++d )
      {
         const int ax = qx + dirX( d ), ay = qy + dirY( d ), az = qz + dirZ( d );
         v[d]         = xs[inCell( N, ax, ay, az ) ? cidx( w, ax, ay, az ) : q];
      }
#pragma unroll
      for ( int d = 0; d < NDIR; ++d )
         Ax += wv[d] * v[d];
   }
   if ( !on )
      return;
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

// Macro-face points: the two incident cells' one-sided stencils, with the 15 neighbour offsets of a
// point derived incrementally from its own index (no per-direction index arithmetic). Directions
// with a zero weight (those leaving the cell) read the point itself.
template < typename T, int MODE >
__global__ __launch_bounds__( kBlock ) void kFace( const LowerChunk* __restrict__ chunks, int nChunks,
                                                   const LowerPrim* __restrict__ prims, const LowerEntry* __restrict__ entries,
                                                   const T* __restrict__ entryW, const T* __restrict__ primInv, int L,
                                                   i64 stride, int flag, const T* __restrict__ x, const T* __restrict__ in1,
                                                   T* __restrict__ out0, T* __restrict__ out1, T a )
{
   const int nb = ( nChunks + 3 ) / 4;
   const int ci = remapXcd( int( blockIdx.x ), nb ) * ( kBlock / 64 ) + __builtin_amdgcn_readfirstlane( int( threadIdx.x >> 6 ) );
   if ( ci >= nChunks )
      return;
   const LowerChunk ch = chunks[ci];
   const LowerPrim  p  = prims[ch.prim];
   if ( !( p.bc & flag ) )
      return;
   const int  lane = int( threadIdx.x & 63 );
   const bool on   = lane < ch.n;
   const int  N = 1 << L, w = N + 1;
   const int  i = ch.i0 + ( on ? lane : 0 ), j = ch.j;
   T          v[2][NDIR];
   i64        o[2];
#pragma unroll
   for ( int k = 0; k < 2; ++k )
   {
      const int        ek = p.entry0 + ( k < p.nCells ? k : 0 );
      const LowerEntry e  = entries[ek];
      int              qx, qy, qz;
      entryPos( e, N, i, j, qx, qy, qz );
      const int W  = w - qz;
      const int tc = T2i( W ) - qy, bc = T2i( W + 1 ) - qy, nn = W - qy;
      const int off[NDIR] = { 0,  -1,      1,          nn,     -( nn + 1 ),      nn - 1, -nn, tc, tc - 1, tc - nn, tc - nn + 1,
                              -bc, -bc + nn + 1, -bc + 1, -bc + nn };
      o[k]                = i64( e.cell ) * stride + cidx( w, qx, qy, qz );
      const T* xs         = x + o[k];
      const T* wv         = entryW + i64( ek ) * 16;
#pragma unroll
      for ( int d = 0; d < NDIR; ++d )
         v[k][d] = xs[wv[d] != T( 0 ) ? off[d] : 0];
   }
   T Ax = T( 0 );
#pragma unroll
   for ( int k = 0; k < 2; ++k )
   {
      const T* wv = entryW + i64( p.entry0 + ( k < p.nCells ? k : 0 ) ) * 16;
      const T  s  = k < p.nCells ? T( 1 ) : T( 0 );
#pragma unroll
      for ( int d = 0; d < NDIR; ++d )
         Ax += ( s * wv[d] ) * v[k][d];
   }
   if ( !on )
      return;
   T i1 = T( 0 ), ov = T( 0 ), r0, r1;
   if constexpr ( MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
      i1 = in1[o[0]];
   if constexpr ( MODE == MODE_APPLY_ADD )
      ov = out0[o[0]];
   epilogue< T, MODE >( Ax, v[0][0], i1, ov, primInv[ch.prim], a, r0, r1 );
   out0[o[0]] = r0;
   if constexpr ( MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
      out1[o[0]] = r1;
   if ( p.nCells > 1 )
   {
      out0[o[1]] = r0;
      if constexpr ( MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
         out1[o[1]] = r1;
   }
}

// ------------------------------------------------------------------ launch
// cell-interior configuration of a level: set by the tuner, else HYG_CELL_CFG, else rows = 2
int cellConfig( LevelGeom& g, int dtype )
{
   int& c = g.cellCfg[dtype == 8 ? 0 : 1];
   if ( c < 0 )
   {
      const char* e = getenv( "HYG_CELL_CFG" );
      c             = e ? std::max( 0, std::min( kNumCellCfgs - 1, atoi( e ) ) ) : 1;
   }
   return c;
}
void setCellConfig( LevelGeom& g, int dtype, int cfg ) { g.cellCfg[dtype == 8 ? 0 : 1] = cfg; }

template < typename T, int MODE >
static void stencilT( const MeshCtx& m, LevelGeom& g, const hyg_op_s::Level& ol, int L, const StencilCall& c, hipStream_t s )
{
   const int nC  = m.mesh->nC();
   const T*  cst = std::is_same< T, double >::value ? (const T*) ol.cellSt : (const T*) ol.cellStF;
   const T*  ew  = std::is_same< T, double >::value ? (const T*) ol.entryW : (const T*) ol.entryWF;
   const T*  pi  = std::is_same< T, double >::value ? (const T*) ol.primInv : (const T*) ol.primInvF;
   const int cfg = cellConfig( g, std::is_same< T, double >::value ? 8 : 4 );
   if ( ( c.flag & 1 ) && g.nZSegs[cfg] > 0 )
   {
      const int nb = ( ( g.nZSegs[cfg] + 3 ) / 4 ) * nC;
      auto      kf = cfg == 0 ? kCellZ< T, MODE, 1 > : ( cfg == 1 ? kCellZ< T, MODE, 2 > : kCellZ< T, MODE, 4 > );
      hipLaunchKernelGGL( kf, dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, s, g.zSegs[cfg], g.nZSegs[cfg], nC, L, g.stride, cst,
                          (const T*) c.x, (const T*) c.in1, (T*) c.out0, (T*) c.out1, T( c.a ) );
   }
   const int nLow = g.firstFaceChunk, nFace = g.nChunks - g.firstFaceChunk;
   if ( nLow > 0 )
      hipLaunchKernelGGL( ( kLower< T, MODE > ), dim3( xcdGrid( ceilDiv( nLow, kBlock / 64 ) ) ), dim3( kBlock ), 0, s, g.chunks,
                          nLow, m.prims, m.entries, ew, pi, L, g.stride, c.flag, (const T*) c.x, (const T*) c.in1,
                          (T*) c.out0, (T*) c.out1, T( c.a ) );
   if ( nFace > 0 )
      hipLaunchKernelGGL( ( kFace< T, MODE > ), dim3( xcdGrid( ceilDiv( nFace, kBlock / 64 ) ) ), dim3( kBlock ), 0, s,
                          g.chunks + g.firstFaceChunk, nFace, m.prims, m.entries, ew, pi, L, g.stride, c.flag, (const T*) c.x,
                          (const T*) c.in1, (T*) c.out0, (T*) c.out1, T( c.a ) );
}

template < typename T >
static void stencilMode( const MeshCtx& m, LevelGeom& g, const hyg_op_s::Level& ol, int L, const StencilCall& c, hipStream_t s )
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

void launchStencil( const MeshCtx& m, LevelGeom& g, const hyg_op_s::Level& ol, int L, int dtype, const StencilCall& c,
                    hipStream_t s )
{
   if ( dtype == 8 )
      stencilMode< double >( m, g, ol, L, c, s );
   else
      stencilMode< float >( m, g, ol, L, c, s );
}

} // namespace hyg
