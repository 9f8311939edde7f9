// gfx950 kernels of the 15-point P1 stencil operations (apply, residual, Jacobi, Chebyshev steps).
//
// Cell interior. A block owns a band of interior rows [y0, y0 + ROWS) of planes [z0, z1) of one
// macro-cell. For every plane p it stages the band slab, rows y0-1 .. y0+ROWS of plane p (one
// contiguous range of the cell array), in a ring of three LDS slots; while plane z is computed
// from the slabs of planes z-1, z, z+1, the slab of plane z+2 is already in flight in registers and
// replaces the slab of plane z-1 afterwards. Each input value is therefore read from HBM once per
// band plus the two halo rows shared with the neighbouring bands, independent of cache behaviour.
// Within a plane the band's points are cut into windows of 64 consecutive points; each lane reads
// 7 values (its own point in rows (y,z) (y+1,z) (y-1,z) (y,z+1) (y-1,z+1) (y,z-1) (y+1,z-1)) and
// takes the x-1 / x+1 neighbours from the adjacent lanes with DPP shifts; lanes 1..62 produce results.
// Bands are dispatched cell-major so that each XCD (block b runs on XCD b mod 8) works on its own
// cells, with the neighbouring bands of a plane running side by side.
//
// Macro-faces, -edges and -vertices. One wavefront per chunk of up to 64 points of a primitive;
// each point sums the one-sided stencils of its incident cells and writes the result to every
// cell's replica of the point.
#include "device_common.hpp"

namespace hyg {

// ------------------------------------------------------------------ cell interior, LDS-staged bands
__device__ __forceinline__ int slabStart( int N, int p, int y0 ) { return rowStart( N + 1 - p, y0 - 1 ); }

template < int ROWS >
__device__ __forceinline__ int slabEnd( int N, int p, int y0 )
{
   const int yb = min( y0 + ROWS, N - p );
   return yb < y0 - 1 ? slabStart( N, p, y0 ) : rowStart( N + 1 - p, yb + 1 );
}

template < typename T, int MODE, int ROWS, int NT >
__global__ __launch_bounds__( NT ) void kCellBand( const BandTile* __restrict__ tiles, int nTiles, int nC, int L, i64 stride,
                                                   const T* __restrict__ cellSt, const T* __restrict__ x,
                                                   const T* __restrict__ in1, T* __restrict__ out0, T* __restrict__ out1, T a )
{
   extern __shared__ char ldsRaw[];
   T*            lds  = reinterpret_cast< T* >( ldsRaw );
   const int     N    = 1 << L, w1 = N + 1;
   const int     slab = ( ROWS + 2 ) * w1 + 64; // + 64: lanes past a row end read (unused) padding
   constexpr int NW   = NT / 64;
   // register prefetch of one slab for levels up to kBandMaxLevel
   constexpr int PER = ( ( ROWS + 2 ) * ( ( 1 << kBandMaxLevel ) + 1 ) + NT - 1 ) / NT;

   const int nb = nTiles * nC;
   const int lb = remapXcd( int( blockIdx.x ), nb );
   if ( lb >= nb )
      return;
   const int      c    = lb / nTiles;
   const BandTile tile = tiles[lb - c * nTiles];
   const int      y0   = tile.y0;
   const int      tid  = int( threadIdx.x );
   const int      lane = tid & 63, wave = tid >> 6;
   const i64      base = i64( c ) * stride;
   const T*       xs   = x + base;
   const T*       st   = cellSt + 16 * c;
   T              w[16];
#pragma unroll
   for ( int d = 0; d < 16; ++d )
      w[d] = st[d];

   for ( int p = tile.z0 - 1; p <= tile.z0 + 1; ++p )
   {
      const int a0 = slabStart( N, p, y0 ), n = slabEnd< ROWS >( N, p, y0 ) - a0;
      const int pb = S3i( w1 ) - S3i( w1 - p ) + a0;
      T*        dst = lds + ( p % 3 ) * slab;
      for ( int k = tid; k < n; k += NT )
         dst[k] = xs[pb + k];
   }
   __syncthreads();

   for ( int z = tile.z0; z < tile.z1; ++z )
   {
      // prefetch the slab of plane z+2 into registers
      const int  pn = z + 2;
      const bool pf = pn <= tile.z1 && pn <= N;
      T          pre[PER];
      int        npre = 0;
      if ( pf )
      {
         const int a0 = slabStart( N, pn, y0 ), n = slabEnd< ROWS >( N, pn, y0 ) - a0;
         const int pb = S3i( w1 ) - S3i( w1 - pn ) + a0;
         npre         = n;
#pragma unroll
         for ( int q = 0; q < PER; ++q )
         {
            const int k = tid + q * NT;
            pre[q]      = k < n ? xs[pb + k] : T( 0 );
         }
      }
      const int W = w1 - z, M = N - z;
      const int yl = min( y0 + ROWS - 1, M - 2 ); // last interior row of the band in plane z
      const T*  sm = lds + ( ( z + 2 ) % 3 ) * slab; // plane z-1
      const T*  s0 = lds + ( z % 3 ) * slab;
      const T*  sp = lds + ( ( z + 1 ) % 3 ) * slab;
      const int bm = slabStart( N, z - 1, y0 ), b0 = slabStart( N, z, y0 ), bp = slabStart( N, z + 1, y0 );
      const int p0 = S3i( w1 ) - S3i( w1 - z );
      // each wave owns whole rows; all row offsets below are wave-uniform
      for ( int y = y0 + wave; y <= yl; y += NW )
      {
         const int len = M - 1 - y; // interior points x = 1 .. len
         const int rC  = rowStart( W, y ) - b0;
         const int rN  = rowStart( W, y + 1 ) - b0;
         const int rS  = rowStart( W, y - 1 ) - b0;
         const int rT  = rowStart( W - 1, y ) - bp;
         const int rTS = rowStart( W - 1, y - 1 ) - bp;
         const int rB  = rowStart( W + 1, y ) - bm;
         const int rBN = rowStart( W + 1, y + 1 ) - bm;
         const i64 go  = base + p0 + rowStart( W, y );
         for ( int x0 = 1; x0 <= len; x0 += 62 )
         {
            const int xx   = x0 - 1 + lane;
            const T   vC   = s0[rC + xx];
            const T   vN   = s0[rN + xx];
            const T   vS   = s0[rS + xx];
            const T   vT   = sp[rT + xx];
            const T   vTS  = sp[rTS + xx];
            const T   vB   = sm[rB + xx];
            const T   vBN  = sm[rBN + xx];
            const T   vW   = laneShift< kDppWaveShr1 >( vC );
            const T   vE   = laneShift< kDppWaveShl1 >( vC );
            const T   vNW  = laneShift< kDppWaveShr1 >( vN );
            const T   vSE  = laneShift< kDppWaveShl1 >( vS );
            const T   vTW  = laneShift< kDppWaveShr1 >( vT );
            const T   vTSE = laneShift< kDppWaveShl1 >( vTS );
            const T   vBE  = laneShift< kDppWaveShl1 >( vB );
            const T   vBNW = laneShift< kDppWaveShr1 >( vBN );
            if ( lane >= 1 && lane <= 62 && xx <= len )
            {
               T Ax = w[0] * vC;
               Ax += w[1] * vW;
               Ax += w[2] * vE;
               Ax += w[3] * vN;
               Ax += w[4] * vS;
               Ax += w[5] * vNW;
               Ax += w[6] * vSE;
               Ax += w[7] * vT;
               Ax += w[8] * vTW;
               Ax += w[9] * vTS;
               Ax += w[10] * vTSE;
               Ax += w[11] * vB;
               Ax += w[12] * vBN;
               Ax += w[13] * vBE;
               Ax += w[14] * vBNW;
               const i64 o  = go + xx;
               T         i1 = T( 0 ), ov = T( 0 ), r0, r1;
               if constexpr ( MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
                  i1 = in1[o];
               if constexpr ( MODE == MODE_APPLY_ADD )
                  ov = out0[o];
               epilogue< T, MODE >( Ax, vC, i1, ov, w[15], a, r0, r1 );
               storeStream( out0 + o, r0 );
               if constexpr ( MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
                  storeStream( out1 + o, r1 );
            }
         }
      }
      __syncthreads(); // every wave is done with the slab of plane z-1
      if ( pf )
      {
         T* dst = lds + ( pn % 3 ) * slab;
#pragma unroll
         for ( int q = 0; q < PER; ++q )
         {
            const int k = tid + q * NT;
            if ( k < npre )
               dst[k] = pre[q];
         }
      }
      __syncthreads();
   }
}

// ------------------------------------------------------------------ cell interior, z-marching wavefronts
// Each wavefront owns the row segment (y, x0-1 .. x0+62) and marches up its planes. It keeps rows
// y-1, y, y+1 of planes z-1, z, z+1 in registers and has the rows of planes z+2 and z+3 in flight, so
// every plane step issues three coalesced row loads and no barrier is needed. The four wavefronts
// of a block march adjacent rows of the same column, so the rows they share are L1 hits.
template < typename T >
struct Rows3
{
   T m, c, p; // rows y-1, y, y+1 at x = x0-1+lane
};

template < typename T >
__device__ __forceinline__ void loadRows3( const T* __restrict__ xs, int w1, int N, int pl, int y, int xx, Rows3< T >& r )
{
   // plane pl, rows y-1 .. y+1; a row r of plane pl holds x = 0 .. N - pl - r
   const int W  = w1 - pl;
   const int pb = S3i( w1 ) - S3i( w1 - pl );
   const int lm = N - pl - ( y - 1 ), lc = lm - 1, lp = lm - 2;
   r.m          = xx <= lm ? xs[pb + rowStart( W, y - 1 ) + xx] : T( 0 );
   r.c          = xx <= lc ? xs[pb + rowStart( W, y ) + xx] : T( 0 );
   r.p          = ( lp >= 0 && xx <= lp ) ? xs[pb + rowStart( W, y + 1 ) + xx] : T( 0 );
}

template < typename T, int MODE >
__global__ __launch_bounds__( kBlock ) void kCellZ( const ZSeg* __restrict__ segs, int nSegs, int nC, int L, i64 stride,
                                                    const T* __restrict__ cellSt, const T* __restrict__ x,
                                                    const T* __restrict__ in1, T* __restrict__ out0, T* __restrict__ out1, T a )
{
   const int nbx = ( nSegs + 3 ) / 4; // blocks per cell
   const int nb  = nbx * nC;
   const int lb  = remapXcd( int( blockIdx.x ), nb );
   if ( lb >= nb )
      return;
   const int c  = lb / nbx;
   const int si = ( lb - c * nbx ) * 4 + int( threadIdx.x >> 6 );
   if ( si >= nSegs )
      return;
   const ZSeg sg   = segs[si];
   const int  N    = 1 << L, w1 = N + 1, y = sg.y;
   const int  lane = int( threadIdx.x & 63 );
   const int  xx   = sg.x0 - 1 + lane;
   const i64  base = i64( c ) * stride;
   const T*   xs   = x + base;
   const T*   st   = cellSt + 16 * c;
   T          w[16];
#pragma unroll
   for ( int d = 0; d < 16; ++d )
      w[d] = st[d];

   Rows3< T > pm, p0, pp, q2, q3;
   loadRows3( xs, w1, N, sg.z0 - 1, y, xx, pm );
   loadRows3( xs, w1, N, sg.z0, y, xx, p0 );
   loadRows3( xs, w1, N, sg.z0 + 1, y, xx, pp );
   if ( sg.z0 + 2 <= N )
      loadRows3( xs, w1, N, sg.z0 + 2, y, xx, q2 );
   for ( int z = sg.z0; z < sg.z1; ++z )
   {
      if ( z + 3 <= N && z + 3 <= sg.z1 )
         loadRows3( xs, w1, N, z + 3, y, xx, q3 );
      // plane z: (y,z) = p0.c, (y+1,z) = p0.p, (y-1,z) = p0.m; (y,z+1) = pp.c, (y-1,z+1) = pp.m;
      // (y,z-1) = pm.c, (y+1,z-1) = pm.p
      const T vW   = laneShift< kDppWaveShr1 >( p0.c );
      const T vE   = laneShift< kDppWaveShl1 >( p0.c );
      const T vNW  = laneShift< kDppWaveShr1 >( p0.p );
      const T vSE  = laneShift< kDppWaveShl1 >( p0.m );
      const T vTW  = laneShift< kDppWaveShr1 >( pp.c );
      const T vTSE = laneShift< kDppWaveShl1 >( pp.m );
      const T vBE  = laneShift< kDppWaveShl1 >( pm.c );
      const T vBNW = laneShift< kDppWaveShr1 >( pm.p );
      const int M  = N - z;
      if ( lane >= 1 && lane <= 62 && xx <= M - 1 - y )
      {
         T Ax = w[0] * p0.c;
         Ax += w[1] * vW;
         Ax += w[2] * vE;
         Ax += w[3] * p0.p;
         Ax += w[4] * p0.m;
         Ax += w[5] * vNW;
         Ax += w[6] * vSE;
         Ax += w[7] * pp.c;
         Ax += w[8] * vTW;
         Ax += w[9] * pp.m;
         Ax += w[10] * vTSE;
         Ax += w[11] * pm.c;
         Ax += w[12] * pm.p;
         Ax += w[13] * vBE;
         Ax += w[14] * vBNW;
         const i64 o  = base + ( S3i( w1 ) - S3i( w1 - z ) ) + rowStart( w1 - z, y ) + xx;
         T         i1 = T( 0 ), ov = T( 0 ), r0, r1;
         if constexpr ( MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
            i1 = in1[o];
         if constexpr ( MODE == MODE_APPLY_ADD )
            ov = out0[o];
         epilogue< T, MODE >( Ax, p0.c, i1, ov, w[15], a, r0, r1 );
         storeStream( out0 + o, r0 );
         if constexpr ( MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
            storeStream( out1 + o, r1 );
      }
      pm = p0;
      p0 = pp;
      pp = q2;
      q2 = q3;
   }
}

// ------------------------------------------------------------------ cell interior, levels above kBandMaxLevel
// Blocks of 4 waves x 2 windows of consecutive points of one plane, read through the caches.
template < typename T, int MODE >
__global__ __launch_bounds__( kBlock ) void kCellFlat( const StencilTile* __restrict__ tiles, int nTiles, int nC, int L, i64 stride,
                                                       const T* __restrict__ cellSt, const T* __restrict__ x,
                                                       const T* __restrict__ in1, T* __restrict__ out0, T* __restrict__ out1, T a )
{
   const int nb = nTiles * nC;
   const int lb = remapXcd( int( blockIdx.x ), nb );
   if ( lb >= nb )
      return;
   const int         c    = lb / nTiles;
   const StencilTile tile = tiles[lb - c * nTiles];
   const int         N = 1 << L, z = tile.z, W = N + 1 - z, M = N - z;
   const int         TW = ( W * ( W + 1 ) ) / 2, TW1 = ( ( W + 1 ) * ( W + 2 ) ) / 2;
   const int         lane = int( threadIdx.x & 63 ), wave = int( threadIdx.x >> 6 );
   const i64         base = i64( c ) * stride;
   const T*          xs   = x + base;
   const T*          st   = cellSt + 16 * c;
#pragma unroll 1
   for ( int u = 0; u < 4; ++u )
   {
      const int s  = tile.s0 + 62 * ( wave + 4 * u ) - 1 + lane;
      const bool pt = s < TW;
      int        px = 0, py = 0;
      if ( pt )
         planeXY( W, s, px, py );
      const int iC = tile.p0 + s, iT = iC + TW - py, iB = iC - TW1 + py;
      const T   vC  = pt ? xs[iC] : T( 0 );
      const T   vN  = ( pt && px + py + 1 <= M ) ? xs[iC + W - py] : T( 0 );
      const T   vS  = ( pt && py >= 1 ) ? xs[iC - ( W - py + 1 )] : T( 0 );
      const T   vT  = ( pt && px + py <= M - 1 ) ? xs[iT] : T( 0 );
      const T   vTS = ( pt && py >= 1 ) ? xs[iT - ( W - py )] : T( 0 );
      const T   vB  = pt ? xs[iB] : T( 0 );
      const T   vBN = pt ? xs[iB + W + 1 - py] : T( 0 );
      const T   vW   = laneShift< kDppWaveShr1 >( vC );
      const T   vE   = laneShift< kDppWaveShl1 >( vC );
      const T   vNW  = laneShift< kDppWaveShr1 >( vN );
      const T   vSE  = laneShift< kDppWaveShl1 >( vS );
      const T   vTW  = laneShift< kDppWaveShr1 >( vT );
      const T   vTSE = laneShift< kDppWaveShl1 >( vTS );
      const T   vBE  = laneShift< kDppWaveShl1 >( vB );
      const T   vBNW = laneShift< kDppWaveShr1 >( vBN );
      if ( pt && lane >= 1 && lane <= 62 && px >= 1 && py >= 1 && px + py <= M - 1 )
      {
         T Ax = st[0] * vC + st[1] * vW + st[2] * vE + st[3] * vN + st[4] * vS + st[5] * vNW + st[6] * vSE + st[7] * vT +
                st[8] * vTW + st[9] * vTS + st[10] * vTSE + st[11] * vB + st[12] * vBN + st[13] * vBE + st[14] * vBNW;
         const i64 o  = base + iC;
         T         i1 = T( 0 ), ov = T( 0 ), r0, r1;
         if constexpr ( MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
            i1 = in1[o];
         if constexpr ( MODE == MODE_APPLY_ADD )
            ov = out0[o];
         epilogue< T, MODE >( Ax, vC, i1, ov, st[15], a, r0, r1 );
         storeStream( out0 + o, r0 );
         if constexpr ( MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
            storeStream( out1 + o, r1 );
      }
   }
}

// ------------------------------------------------------------------ macro-face / -edge / -vertex points
template < typename T, int MODE >
__global__ __launch_bounds__( kBlock ) void kLower( const LowerChunk* __restrict__ chunks, int nChunks,
                                                    const LowerPrim* __restrict__ prims, const LowerEntry* __restrict__ entries,
                                                    const T* __restrict__ entryW, const T* __restrict__ primInv, int L,
                                                    i64 stride, int flag, const T* __restrict__ x, const T* __restrict__ in1,
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
         const T   wd = wv[d];
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

// ------------------------------------------------------------------ launch configuration
// Band configurations (rows per band, threads per block); the default is chosen per value type and
// can be replaced by the tuner (hyg_tune_stencil).
struct BandCfg
{
   int rows, threads;
};
static const BandCfg kBandCfgs[] = { { 4, 256 }, { 8, 256 }, { 8, 512 }, { 16, 512 }, { 16, 1024 } };
constexpr int        kNumBandCfgs = int( sizeof( kBandCfgs ) / sizeof( kBandCfgs[0] ) );

static int g_bandCfg[2] = { -1, -1 }; // per value type (0: FP64, 1: FP32)

int stencilConfigCount() { return kNumBandCfgs; }

// cell-interior kernel family: 0 LDS-staged bands, 1 z-marching wavefronts
static int cellKernelKind()
{
   static int k = -1;
   if ( k < 0 )
   {
      const char* e = getenv( "HYG_CELL_KERNEL" );
      k             = e ? atoi( e ) : 1;
   }
   return k;
}
void setStencilConfig( int dtype, int cfg ) { g_bandCfg[dtype == 8 ? 0 : 1] = cfg; }
int  getStencilConfig( int dtype )
{
   int& c = g_bandCfg[dtype == 8 ? 0 : 1];
   if ( c < 0 )
   {
      const char* e = getenv( "HYG_BAND_CFG" );
      c             = e ? std::max( 0, std::min( kNumBandCfgs - 1, atoi( e ) ) ) : 2;
   }
   return c;
}

template < typename T, int MODE, int ROWS, int NT >
static void launchBand( const MeshCtx& m, LevelGeom& g, const T* cst, int L, const StencilCall& c, hipStream_t s )
{
   const std::vector< BandTile >& tl    = g.bandTilesFor( ROWS );
   const BandTile*                tiles = g.bandTilesDev( ROWS );
   const int                      nC    = m.mesh->nC();
   const int                      nb    = int( tl.size() ) * nC;
   if ( nb == 0 )
      return;
   const size_t lds = size_t( 3 ) * ( ( ROWS + 2 ) * ( ( 1 << L ) + 1 ) + 64 ) * sizeof( T );
   static bool  ldsAttr = false; // dynamic LDS beyond 64 KiB must be enabled per kernel
   if ( !ldsAttr )
   {
      (void) hipFuncSetAttribute( reinterpret_cast< const void* >( &kCellBand< T, MODE, ROWS, NT > ),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 );
      ldsAttr = true;
   }
   hipLaunchKernelGGL( ( kCellBand< T, MODE, ROWS, NT > ), dim3( xcdGrid( nb ) ), dim3( NT ), lds, s, tiles, int( tl.size() ), nC,
                       L, g.stride, cst, (const T*) c.x, (const T*) c.in1, (T*) c.out0, (T*) c.out1, T( c.a ) );
}

template < typename T, int MODE >
static void stencilT( const MeshCtx& m, LevelGeom& g, const hyg_op_s::Level& ol, int L, const StencilCall& c, hipStream_t s )
{
   const int nC  = m.mesh->nC();
   const T*  cst = std::is_same< T, double >::value ? (const T*) ol.cellSt : (const T*) ol.cellStF;
   const T*  ew  = std::is_same< T, double >::value ? (const T*) ol.entryW : (const T*) ol.entryWF;
   const T*  pi  = std::is_same< T, double >::value ? (const T*) ol.primInv : (const T*) ol.primInvF;
   if ( ( c.flag & 1 ) && g.nStTiles > 0 )
   {
      if ( cellKernelKind() == 1 && g.nZSegs > 0 )
      {
         const int nb = ( ( g.nZSegs + 3 ) / 4 ) * nC;
         hipLaunchKernelGGL( ( kCellZ< T, MODE > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, s, g.zSegs, g.nZSegs, nC, L, g.stride,
                             cst, (const T*) c.x, (const T*) c.in1, (T*) c.out0, (T*) c.out1, T( c.a ) );
      }
      else if ( L <= kBandMaxLevel )
      {
         switch ( getStencilConfig( std::is_same< T, double >::value ? 8 : 4 ) )
         {
         case 0:
            launchBand< T, MODE, 4, 256 >( m, g, cst, L, c, s );
            break;
         case 1:
            launchBand< T, MODE, 8, 256 >( m, g, cst, L, c, s );
            break;
         case 2:
            launchBand< T, MODE, 8, 512 >( m, g, cst, L, c, s );
            break;
         case 3:
            launchBand< T, MODE, 16, 512 >( m, g, cst, L, c, s );
            break;
         default:
            launchBand< T, MODE, 16, 1024 >( m, g, cst, L, c, s );
            break;
         }
      }
      else
      {
         const int nb = g.nStTiles * nC;
         hipLaunchKernelGGL( ( kCellFlat< T, MODE > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, s, g.stTiles, g.nStTiles, nC, L,
                             g.stride, cst, (const T*) c.x, (const T*) c.in1, (T*) c.out0, (T*) c.out1, T( c.a ) );
      }
   }
   if ( g.nChunks > 0 )
      hipLaunchKernelGGL( ( kLower< T, MODE > ), dim3( ceilDiv( g.nChunks, kBlock / 64 ) ), dim3( kBlock ), 0, s, g.chunks,
                          g.nChunks, m.prims, m.entries, ew, pi, L, g.stride, c.flag, (const T*) c.x, (const T*) c.in1,
                          (T*) c.out0, (T*) c.out1, T( c.a ) );
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
