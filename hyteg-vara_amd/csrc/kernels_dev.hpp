// Device code of the pointwise, reduction and grid-transfer kernels (bodies shared by the single-problem and the batched launches).
#pragma once

#include "device_common.hpp"

namespace hyg {

// ------------------------------------------------------------------ pointwise kernels over all cell slots
// A plane tile holds kPtsPerThread x 256 consecutive points of one plane; thread t handles points
// s0 + t + 256 u. Positions are resolved first, then all loads are issued, then results stored.
constexpr int kPtsPerThread = 4;

struct SlotPos
{
   int  c, x, y, z, slot;
   i64  o;
   bool ok;
};

__device__ __forceinline__ int tileCell( const PlaneTile* tiles, int nTiles, int lb, PlaneTile& tile )
{
   const int c = lb / nTiles;
   tile        = tiles[lb - c * nTiles];
   return c;
}

__device__ __forceinline__ void slotAt( int c, const PlaneTile& tile, int N, i64 stride, int s, SlotPos& sp )
{
   const int w = N + 1, W = w - tile.z;
   sp.ok       = s < ( W * ( W + 1 ) ) / 2;
   sp.c        = c;
   sp.z        = tile.z;
   sp.x = sp.y = 0;
   if ( sp.ok )
      planeXY( W, s, sp.x, sp.y );
   sp.slot = classifySlot( N, sp.x, sp.y, sp.z );
   sp.o    = i64( c ) * stride + ( S3i( w ) - S3i( w - tile.z ) ) + s;
}

template < typename T >
__device__ __forceinline__ void bLincomb( int bid, const PlaneTile* __restrict__ tiles, int nTiles, int nC, int L, i64 stride,
                                                      const CellInfo* __restrict__ cinfo, int flag, T* __restrict__ dst, int nsrc,
                                                      const T* __restrict__ s0, const T* __restrict__ s1,
                                                      const T* __restrict__ s2, T a0, T a1, T a2, int accumulate, T value )
{
   const int lb = remapXcd( bid, nTiles * nC );
   if ( lb >= nTiles * nC )
      return;
   PlaneTile      tile;
   const int      c  = tileCell( tiles, nTiles, lb, tile );
   const uint64_t bw = cinfo[c].bcword;
   SlotPos        sp[kPtsPerThread];
   bool           on[kPtsPerThread];
#pragma unroll
   for ( int u = 0; u < kPtsPerThread; ++u )
   {
      slotAt( c, tile, 1 << L, stride, tile.s0 + int( threadIdx.x ) + kBlock * u, sp[u] );
      on[u] = sp[u].ok && ( int( ( bw >> ( 4 * sp[u].slot ) ) & 15u ) & flag );
   }
   T r[kPtsPerThread];
#pragma unroll
   for ( int u = 0; u < kPtsPerThread; ++u )
   {
      T v = ( accumulate && on[u] ) ? dst[sp[u].o] : T( 0 );
      if ( nsrc == 0 )
         v = value;
      if ( nsrc > 0 && on[u] )
         v += a0 * s0[sp[u].o];
      if ( nsrc > 1 && on[u] )
         v += a1 * s1[sp[u].o];
      if ( nsrc > 2 && on[u] )
         v += a2 * s2[sp[u].o];
      r[u] = v;
   }
#pragma unroll
   for ( int u = 0; u < kPtsPerThread; ++u )
      if ( on[u] )
         dst[sp[u].o] = r[u];
}

template < typename T >
__global__ __launch_bounds__( kBlock ) void kLincomb( const PlaneTile* __restrict__ tiles, int nTiles, int nC, int L, i64 stride,
                                                      const CellInfo* __restrict__ cinfo, int flag, T* __restrict__ dst, int nsrc,
                                                      const T* __restrict__ s0, const T* __restrict__ s1,
                                                      const T* __restrict__ s2, T a0, T a1, T a2, int accumulate, T value )
{
   bLincomb< T >( int( blockIdx.x ), tiles, nTiles, nC, L, stride, cinfo, flag, dst, nsrc, s0, s1, s2, a0, a1, a2, accumulate, value );
}

// Conjugate-gradient updates with scalars held in device memory (so that a captured graph needs no
// host round trip): sc[rsOld], sc[pAp] from the dot kernels. A step is taken only while the residual
// norm is not below tol (tol2 = tol^2), matching CGSolver's early exit.
template < typename T >
__global__ __launch_bounds__( kBlock ) void kCgUpdate( const PlaneTile* __restrict__ tiles, int nTiles, int nC, int L, i64 stride,
                                                       const CellInfo* __restrict__ cinfo, int flag, T* __restrict__ x,
                                                       T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ ap,
                                                       const double* __restrict__ sc, int rsOld, int pAp, double tol2 )
{
   const int lb = remapXcd( blockIdx.x, nTiles * nC );
   if ( lb >= nTiles * nC )
      return;
   const double rs = sc[rsOld];
   const T      al = rs >= tol2 && sc[pAp] != 0.0 ? T( rs / sc[pAp] ) : T( 0 );
   PlaneTile    tile;
   const int    c  = tileCell( tiles, nTiles, lb, tile );
   const uint64_t bw = cinfo[c].bcword;
#pragma unroll
   for ( int u = 0; u < kPtsPerThread; ++u )
   {
      SlotPos sp;
      slotAt( c, tile, 1 << L, stride, tile.s0 + int( threadIdx.x ) + kBlock * u, sp );
      if ( sp.ok && ( int( ( bw >> ( 4 * sp.slot ) ) & 15u ) & flag ) )
      {
         x[sp.o] += al * p[sp.o];
         r[sp.o] -= al * ap[sp.o];
      }
   }
}

// p = r + beta p, beta = sc[rsNew] / sc[rsOld] (0 once converged)
template < typename T >
__global__ __launch_bounds__( kBlock ) void kCgDirection( const PlaneTile* __restrict__ tiles, int nTiles, int nC, int L, i64 stride,
                                                          const CellInfo* __restrict__ cinfo, int flag, T* __restrict__ p,
                                                          const T* __restrict__ r, const double* __restrict__ sc, int rsOld,
                                                          int rsNew, double tol2 )
{
   const int lb = remapXcd( blockIdx.x, nTiles * nC );
   if ( lb >= nTiles * nC )
      return;
   const double o  = sc[rsOld], nw = sc[rsNew];
   const T      be = nw >= tol2 && o > 0.0 ? T( nw / o ) : T( 0 );
   PlaneTile    tile;
   const int    c  = tileCell( tiles, nTiles, lb, tile );
   const uint64_t bw = cinfo[c].bcword;
#pragma unroll
   for ( int u = 0; u < kPtsPerThread; ++u )
   {
      SlotPos sp;
      slotAt( c, tile, 1 << L, stride, tile.s0 + int( threadIdx.x ) + kBlock * u, sp );
      if ( sp.ok && ( int( ( bw >> ( 4 * sp.slot ) ) & 15u ) & flag ) )
         p[sp.o] = r[sp.o] + be * p[sp.o];
   }
}

template < typename T >
__global__ __launch_bounds__( kBlock ) void kMult( const PlaneTile* __restrict__ tiles, int nTiles, int nC, int L, i64 stride,
                                                   const CellInfo* __restrict__ cinfo, int flag, T* dst, const T* a, const T* b )
{
   const int lb = remapXcd( blockIdx.x, nTiles * nC );
   if ( lb >= nTiles * nC )
      return;
   PlaneTile      tile;
   const int      c  = tileCell( tiles, nTiles, lb, tile );
   const uint64_t bw = cinfo[c].bcword;
#pragma unroll
   for ( int u = 0; u < kPtsPerThread; ++u )
   {
      SlotPos sp;
      slotAt( c, tile, 1 << L, stride, tile.s0 + int( threadIdx.x ) + kBlock * u, sp );
      if ( sp.ok && ( int( ( bw >> ( 4 * sp.slot ) ) & 15u ) & flag ) )
         dst[sp.o] = a[sp.o] * b[sp.o];
   }
}

__device__ __forceinline__ double waveSum( double v )
{
#pragma unroll
   for ( int off = 32; off > 0; off >>= 1 )
      v += __shfl_xor( v, off, 64 );
   return v;
}

template < int NT >
__device__ __forceinline__ double blockSum( double v )
{
   __shared__ double red[NT / 64];
   v = waveSum( v );
   if ( ( threadIdx.x & 63 ) == 0 )
      red[threadIdx.x >> 6] = v;
   __syncthreads();
   double t = 0.0;
   if ( threadIdx.x == 0 )
      for ( int k = 0; k < NT / 64; ++k )
         t += red[k];
   return t;
}

template < typename T >
__device__ __forceinline__ void bDot( int bid, const PlaneTile* __restrict__ tiles, int nTiles, int nC, int L, i64 stride,
                                                  const CellInfo* __restrict__ cinfo, int flag, const T* __restrict__ a,
                                                  const T* __restrict__ b, double* __restrict__ partials )
{
   PlaneTile      tile;
   const int      c  = tileCell( tiles, nTiles, bid, tile );
   const uint64_t bw = cinfo[c].bcword;
   const uint32_t om = cinfo[c].ownmask;
   double         v  = 0.0;
#pragma unroll
   for ( int u = 0; u < kPtsPerThread; ++u )
   {
      SlotPos sp;
      slotAt( c, tile, 1 << L, stride, tile.s0 + int( threadIdx.x ) + kBlock * u, sp );
      if ( sp.ok && ( int( ( bw >> ( 4 * sp.slot ) ) & 15u ) & flag ) && ( ( om >> sp.slot ) & 1u ) )
         v += double( a[sp.o] ) * double( b[sp.o] );
   }
   v = blockSum< kBlock >( v );
   if ( threadIdx.x == 0 )
      partials[bid] = v;
}

template < typename T >
__global__ __launch_bounds__( kBlock ) void kDot( const PlaneTile* __restrict__ tiles, int nTiles, int nC, int L, i64 stride,
                                                  const CellInfo* __restrict__ cinfo, int flag, const T* __restrict__ a,
                                                  const T* __restrict__ b, double* __restrict__ partials )
{
   bDot< T >( int( blockIdx.x ), tiles, nTiles, nC, L, stride, cinfo, flag, a, b, partials );
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
   const int lb = remapXcd( blockIdx.x, nTiles * nC );
   if ( lb >= nTiles * nC )
      return;
   PlaneTile tile;
   const int c = tileCell( tiles, nTiles, lb, tile );
#pragma unroll
   for ( int u = 0; u < kPtsPerThread; ++u )
   {
      SlotPos sp;
      slotAt( c, tile, 1 << L, stride, tile.s0 + int( threadIdx.x ) + kBlock * u, sp );
      if ( sp.ok )
         dst[sp.o] = sp.slot == 0 ? cellSt[16 * c + 15] : primInv[cinfo[c].prim[sp.slot]];
   }
}

// ------------------------------------------------------------------ grid transfer
// Prolongation: every fine point is a coarse point (weight 1) or the midpoint of a coarse micro-edge
// (weights 1/2); both parents lie in the same macro-cell, so each replica is computed in its own cell.
template < typename T >
__device__ __forceinline__ void bProlongate( int bid, const PlaneTile* __restrict__ tiles, int nTiles, int nC, int Lf,
                                                         i64 strideF, i64 strideC, const CellInfo* __restrict__ cinfo, int flag,
                                                         T* __restrict__ fine, const T* __restrict__ coarse, int add )
{
   const int lb = remapXcd( bid, nTiles * nC );
   if ( lb >= nTiles * nC )
      return;
   PlaneTile      tile;
   const int      c  = tileCell( tiles, nTiles, lb, tile );
   const uint64_t bw = cinfo[c].bcword;
   const int      wc = ( 1 << ( Lf - 1 ) ) + 1;
   const T*       cs = coarse + i64( c ) * strideC;
   // micro-edge direction of each parity class px + 2 py + 4 pz (the refinement's edges are
   // E, N, SE, TC, BE, BN, TSE up to sign): (1,0,0) (0,1,0) (1,-1,0) (0,0,1) (1,0,-1) (0,1,-1) (1,-1,1)
   constexpr int EX[8] = { 0, 1, 0, 1, 0, 1, 0, 1 };
   constexpr int EY[8] = { 0, 0, 1, -1, 0, 0, 1, -1 };
   constexpr int EZ[8] = { 0, 0, 0, 0, 1, -1, -1, 1 };
   SlotPos       sp[kPtsPerThread];
   bool          on[kPtsPerThread];
   T             va[kPtsPerThread], vb[kPtsPerThread], vf[kPtsPerThread];
#pragma unroll
   for ( int u = 0; u < kPtsPerThread; ++u )
   {
      slotAt( c, tile, 1 << Lf, strideF, tile.s0 + int( threadIdx.x ) + kBlock * u, sp[u] );
      // cell-interior points are prolongated whatever the flag (the reference's cell loop has no flag test,
      // P1toP1LinearProlongation.cpp:208-260); points of macro-faces, -edges and -vertices follow the flag
      on[u]        = sp[u].ok && ( sp[u].slot == 0 || ( int( ( bw >> ( 4 * sp[u].slot ) ) & 15u ) & flag ) );
      const int pc = ( sp[u].x & 1 ) | ( ( sp[u].y & 1 ) << 1 ) | ( ( sp[u].z & 1 ) << 2 );
      const int ex = EX[pc], ey = EY[pc], ez = EZ[pc];
      va[u] = on[u] ? cs[cidx( wc, ( sp[u].x - ex ) >> 1, ( sp[u].y - ey ) >> 1, ( sp[u].z - ez ) >> 1 )] : T( 0 );
      vb[u] = ( on[u] && pc ) ? cs[cidx( wc, ( sp[u].x + ex ) >> 1, ( sp[u].y + ey ) >> 1, ( sp[u].z + ez ) >> 1 )] : T( 0 );
      vf[u] = ( on[u] && add ) ? fine[sp[u].o] : T( 0 );
   }
#pragma unroll
   for ( int u = 0; u < kPtsPerThread; ++u )
   {
      if ( !on[u] )
         continue;
      const int pc = ( sp[u].x & 1 ) | ( ( sp[u].y & 1 ) << 1 ) | ( ( sp[u].z & 1 ) << 2 );
      const T   v  = pc ? T( 0.5 ) * ( va[u] + vb[u] ) : va[u];
      fine[sp[u].o] = add ? vf[u] + v : v;
   }
}

template < typename T >
__global__ __launch_bounds__( kBlock ) void kProlongate( const PlaneTile* __restrict__ tiles, int nTiles, int nC, int Lf,
                                                         i64 strideF, i64 strideC, const CellInfo* __restrict__ cinfo, int flag,
                                                         T* __restrict__ fine, const T* __restrict__ coarse, int add )
{
   bProlongate< T >( int( blockIdx.x ), tiles, nTiles, nC, Lf, strideF, strideC, cinfo, flag, fine, coarse, add );
}

// Restriction, coarse points inside a cell: all 14 fine neighbours of 2q are cell-interior.
template < typename T >
__device__ __forceinline__ void bRestrictCell( int bid, const PlaneTile* __restrict__ tiles, int tile0, int nTiles, int nC,
                                                           int Lc, i64 strideC, i64 strideF, T* __restrict__ coarse,
                                                           const T* __restrict__ fine )
{
   const int nb = nTiles * nC;
   const int lb = remapXcd( bid, nb );
   if ( lb >= nb )
      return;
   const int       c    = lb / nTiles;
   const PlaneTile tile = tiles[tile0 + ( lb - c * nTiles )];
   const int       Nc = 1 << Lc, wc = Nc + 1, z = tile.z, W = wc - z;
   const int       wf = 2 * Nc + 1;
   const T*        fs = fine + i64( c ) * strideF;
#pragma unroll
   for ( int u = 0; u < kPtsPerThread; ++u )
   {
      const int s = tile.s0 + int( threadIdx.x ) + kBlock * u;
      if ( s >= ( W * ( W + 1 ) ) / 2 )
         continue;
      int x, y;
      planeXY( W, s, x, y );
      if ( x < 1 || y < 1 || x + y + z > Nc - 1 )
         continue;
      const int fx = 2 * x, fy = 2 * y, fz = 2 * z;
      T         v  = fs[cidx( wf, fx, fy, fz )];
      T         h  = T( 0 );
#pragma unroll
      for ( int d = 1; d < NDIR; ++d )
         h += fs[cidx( wf, fx + dirX( d ), fy + dirY( d ), fz + dirZ( d ) )];
      coarse[i64( c ) * strideC + ( S3i( wc ) - S3i( wc - z ) ) + s] = v + T( 0.5 ) * h;
   }
}

template < typename T >
__global__ __launch_bounds__( kBlock ) void kRestrictCell( const PlaneTile* __restrict__ tiles, int tile0, int nTiles, int nC,
                                                           int Lc, i64 strideC, i64 strideF, T* __restrict__ coarse,
                                                           const T* __restrict__ fine )
{
   bRestrictCell< T >( int( blockIdx.x ), tiles, tile0, nTiles, nC, Lc, strideC, strideF, coarse, fine );
}

// Restriction, coarse points on macro-faces/-edges/-vertices: sum over the incident cells of the
// in-cell fine neighbours, each scaled by 1 / (#cells sharing the fine point's primitive).
template < typename T >
__device__ __forceinline__ void bRestrictLower( int bid, const LowerChunk* __restrict__ chunks, int nChunks,
                                                            const LowerPrim* __restrict__ prims,
                                                            const LowerEntry* __restrict__ entries,
                                                            const CellInfo* __restrict__ cinfo, int Lc, i64 strideC,
                                                            i64 strideF, int flag, T* __restrict__ coarse,
                                                            const T* __restrict__ fine )
{
   const int ci = bid * ( kBlock / 64 ) + int( threadIdx.x >> 6 );
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

template < typename T >
__global__ __launch_bounds__( kBlock ) void kRestrictLower( const LowerChunk* __restrict__ chunks, int nChunks,
                                                            const LowerPrim* __restrict__ prims,
                                                            const LowerEntry* __restrict__ entries,
                                                            const CellInfo* __restrict__ cinfo, int Lc, i64 strideC,
                                                            i64 strideF, int flag, T* __restrict__ coarse,
                                                            const T* __restrict__ fine )
{
   bRestrictLower< T >( int( blockIdx.x ), chunks, nChunks, prims, entries, cinfo, Lc, strideC, strideF, flag, coarse, fine );
}

} // namespace hyg
