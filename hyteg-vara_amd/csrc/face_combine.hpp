// Kernels and helpers shared by the constant-coefficient, the elementwise and the P2 operators: face and
// edge/vertex values from the one-sided partial sums of their cells (sum, epilogue, write every replica),
// the pointwise inverse, and the enumeration of the points of a macro-cell.
#pragma once

#include "device_common.hpp"

namespace hyg {

template < typename T, int MODE >
__device__ __forceinline__ void bFaceCombine( int bid, const LowerChunk* __restrict__ chunks, int nChunks,
                                                          const LowerPrim* __restrict__ prims,
                                                          const LowerEntry* __restrict__ entries, int firstFacePrim,
                                                          const T* __restrict__ part, const T* __restrict__ primInv,
                                                          const T* __restrict__ dinvArr, int L,
                                                          i64 stride, int flag, const T* __restrict__ x,
                                                          const T* __restrict__ in1, T* __restrict__ out0, T* __restrict__ out1, T a )
{
   const int nb = ( nChunks + 3 ) / 4;
   const int ci = remapXcd( bid, nb ) * ( kBlock / 64 ) + __builtin_amdgcn_readfirstlane( int( threadIdx.x >> 6 ) );
   if ( ci >= nChunks )
      return;
   const LowerChunk ch = chunks[ci];
   const LowerPrim  p  = prims[ch.prim];
   const int        lane = int( threadIdx.x & 63 );
   if ( !( p.bc & flag ) || lane >= ch.n )
      return;
   const int N = 1 << L, NP = N + 1, w = N + 1;
   const int i = ch.i0 + lane, j = ch.j, f = ch.prim - firstFacePrim;
   T         Ax = T( 0 );
   i64       o[2];
   for ( int k = 0; k < p.nCells; ++k )
   {
      const LowerEntry e = entries[p.entry0 + k];
      int              qx, qy, qz;
      entryPos( e, N, i, j, qx, qy, qz );
      o[k] = i64( e.cell ) * stride + cidx( w, qx, qy, qz );
      Ax += part[i64( 2 * f + k ) * NP * NP + facePartIndex( entrySlot( e ), N, qx, qy, qz )];
   }
   T xc = T( 0 ), i1 = T( 0 ), ov = T( 0 ), r0, r1;
   if constexpr ( MODE == MODE_JACOBI || MODE == MODE_CHEB0 )
      xc = x[o[0]];
   if constexpr ( MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
      i1 = in1[o[0]];
   if constexpr ( MODE == MODE_APPLY_ADD )
      ov = out0[o[0]];
   const T dinv = dinvArr ? dinvArr[o[0]] : ( primInv ? primInv[ch.prim] : T( 1 ) );
   epilogue< T, MODE >( Ax, xc, i1, ov, dinv, a, r0, r1 );
   for ( int k = 0; k < p.nCells; ++k )
   {
      out0[o[k]] = r0;
      if constexpr ( MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
         out1[o[k]] = r1;
   }
}

template < typename T, int MODE >
__global__ __launch_bounds__( kBlock ) void kFaceCombine( const LowerChunk* __restrict__ chunks, int nChunks,
                                                          const LowerPrim* __restrict__ prims,
                                                          const LowerEntry* __restrict__ entries, int firstFacePrim,
                                                          const T* __restrict__ part, const T* __restrict__ primInv,
                                                          const T* __restrict__ dinvArr, int L,
                                                          i64 stride, int flag, const T* __restrict__ x,
                                                          const T* __restrict__ in1, T* __restrict__ out0, T* __restrict__ out1, T a )
{
   bFaceCombine< T, MODE >( int( blockIdx.x ), chunks, nChunks, prims, entries, firstFacePrim, part, primInv, dinvArr, L, stride, flag, x, in1, out0, out1, a );
}

// colour of point (i, j) of a macro-face, -edge or -vertex in the multi-colour SOR ordering: the sum over the
// vertices of a cell containing it of (vertex colour) x (barycentric lattice coordinate), mod the number of
// colours; every containing cell gives the same value, the first entry's is used
__device__ __forceinline__ int sorColour( const CellInfo* __restrict__ cinfo, const LowerEntry& e, int N, int i, int j, int m )
{
   int qx, qy, qz;
   entryPos( e, N, i, j, qx, qy, qz );
   const CellInfo& ci = cinfo[e.cell];
   return ( ci.colour[0] * ( N - qx - qy - qz ) + ci.colour[1] * qx + ci.colour[2] * qy + ci.colour[3] * qz ) % m;
}

// one colour pass of multi-colour SOR at macro-face points (constant operator): x += a D^-1 (b - A x) at the
// points of colour sorK, in place, every replica
template < typename T >
__global__ __launch_bounds__( kBlock ) void kFaceCombineSor( const LowerChunk* __restrict__ chunks, int nChunks,
                                                             const LowerPrim* __restrict__ prims, const LowerEntry* __restrict__ entries,
                                                             int firstFacePrim, const T* __restrict__ part, const T* __restrict__ primInv,
                                                             const CellInfo* __restrict__ cinfo, int sorK, int sorM, int L, i64 stride,
                                                             int flag, T* __restrict__ x, const T* __restrict__ b, T a )
{
   const int nb = ( nChunks + 3 ) / 4;
   const int ci = remapXcd( int( blockIdx.x ), nb ) * ( kBlock / 64 ) + __builtin_amdgcn_readfirstlane( int( threadIdx.x >> 6 ) );
   if ( ci >= nChunks )
      return;
   const LowerChunk ch   = chunks[ci];
   const LowerPrim  p    = prims[ch.prim];
   const int        lane = int( threadIdx.x & 63 );
   if ( !( p.bc & flag ) || lane >= ch.n )
      return;
   const int N = 1 << L, NP = N + 1, w = N + 1;
   const int i = ch.i0 + lane, j = ch.j, f = ch.prim - firstFacePrim;
   if ( sorColour( cinfo, entries[p.entry0], N, i, j, sorM ) != sorK )
      return;
   T   Ax = T( 0 );
   i64 o[2];
   for ( int k = 0; k < p.nCells; ++k )
   {
      const LowerEntry e = entries[p.entry0 + k];
      int              qx, qy, qz;
      entryPos( e, N, i, j, qx, qy, qz );
      o[k] = i64( e.cell ) * stride + cidx( w, qx, qy, qz );
      Ax += part[i64( 2 * f + k ) * NP * NP + facePartIndex( entrySlot( e ), N, qx, qy, qz )];
   }
   const T v = x[o[0]] + a * primInv[ch.prim] * ( b[o[0]] - Ax );
   for ( int k = 0; k < p.nCells; ++k )
      x[o[k]] = v;
}

// micro-cell types WHITE_UP, BLUE_UP, GREEN_UP, WHITE_DOWN, BLUE_DOWN, GREEN_DOWN: vertex offsets
// from the micro-cell's anchor (volumedofspace/CellDoFIndexing.hpp:155-198)
constexpr int kMC[6][4][3] = { { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } },
                                  { { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 }, { 1, 0, 1 } },
                                  { { 1, 0, 0 }, { 0, 1, 0 }, { 1, 0, 1 }, { 0, 0, 1 } },
                                  { { 1, 1, 0 }, { 1, 1, 1 }, { 0, 1, 1 }, { 1, 0, 1 } },
                                  { { 1, 0, 1 }, { 0, 1, 1 }, { 0, 0, 1 }, { 0, 1, 0 } },
                                  { { 0, 1, 0 }, { 1, 1, 0 }, { 1, 0, 1 }, { 0, 1, 1 } } };
// largest coordinate sum of a type's vertex offsets: the element anchored at a lies in the cell iff
// a >= 0 and |a| + kMCMax[t] <= N
constexpr int kMCMax[6] = { 1, 2, 2, 3, 2, 2 };

// (x, y, z) of the s-th point of a tetrahedral lattice with w points per edge
__device__ __forceinline__ void tetXYZ( int w, int s, int& x, int& y, int& z )
{
   // plane z: S3(w) - S3(w - z) <= s < S3(w) - S3(w - z - 1)
   int lo = 0, hi = w - 1;
   const int S = S3i( w );
   while ( lo < hi )
   {
      const int mid = ( lo + hi + 1 ) >> 1;
      if ( S - S3i( w - mid ) <= s )
         lo = mid;
      else
         hi = mid - 1;
   }
   z = lo;
   planeXY( w - z, s - ( S - S3i( w - z ) ), x, y );
}

// ------------------------------------------------------------------ edges and vertices: combine
template < typename T, int MODE >
__device__ __forceinline__ void bLowCombine( int bid, const LowerChunk* __restrict__ chunks, int nChunks,
                                                         const LowerPrim* __restrict__ prims, const LowerEntry* __restrict__ entries,
                                                         const T* __restrict__ lowPart, const T* __restrict__ dinv, int L,
                                                         i64 stride, int flag, const T* __restrict__ x, const T* __restrict__ in1,
                                                         T* __restrict__ out0, T* __restrict__ out1, T a )
{
   const int ci = bid * ( kBlock / 64 ) + int( threadIdx.x >> 6 );
   if ( ci >= nChunks )
      return;
   const LowerChunk ch   = chunks[ci];
   const LowerPrim  p    = prims[ch.prim];
   const int        lane = int( threadIdx.x & 63 );
   if ( !( p.bc & flag ) || lane >= ch.n )
      return;
   const int N = 1 << L, w = N + 1, NP = N + 1;
   const int i = ch.i0 + lane;
   T         Ax = T( 0 );
   for ( int k = 0; k < p.nCells; ++k )
      Ax += lowPart[i64( p.entry0 + k ) * NP + i];
   int qx, qy, qz;
   entryPos( entries[p.entry0], N, i, 0, qx, qy, qz );
   const i64 o0 = i64( entries[p.entry0].cell ) * stride + cidx( w, qx, qy, qz );
   T         xc = T( 0 ), i1 = T( 0 ), r0, r1;
   if constexpr ( MODE == MODE_JACOBI || MODE == MODE_CHEB0 )
      xc = x[o0];
   if constexpr ( MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
      i1 = in1[o0];
   if constexpr ( MODE == MODE_APPLY_ADD )
      i1 = out0[o0];
   epilogue< T, MODE >( Ax, xc, i1, i1, dinv ? dinv[o0] : T( 1 ), a, r0, r1 );
   for ( int k = 0; k < p.nCells; ++k )
   {
      const LowerEntry e = entries[p.entry0 + k];
      entryPos( e, N, i, 0, qx, qy, qz );
      const i64 o = i64( e.cell ) * stride + cidx( w, qx, qy, qz );
      out0[o]     = r0;
      if constexpr ( MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
         out1[o] = r1;
   }
}

template < typename T, int MODE >
__global__ __launch_bounds__( kBlock ) void kLowCombine( const LowerChunk* __restrict__ chunks, int nChunks,
                                                         const LowerPrim* __restrict__ prims, const LowerEntry* __restrict__ entries,
                                                         const T* __restrict__ lowPart, const T* __restrict__ dinv, int L,
                                                         i64 stride, int flag, const T* __restrict__ x, const T* __restrict__ in1,
                                                         T* __restrict__ out0, T* __restrict__ out1, T a )
{
   bLowCombine< T, MODE >( int( blockIdx.x ), chunks, nChunks, prims, entries, lowPart, dinv, L, stride, flag, x, in1, out0, out1, a );
}

// 1 / d of every value (0 stays 0: padding and unused slots)
template < typename T >
__global__ void kInvertAll( i64 n, T* __restrict__ d )
{
   for ( i64 i = i64( blockIdx.x ) * blockDim.x + threadIdx.x; i < n; i += i64( gridDim.x ) * blockDim.x )
      d[i] = d[i] != T( 0 ) ? T( 1 ) / d[i] : T( 0 );
}

} // namespace hyg
