// Face values from the one-sided partial sums of their cells (shared by the constant-coefficient and
// the elementwise operators): sum the partial sums, apply the epilogue, write every replica.
#pragma once

#include "device_common.hpp"

namespace hyg {

template < typename T, int MODE >
__global__ __launch_bounds__( kBlock ) void kFaceCombine( const LowerChunk* __restrict__ chunks, int nChunks,
                                                          const LowerPrim* __restrict__ prims,
                                                          const LowerEntry* __restrict__ entries, int firstFacePrim,
                                                          const T* __restrict__ part, const T* __restrict__ primInv,
                                                          const T* __restrict__ dinvArr, int L,
                                                          i64 stride, int flag, const T* __restrict__ x,
                                                          const T* __restrict__ in1, T* __restrict__ out0, T* __restrict__ out1, T a )
{
   const int nb = ( nChunks + 3 ) / 4;
   const int ci = remapXcd( int( blockIdx.x ), nb ) * ( kBlock / 64 ) + __builtin_amdgcn_readfirstlane( int( threadIdx.x >> 6 ) );
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

} // namespace hyg
