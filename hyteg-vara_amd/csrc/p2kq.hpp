// On-the-fly variable-coefficient P2 kernels (P2ElementwiseOperator< p2_div_k_grad_affine_q4 >,
// P2ElementwiseOperator.cpp:100-310): cell-interior and cell-face points. Compiled once per value type
// (p2kq_f64.hip, p2kq_f32.hip).
#pragma once
#include "p2_common.hpp"

namespace hyg {

// ------------------------------------------------------------------ variable coefficient, assembled on the fly
// Each point sums, over the micro-cells that have it as node l (known at compile time per parity class; for a
// cell-face point only those on the cell's side), row l of sum_q k_q M_{t,q} (the rows staged in LDS by p2StageRows,
// read with wave-uniform loads) times the ten node values, for up to three components at once, so a vector
// function's components share the assembly. No per-point stencil array is kept, which keeps the register footprint
// small. Cell-interior points go by class sub-lattice as the stored stencils do, cell-face points per face and
// sub-lattice (one-sided sums into each component's partial-sum block), the points on cell edges and vertices
// element by element (kP2Apply).
// Each point sums, over the micro-cells that have it as node l (on the cell's side of face F for a face point), row
// l of sum_q k_q M_{t,q} times the ten node values, for up to three components (val[k] += ...). One micro-cell per
// iteration of a loop that is not unrolled: the micro-cell's data are wave-uniform (scalar loads), and no loads of
// one micro-cell can be hoisted above another's, so the register footprint is that of one micro-cell. rb: starts of
// the rows py + dy of the planes pz + dz (|dy|, |dz| <= 2) relative to the cell, base = cell offset + px.
template < typename T, int CLS, int F >
__device__ __forceinline__ void p2KQDirect( int c, int N, int px, int py, int pz, const T* __restrict__ kq, i64 kqStride,
                                            const T* __restrict__ Ms, const int ( &rb )[5][5], i64 base, const P2Comps< T >& cp, int nc,
                                            T ( &val )[3], bool onFace = false )
{
   const int          Nc = N >> 1, SA = S3i( Nc + 1 );
   const P2KQEntries& E  = kP2KQTab.c[CLS][F + 1];
   int                rf[25];
#pragma unroll
   for ( int a = 0; a < 25; ++a )
      rf[a] = rb[a / 5][a % 5];
#pragma unroll 1
   for ( int k = 0; k < E.n; ++k )
   {
      const P2KQEntry& e  = E.e[k];
      const int        ax = ( px - e.nd[0] ) >> 1, ay = ( py - e.nd[1] ) >> 1, az = ( pz - e.nd[2] ) >> 1;
      // a point on a cell face takes the micro-cells inside the cell (its one-sided sum)
      if ( onFace && ( ax < 0 || ay < 0 || az < 0 || ax + ay + az + e.mcmax > Nc ) )
         continue;
      const T* kv = kq + i64( c * 6 + e.t ) * SA + cidx( Nc + 1, ax, ay, az );
      T        kk[11];
#pragma unroll
      for ( int q = 0; q < 11; ++q )
         kk[q] = kv[i64( q ) * kqStride];
#pragma unroll
      for ( int jj = 0; jj < 10; ++jj )
      {
         const T* row = Ms + ( e.ord * 10 + jj ) * kP2RowPad;
         T        r   = T( 0 );
#pragma unroll
         for ( int q = 0; q < 11; ++q )
            r += kk[q] * row[q];
         const int nk = e.node[jj];
         const i64 o  = base + rf[nk & 255] + ( ( nk >> 8 ) - 2 );
         // explicit fused multiply-adds: every component (and the single-component pass) rounds alike
         val[0] = fma( r, cp.x[0][o], val[0] );
         if ( nc > 1 )
         {
            val[1] = fma( r, cp.x[1][o], val[1] );
            val[2] = fma( r, cp.x[2][o], val[2] );
         }
      }
   }
}
template < typename T >
__device__ __forceinline__ void p2EmitInterior( int mode, int nc, const T ( &val )[3], i64 q, const P2Comps< T >& cp,
                                                const T* __restrict__ dinv, T a )
{
#pragma unroll
   for ( int k = 0; k < 3; ++k )
   {
      if ( k >= nc )
         break;
      T r0 = T( 0 ), r1 = T( 0 );
      switch ( mode )
      {
      case MODE_APPLY: epilogue< T, MODE_APPLY >( val[k], T( 0 ), T( 0 ), T( 0 ), T( 1 ), a, r0, r1 ); break;
      case MODE_APPLY_ADD:
      {
         const T o = cp.out0[k][q];
         epilogue< T, MODE_APPLY_ADD >( val[k], T( 0 ), o, o, T( 1 ), a, r0, r1 );
         break;
      }
      case MODE_RESIDUAL:
      {
         const T i1 = cp.in1[k][q];
         epilogue< T, MODE_RESIDUAL >( val[k], T( 0 ), i1, i1, T( 1 ), a, r0, r1 );
         break;
      }
      case MODE_JACOBI:
      {
         const T i1 = cp.in1[k][q];
         epilogue< T, MODE_JACOBI >( val[k], cp.x[k][q], i1, i1, dinv[q], a, r0, r1 );
         break;
      }
      case MODE_CHEB0:
      {
         const T i1 = cp.in1[k][q];
         epilogue< T, MODE_CHEB0 >( val[k], cp.x[k][q], i1, i1, dinv[q], a, r0, r1 );
         break;
      }
      default:
      {
         const T i1 = cp.in1[k][q];
         epilogue< T, MODE_CHEBK >( val[k], cp.x[k][q], i1, i1, dinv[q], a, r0, r1 );
         break;
      }
      }
      cp.out0[k][q] = r0;
      if ( mode == MODE_CHEB0 || mode == MODE_CHEBK )
         cp.out1[k][q] = r1;
   }
}
// All points of class CLS of a cell that are not on a cell edge: the class sub-lattice {2 q + kappa : 2 |q| + |kappa| <= N}
// (p2ClassExtent), so the points on the x = 0 and diagonal faces sit in the waves of their rows' interior points and
// are read and written with them. Interior points get the mode's epilogue, face points store their one-sided sums.
template < typename T, int CLS >
__device__ __forceinline__ void p2KQIntPt( int c, int s, int N, i64 stride, const T* __restrict__ el, const T* __restrict__ kq,
                                           i64 kqStride, const T* __restrict__ dinv, const P2Comps< T >& cp, int nc, int mode, T a,
                                           T* __restrict__ Ms, const CellFaceSlot* __restrict__ cfs, int flag, T* __restrict__ part,
                                           i64 partStride )
{
   constexpr int kx = CLS & 1, ky = ( CLS >> 1 ) & 1, kz = ( CLS >> 2 ) & 1;
   p2StageRows< T, CLS >( el, c, Ms );
   const int M = p2ClassExtent( N, CLS );
   if ( s >= S3i( M + 1 ) )
      return;
   int qx, qy, qz;
   tetXYZ( M + 1, s, qx, qy, qz );
   const int px = 2 * qx + kx, py = 2 * qy + ky, pz = 2 * qz + kz;
   const int slot = classifySlot( N, px, py, pz );
   if ( slot > 4 || ( slot == 0 && !( flag & 1 ) ) ) // cell edges and vertices: kP2Apply
      return;
   i64 pi = 0;
   if ( slot > 0 )
   {
      const CellFaceSlot fs = cfs[4 * c + slot - 1];
      if ( !( fs.bc & flag ) )
         return;
      pi = i64( fs.block ) * ( N + 1 ) * ( N + 1 ) + facePartIndex( slot - 1, N, px, py, pz );
   }
   const int w = N + 1;
   int       rb[5][5];
#pragma unroll
   for ( int dz = -2; dz <= 2; ++dz )
   {
      const int Wz = w - ( pz + dz );
      const int pb = S3i( w ) - S3i( Wz );
#pragma unroll
      for ( int dy = -2; dy <= 2; ++dy )
         rb[dz + 2][dy + 2] = pb + rowStart( Wz, py + dy );
   }
   T val[3] = { T( 0 ), T( 0 ), T( 0 ) };
   p2KQDirect< T, CLS, -1 >( c, N, px, py, pz, kq, kqStride, Ms, rb, i64( c ) * stride + px, cp, nc, val, slot > 0 );
   if ( slot == 0 )
      p2EmitInterior< T >( mode, nc, val, i64( c ) * stride + rb[2][2] + px, cp, dinv, a );
   else
   {
#pragma unroll
      for ( int k = 0; k < 3; ++k )
      {
         if ( k >= nc )
            break;
         part[i64( k ) * partStride + pi] = val[k];
      }
   }
}
template < typename T, int CLS >
__device__ __forceinline__ void p2KQIntDispatch( int cls, int c, int s, int N, i64 stride, const T* el, const T* kq, i64 kqStride,
                                                 const T* dinv, const P2Comps< T >& cp, int nc, int mode, T a, T* Ms,
                                                 const CellFaceSlot* cfs, int flag, T* part, i64 partStride )
{
   if ( cls == CLS )
      p2KQIntPt< T, CLS >( c, s, N, stride, el, kq, kqStride, dinv, cp, nc, mode, a, Ms, cfs, flag, part, partStride );
   else if constexpr ( CLS < 7 )
      p2KQIntDispatch< T, CLS + 1 >( cls, c, s, N, stride, el, kq, kqStride, dinv, cp, nc, mode, a, Ms, cfs, flag, part, partStride );
}
template < typename T >
__global__ __launch_bounds__( kBlock ) void kP2KQInt( int L, i64 stride, P2ClassTiles ct, const T* __restrict__ el,
                                                      const T* __restrict__ kq, i64 kqStride, const T* __restrict__ dinv,
                                                      P2Comps< T > cp, int nc, int mode, T a, const CellFaceSlot* __restrict__ cfs,
                                                      int flag, T* __restrict__ part, i64 partStride, int nC )
{
   __shared__ T Ms[kP2MaxPairs * 10 * kP2RowPad];
   const int    per = ct.start[8];
   // each XCD takes a contiguous range of cells (the hardware deals workgroups out round-robin over the 8 XCDs), so
   // a cell's coefficient values and source stay in one XCD's L2 while its class blocks run
   const int b = remapXcd( int( blockIdx.x ), nC * per );
   if ( b >= nC * per ) // the whole block (grid padded to a multiple of 8)
      return;
   const int c = b / per;
   const int t = b - c * per;
   int          cls = 0;
   while ( cls < 7 && t >= ct.start[cls + 1] )
      ++cls;
   p2KQIntDispatch< T, 0 >( cls, c, ( t - ct.start[cls] ) * kBlock + int( threadIdx.x ), 1 << L, stride, el, kq, kqStride, dinv, cp,
                            nc, mode, a, Ms, cfs, flag, part, partStride );
}
template < typename T >
void launchP2KQPoints( int L, int nC, i64 stride, const T* el, const T* kq, i64 kqStride, const T* dinv, const P2Comps< T >& cp, int nc,
                       int mode, T a, const CellFaceSlot* cfs, int flag, T* part, i64 partStride, hipStream_t s )
{
   const int    N = 1 << L;
   P2ClassTiles ct;
   ct.start[0] = 0;
   for ( int k = 0; k < 8; ++k )
      ct.start[k + 1] = ct.start[k] + ceilDiv( S3i( p2ClassExtent( N, k ) + 1 ), kBlock );
   hipLaunchKernelGGL( ( kP2KQInt< T > ), dim3( xcdGrid( nC * ct.start[8] ) ), dim3( kBlock ), 0, s, L, stride, ct, el, kq, kqStride, dinv,
                       cp, nc, mode, a, cfs, flag, part, partStride, nC );
}

} // namespace hyg
