// Device code of the constant P1 stencil sweep (kernel bodies shared by the single-problem and the batched launches).
#pragma once

#include "face_combine.hpp"

namespace hyg {

// ------------------------------------------------------------------ cell interior, z-marching wavefronts
// R output rows and P points per lane: one plane of the register ring holds rows y0-1 .. y0+R of x
// and rows y0 .. y0+R-1 of the epilogue input; lane l holds x = x0-1+P*l .. x0-1+P*l+P-1.
template < typename T, int R, int P >
struct PlaneRows
{
   T r[R + 2][P];
   T e[R][P];
};

// directions without an x shift and with x+1 (E group)
constexpr int kDirU[7] = { 0, 3, 4, 7, 9, 11, 12 };
constexpr int kDirE[4] = { 2, 6, 10, 13 };

// one row of a stencil sweep: the 7 unshifted, 4 x-1 (W, NW, TW, BNW) and 4 x+1 (E, SE, TSE, BE)
// terms for the P points of a lane; wu / we are the per-lane weights of the first point's
// unshifted terms and of the E-group terms formed for the first point (see cellZ).
template < typename T, int P >
__device__ __forceinline__ void stencilRow( const T* w, const T* wu, const T* we, const T ( &bc )[P], const T ( &bn )[P],
                                            const T ( &bs )[P], const T ( &cc )[P], const T ( &cs )[P], const T ( &ac )[P],
                                            const T ( &an )[P], T ( &ax )[P] )
{
   if constexpr ( P == 1 )
   {
      T a = wu[0] * bc[0];
      a += wu[1] * bn[0];
      a += wu[2] * bs[0];
      a += wu[3] * cc[0];
      a += wu[4] * cs[0];
      a += wu[5] * ac[0];
      a += wu[6] * an[0];
      T sr = w[1] * bc[0];
      sr += w[5] * bn[0];
      sr += w[8] * cc[0];
      sr += w[14] * an[0];
      T sl = we[0] * bc[0];
      sl += we[1] * bs[0];
      sl += we[2] * cs[0];
      sl += we[3] * ac[0];
      a += laneMove< kDppWaveShr1 >( sr );
      a += laneMove< kDppWaveShl1 >( sl );
      ax[0] = a;
   }
   else
   {
      // point 0: x-1 is point 1 of lane-1 (shifted), x+1 is point 1 of this lane;
      // point 1: x-1 is point 0 of this lane, x+1 is point 0 of lane+1 (shifted)
      T a0 = wu[0] * bc[0];
      a0 += wu[1] * bn[0];
      a0 += wu[2] * bs[0];
      a0 += wu[3] * cc[0];
      a0 += wu[4] * cs[0];
      a0 += wu[5] * ac[0];
      a0 += wu[6] * an[0];
      a0 += we[0] * bc[1];
      a0 += we[1] * bs[1];
      a0 += we[2] * cs[1];
      a0 += we[3] * ac[1];
      T a1 = w[0] * bc[1];
      a1 += w[3] * bn[1];
      a1 += w[4] * bs[1];
      a1 += w[7] * cc[1];
      a1 += w[9] * cs[1];
      a1 += w[11] * ac[1];
      a1 += w[12] * an[1];
      a1 += w[1] * bc[0];
      a1 += w[5] * bn[0];
      a1 += w[8] * cc[0];
      a1 += w[14] * an[0];
      T sr = w[1] * bc[1];
      sr += w[5] * bn[1];
      sr += w[8] * cc[1];
      sr += w[14] * an[1];
      T sl = w[2] * bc[0];
      sl += w[6] * bs[0];
      sl += w[10] * cs[0];
      sl += w[13] * ac[0];
      ax[0] = a0 + laneMove< kDppWaveShr1 >( sr );
      ax[1] = a1 + laneMove< kDppWaveShl1 >( sl );
   }
}

template < typename T, int MODE, int R, int P, int D >
__device__ __forceinline__ void cellZ( int bid, const ZSeg* __restrict__ segs, int nSegs, int nC, int L, i64 stride,
                                                    const T* __restrict__ cellSt, const CellFaceSlot* __restrict__ cfs,
                                                    const T* __restrict__ entryW, T* __restrict__ part, int flag,
                                                    const T* __restrict__ x, const T* __restrict__ in1, T* __restrict__ out0,
                                                    T* __restrict__ out1, T a, const CellInfo* __restrict__ cinfo, int sorK,
                                                    int sorM )
{
   constexpr bool kIn1 = MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK || MODE == MODE_SOR;
   constexpr bool kE   = kIn1 || MODE == MODE_APPLY_ADD;
   constexpr bool kTwo = MODE == MODE_CHEB0 || MODE == MODE_CHEBK;
   constexpr int  B    = int( sizeof( T ) );
   constexpr int  WO   = 64 * P - 2; // outputs per window: x0 .. x0+WO-1
   const int      nbx  = ( nSegs + 3 ) / 4; // blocks per cell
   const int      nb   = nbx * nC;
   const int      lb   = remapXcd( bid, nb );
   if ( lb >= nb )
      return;
   const int c  = lb / nbx;
   const int si = ( lb - c * nbx ) * 4 + __builtin_amdgcn_readfirstlane( int( threadIdx.x ) >> 6 );
   if ( si >= nSegs )
      return;
   const ZSeg     sg    = segs[si];
   const int      N     = 1 << L, NP = N + 1, y0 = sg.y, z0 = sg.z0, z1 = sg.z1, x0 = sg.x0;
   const unsigned lane  = threadIdx.x & 63u;
   const int      d0    = P * int( lane ); // window offset of the lane's first point (x = x0-1+d0)
   const unsigned vo    = unsigned( d0 * B );
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
   const bool interior = ( flag & 1 ) != 0;
   // multi-colour SOR: point (x, y, z) has the colour (sb + s1 x + s2 y + s3 z) mod sorM
   int s1 = 0, s2 = 0, s3 = 0, sb = 0;
   if constexpr ( MODE == MODE_SOR )
   {
      const CellInfo& ci = cinfo[c];
      s1                 = ( ci.colour[1] - ci.colour[0] + sorM ) % sorM;
      s2                 = ( ci.colour[2] - ci.colour[0] + sorM ) % sorM;
      s3                 = ( ci.colour[3] - ci.colour[0] + sorM ) % sorM;
      sb                 = ( ci.colour[0] * ( N % sorM ) ) % sorM;
   }

   // x = 0 face (slot 2): the first point of lane 0 in the windows starting at a row start, with that
   // face's one-sided weights. Its E-group terms are formed on lane 1 (P = 1, then shifted) or on
   // lane 0 itself (P = 2), which therefore holds that face's E-group weights.
   const CellFaceSlot fx   = cfs[4 * c + 2];
   const bool         onX0 = x0 == 1 && ( fx.bc & flag ) != 0;
   T                  wu[7], we[4];
   {
      const T* wf = entryW + 16 * i64( onX0 ? fx.entry : 0 );
#pragma unroll
      for ( int k = 0; k < 7; ++k )
         wu[k] = ( onX0 && lane == 0u ) ? wf[kDirU[k]] : w[kDirU[k]];
#pragma unroll
      for ( int k = 0; k < 4; ++k )
         we[k] = ( onX0 && lane == ( P == 1 ? 1u : 0u ) ) ? wf[kDirE[k]] : w[kDirE[k]];
   }
   // diagonal face (slot 3): the row ends x = N-y-z of rows y0..y0+R-1 inside the output window
   const CellFaceSlot fd     = cfs[4 * c + 3];
   const int          xeHi   = N - y0 - z0, xeLo = N - ( y0 + R - 1 ) - ( z1 - 1 );
   const bool         onDiag = ( fd.bc & flag ) != 0 && xeHi >= x0 && xeLo <= x0 + WO - 1;
   T                  wd[15], wdu[7], wde[4];
   {
      const T* wf = entryW + 16 * i64( onDiag ? fd.entry : 0 );
#pragma unroll
      for ( int d = 0; d < NDIR; ++d )
         wd[d] = onDiag ? wf[d] : T( 0 );
#pragma unroll
      for ( int k = 0; k < 7; ++k )
         wdu[k] = wd[kDirU[k]];
#pragma unroll
      for ( int k = 0; k < 4; ++k )
         wde[k] = wd[kDirE[k]];
   }
   T f2v[R], dg[R][P]; // face partial sums gathered over the march
#pragma unroll
   for ( int r = 0; r < R; ++r )
   {
      f2v[r] = T( 0 );
#pragma unroll
      for ( int k = 0; k < P; ++k )
         dg[r][k] = T( 0 );
   }

   // load front: offset of (x0-1, y0-1) in plane pl, whose row 0 has Wl points. Every step issues the
   // same loads (the front stops at the last plane needed, min(z1, N), and repeats it), so the
   // wait before a step's arithmetic only covers loads issued two steps earlier.
   const int plMax = min( z1, N );
   int       pl    = z0 - 1;
   int       Wl    = N + 1 - pl;
   int       off   = S3i( N + 1 ) - S3i( Wl ) + rowStart( Wl, y0 - 1 ) + x0 - 1;
   // Lanes past the end of a row (and rows past the end of a plane) read whatever follows in the
   // array, within the descriptor's range (cell array plus tail slack): no output of the cell
   // interior uses such a lane (every neighbour of an interior point lies in the cell), and the
   // diagonal face's one-sided sums weight them by 0. The loads therefore need no per-lane masks.
   auto loadPlane = [&]( PlaneRows< T, R, P >& Q ) {
      int o = off;
#pragma unroll
      for ( int k = 0; k < R + 2; ++k )
      {
         if constexpr ( P == 1 )
            Q.r[k][0] = bufLoad< T >( rx, vo, o * B );
         else
            bufLoad2( rx, vo, o * B, Q.r[k][0], Q.r[k][1] );
         if constexpr ( kE )
         {
            if ( k >= 1 && k <= R )
            {
               if constexpr ( P == 1 )
                  Q.e[k - 1][0] = bufLoad< T >( re, vo, o * B );
               else
                  bufLoad2( re, vo, o * B, Q.e[k - 1][0], Q.e[k - 1][1] );
            }
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
   const int lastD = N - y0 - x0; // window offset of x = N-1-y-z is lastD - (y - y0) - z
   // point k of the lane stores when it is an output of the window and inside the row:
   // d <= lastD - r - z, i.e. sd[k] + r + z <= lastD (sd = d, or a large value when never)
   int sd[P];
#pragma unroll
   for ( int k = 0; k < P; ++k )
   {
      const int d = d0 + k;
      sd[k]       = ( interior && d >= 1 && d <= WO ) ? d : ( 1 << 28 );
   }
   // A, B, C: planes z-1, z, z+1 (resident); planes z+2 .. z+D are in flight; E receives plane z+D+1
   auto step = [&]( const PlaneRows< T, R, P >& A, const PlaneRows< T, R, P >& Bp, const PlaneRows< T, R, P >& C,
                    PlaneRows< T, R, P >& E ) {
      loadPlane( E );
      int o = oz + ( Wz - ( y0 - 1 ) ); // row y0 of plane z
#pragma unroll
      for ( int r = 0; r < R; ++r )
      {
         // plane z: C = row r+1, N = row r+2, S = row r; plane z+1: TC = row r+1, TS = row r;
         // plane z-1: BC = row r+1, BN = row r+2
         T ax[P];
         stencilRow< T, P >( w, wu, we, Bp.r[r + 1], Bp.r[r + 2], Bp.r[r], C.r[r + 1], C.r[r], A.r[r + 1], A.r[r + 2], ax );
#pragma unroll
         for ( int k = 0; k < P; ++k )
         {
            T o0, o1;
            epilogue< T, MODE >( ax[k], Bp.r[r + 1][k], Bp.e[r][k], Bp.e[r][k], w[15], a, o0, o1 );
            bool mine = true;
            if constexpr ( MODE == MODE_SOR )
               mine = ( sb + s1 * ( x0 - 1 + d0 + k ) + s2 * ( y0 + r ) + s3 * z ) % sorM == sorK;
            const unsigned so = mine && sd[k] <= lastD - r - z ? unsigned( ( d0 + k ) * B ) : kDropOffset;
            bufStore( o0, r0, so, o * B );
            if constexpr ( kTwo )
               bufStore( o1, r1, so, o * B );
         }
         if ( onX0 ) // the first point of lane 0 is the x = 0 face point of row y0+r, plane z
            f2v[r] = moveLane( ax[0], 0, z - z0, f2v[r] );
         const int dEnd = lastD + 1 - r - z; // window offset of the row end
         if ( onDiag && dEnd >= 0 && dEnd < 64 * P ) // (wave-uniform) the row end is in the window
         {
            T ad[P];
            stencilRow< T, P >( wd, wdu, wde, Bp.r[r + 1], Bp.r[r + 2], Bp.r[r], C.r[r + 1], C.r[r], A.r[r + 1], A.r[r + 2],
                                ad );
#pragma unroll
            for ( int k = 0; k < P; ++k )
               dg[r][k] = d0 + k == dEnd ? ad[k] : dg[r][k];
         }
         o += Wz - ( y0 + r );
      }
      oz += T2i( Wz ) - ( y0 - 1 );
      --Wz;
      ++z;
   };
   // register ring of NR = D + 3 planes: z-1, z, z+1 resident, D planes in flight. The loop over the
   // ring slots is unrolled, so every slot is a fixed set of registers (no register copies); every path
   // through a step issues the same memory operations, so the loads of D planes stay in flight.
   constexpr int        NR = D + 3;
   PlaneRows< T, R, P > ring[NR];
#pragma unroll
   for ( int u = 0; u < NR; ++u )
#pragma unroll
      for ( int k = 0; k < R; ++k )
#pragma unroll
         for ( int q = 0; q < P; ++q )
            ring[u].e[k][q] = T( 0 );
   loadPlane( ring[0] ); // plane z0-1
   oz = off;             // plane z0 (z0 < z1 <= plMax, so the front advanced)
#pragma unroll
   for ( int u = 1; u <= D + 1; ++u )
      loadPlane( ring[u] ); // planes z0 .. z0+D; the front is now at z0+D+1
   while ( z < z1 )
   {
#pragma unroll
      for ( int u = 0; u < NR; ++u )
      {
         step( ring[u], ring[( u + 1 ) % NR], ring[( u + 2 ) % NR], ring[( u + NR - 1 ) % NR] );
         if ( z >= z1 )
            break;
      }
   }
   // face partial sums of this segment
   const unsigned pb = unsigned( NP * NP * B );
   if ( onX0 )
   {
      const auto rp = makeRsrc( part + i64( fx.block ) * NP * NP, pb );
#pragma unroll
      for ( int r = 0; r < R; ++r )
      {
         // lane k: point (0, y0+r, z0+k) of the face interior (y + z <= N-1)
         const int  zz = z0 + int( lane );
         const bool ok = zz < z1 && y0 + r + zz <= N - 1;
         bufStore( f2v[r], rp, ok ? unsigned( facePartIndex( 2, N, 0, y0 + r, zz ) * B ) : kDropOffset, 0 );
      }
   }
   if ( onDiag )
   {
      const auto rp = makeRsrc( part + i64( fd.block ) * NP * NP, pb );
#pragma unroll
      for ( int r = 0; r < R; ++r )
#pragma unroll
         for ( int k = 0; k < P; ++k )
         {
            // this point held the row end x = x0-1+d of row y0+r at plane z = N - y - x
            const int  d  = d0 + k;
            const int  zz = N - ( y0 + r ) - ( x0 - 1 + d );
            const bool ok = d >= 1 && d <= WO && zz >= z0 && zz < z1 && zz >= 1;
            bufStore( dg[r][k], rp, ok ? unsigned( facePartIndex( 3, N, 0, y0 + r, zz ) * B ) : kDropOffset, 0 );
         }
   }
}

// ------------------------------------------------------------------ z = 0 and y = 0 faces: one-sided sums
// Work item: cell, slot (0: plane z = 0, row `line`; 1: row y = 0 of plane `line`), window start x0.
template < typename T >
__device__ __forceinline__ void faceFlat( int bid, const int4* __restrict__ items, int nItems, int L, i64 stride,
                                                       const CellFaceSlot* __restrict__ cfs, const T* __restrict__ entryW,
                                                       T* __restrict__ part, int flag, const T* __restrict__ x )
{
   constexpr int B  = int( sizeof( T ) );
   const int     nb = ( nItems + 3 ) / 4;
   const int     wi = remapXcd( bid, nb ) * 4 + __builtin_amdgcn_readfirstlane( int( threadIdx.x >> 6 ) );
   if ( wi >= nItems )
      return;
   const int4         it = items[wi];
   const int          c = it.x, slot = it.y, line = it.z, x0 = it.w;
   const CellFaceSlot fs = cfs[4 * c + slot];
   if ( !( fs.bc & flag ) )
      return;
   const int      N = 1 << L, NP = N + 1, w1 = N + 1;
   const unsigned lane = threadIdx.x & 63u, vo = lane * unsigned( B );
   const auto     rx   = makeRsrc( x + i64( c ) * stride, unsigned( ( stride + kTailSlack ) * B ) );
   const T*       wf   = entryW + 16 * i64( fs.entry );
   T              w[15];
#pragma unroll
   for ( int d = 0; d < NDIR; ++d )
      w[d] = wf[d];
   // row (yy, zz) at x = x0-1+lane, 0 past its end
   auto row = [&]( int yy, int zz ) {
      const int W  = w1 - zz;
      const int le = N - zz - yy - ( x0 - 1 ); // last lane of the row
      if ( yy < 0 || zz < 0 || le < 0 )
         return T( 0 );
      const int o = S3i( w1 ) - S3i( W ) + rowStart( W, yy ) + x0 - 1;
      return bufLoad< T >( rx, int( lane ) <= le ? vo : kDropOffset, o * B );
   };
   // the rows of the stencil around (x, yc, zc); rows outside the cell are 0
   const int yc = slot == 0 ? line : 0, zc = slot == 0 ? 0 : line;
   const T   bc = row( yc, zc ), bn = row( yc + 1, zc ), bs = row( yc - 1, zc ), cc = row( yc, zc + 1 ), cs = row( yc - 1, zc + 1 ),
           ac = row( yc, zc - 1 ), an = row( yc + 1, zc - 1 );
   T Ax = w[0] * bc + w[3] * bn + w[4] * bs + w[7] * cc + w[9] * cs + w[11] * ac + w[12] * an;
   T sr = w[1] * bc + w[5] * bn + w[8] * cc + w[14] * an;
   T sl = w[2] * bc + w[6] * bs + w[10] * cs + w[13] * ac;
   Ax += laneMove< kDppWaveShr1 >( sr );
   Ax += laneMove< kDppWaveShl1 >( sl );
   const int  xx = x0 - 1 + int( lane );
   const bool ok = lane - 1u < 62u && xx + yc + zc <= N - 1;
   const auto rp = makeRsrc( part + i64( fs.block ) * NP * NP, unsigned( NP * NP * B ) );
   bufStore( Ax, rp, ok ? unsigned( facePartIndex( slot, N, xx, yc, zc ) * B ) : kDropOffset, 0 );
}

// ------------------------------------------------------------------ macro-edge / -vertex points
// Each lane gathers the one-sided stencil of every incident cell. Directions leaving a cell carry a
// zero weight in that cell's one-sided stencil; they read the point itself instead, so all loads of a
// lane are unconditional and independent and issue back to back.
template < typename T, int MODE >
__device__ __forceinline__ void lower( int bid, const LowerChunk* __restrict__ chunks, int nChunks,
                                                    const LowerPrim* __restrict__ prims, const LowerEntry* __restrict__ entries,
                                                    const T* __restrict__ entryW, const T* __restrict__ primInv, int L,
                                                    i64 stride, int flag, const T* __restrict__ x, const T* __restrict__ in1,
                                                    T* __restrict__ out0, T* __restrict__ out1, T a, const CellInfo* __restrict__ cinfo,
                                                    int sorK, int sorM )
{
   const int nb = ( nChunks + 3 ) / 4;
   const int ci = remapXcd( bid, nb ) * ( kBlock / 64 ) + __builtin_amdgcn_readfirstlane( int( threadIdx.x >> 6 ) );
   if ( ci >= nChunks )
      return;
   const int        lane = int( threadIdx.x & 63 );
   const LowerChunk ch   = chunks[ci];
   const LowerPrim  p    = prims[ch.prim];
   if ( !( p.bc & flag ) )
      return;
   bool      on = lane < ch.n;
   const int N = 1 << L, w = N + 1;
   const int i = ch.i0 + ( on ? lane : 0 ), j = ch.j;
   if constexpr ( MODE == MODE_SOR )
      on = on && sorColour( cinfo, entries[p.entry0], N, i, j, sorM ) == sorK;
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
   if constexpr ( MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_SOR )
      xc = x[o0i];
   if constexpr ( MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK || MODE == MODE_SOR )
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

// ------------------------------------------------------------------ one launch per sweep
// The cell interiors (with the one-sided sums of the x = 0 and diagonal faces), the one-sided sums of
// the z = 0 and y = 0 faces and the macro-edge/-vertex points read only x and in1 and write disjoint
// points, so they share one launch: block ranges [0, gC), [gC, gC + gF) and the rest. The small
// parts are dispatched last and run in the tail of the cell part.
template < typename T >
struct SweepArgs
{
   const ZSeg*         segs;
   int                 nSegs, nC, L, flag;
   i64                 stride;
   const T*            cellSt;
   const CellFaceSlot* cfs;
   const T*            entryW;
   const T*            primInv;
   T*                  part;
   const T*            x;
   const T*            in1;
   T*                  out0;
   T*                  out1;
   T                   a;
   const int4*         items;
   int                 nItems;
   const LowerChunk*   chunks;
   int                 nChunks;
   const LowerPrim*    prims;
   const LowerEntry*   entries;
   int                 gC, gF;
   const CellInfo*     cinfo = nullptr; // multi-colour SOR: cell colours, the colour of this pass and their number
   int                 sorK = 0, sorM = 1;
};

template < typename T, int MODE, int R, int P, int D >
__device__ __forceinline__ void bSweep( int b, const SweepArgs< T >& p )
{
   if ( b < p.gC )
      cellZ< T, MODE, R, P, D >( b, p.segs, p.nSegs, p.nC, p.L, p.stride, p.cellSt, p.cfs, p.entryW, p.part, p.flag, p.x, p.in1,
                              p.out0, p.out1, p.a, p.cinfo, p.sorK, p.sorM );
   else if ( b < p.gC + p.gF )
      faceFlat< T >( b - p.gC, p.items, p.nItems, p.L, p.stride, p.cfs, p.entryW, p.part, p.flag, p.x );
   else
      lower< T, MODE >( b - p.gC - p.gF, p.chunks, p.nChunks, p.prims, p.entries, p.entryW, p.primInv, p.L, p.stride, p.flag,
                        p.x, p.in1, p.out0, p.out1, p.a, p.cinfo, p.sorK, p.sorM );
}

template < typename T, int MODE, int R, int P, int D >
__global__ __launch_bounds__( kBlock ) void kSweep( SweepArgs< T > p )
{
   bSweep< T, MODE, R, P, D >( int( blockIdx.x ), p );
}

} // namespace hyg
