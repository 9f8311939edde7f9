// Device code of the elementwise variable-coefficient P1 operator (kernel bodies shared by the single-problem and the batched launches).
#pragma once

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
// One thread per anchor (x, y, z), x+y+z <= N-1, of one macro-cell; kb[c][t][anchor] for the six types. The tables
// of the block's cell are wave-uniform: they are read with scalar loads and feed the multiply-adds directly.
template < typename T >
__global__ __launch_bounds__( kBlock ) void kKBar( int L, int tilesPerCell, const double* __restrict__ phase, const T* __restrict__ tab,
                                                   int nModes, double k0, T* __restrict__ kb )
{
   typedef const double __attribute__( ( address_space( 4 ) ) ) CD;
   typedef const T __attribute__( ( address_space( 4 ) ) ) CT;
   const int c    = int( blockIdx.x ) / tilesPerCell;
   const int tile = int( blockIdx.x ) - c * tilesPerCell;
   const int N    = 1 << L;
   const CD* mp   = (const CD*) ( phase + i64( c ) * nModes * 5 );
   const CT* tc   = (const CT*) ( tab + i64( c ) * nModes * 60 );
   const int S    = S3i( N );
   const int s    = tile * kBlock + int( threadIdx.x );
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
      const T   am = T( mp[5 * m + 4] );
      const T   ac = am * cs, as = am * sn;
      const CT* tb = tc + 60 * m;
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

// The same integrals for a polynomial coefficient k(x) = sum_i poly[i] x^a y^b z^c over the monomials of degree <= deg
// in the order of the reference's MonomialBasis3D (by total degree; within a degree by decreasing a, then b). One
// thread per anchor; the polynomial is evaluated at the physical position of each of the 30 quadrature points.
template < typename T >
__global__ __launch_bounds__( kBlock ) void kKBarPoly( int L, int tilesPerCell, const ElemCell* __restrict__ cells,
                                                       const double* __restrict__ poly, int deg, T* __restrict__ kb )
{
   const int c    = int( blockIdx.x ) / tilesPerCell;
   const int tile = int( blockIdx.x ) - c * tilesPerCell;
   const int N    = 1 << L;
   const int S    = S3i( N );
   const int s    = tile * kBlock + int( threadIdx.x );
   if ( s >= S )
      return;
   int x, y, z;
   tetXYZ( N, s, x, y, z );
   const ElemCell& ec  = cells[c];
   T*              out = kb + i64( c ) * 6 * S;
   for ( int t = 0; t < 6; ++t )
   {
      if ( x + y + z + kMCMax[t] > N )
         continue;
      double v = 0.0;
      for ( int q = 0; q < 5; ++q )
      {
         double d[3] = { double( x ), double( y ), double( z ) };
         for ( int u = 0; u < 4; ++u )
         {
            const double lam = q == 0 ? 0.25 : ( u == q - 1 ? 0.5 : 1.0 / 6.0 );
            for ( int k = 0; k < 3; ++k )
               d[k] += lam * kMC[t][u][k];
         }
         double X[3];
         for ( int r = 0; r < 3; ++r )
            X[r] = ec.v0[r] + ec.a[r][0] * d[0] + ec.a[r][1] * d[1] + ec.a[r][2] * d[2];
         double kq = 0.0;
         int    i  = 0;
         for ( int dg = 0; dg <= deg; ++dg )
            for ( int ea = dg; ea >= 0; --ea )
               for ( int eb = dg - ea; eb >= 0; --eb, ++i )
               {
                  double mv = 1.0;
                  for ( int e = 0; e < ea; ++e )
                     mv *= X[0];
                  for ( int e = 0; e < eb; ++e )
                     mv *= X[1];
                  for ( int e = 0; e < dg - ea - eb; ++e )
                     mv *= X[2];
                  kq += poly[i] * mv;
               }
         v += kQW6[q] * kq;
      }
      out[t * S + s] = T( v );
   }
}

// ------------------------------------------------------------------ 2. stencil application per cell
// Stencil weights W[d] (C W E N S NW SE TC TW TS TSE BC BN BE BNW) of lattice point (px, py, pz) of a macro-cell
// from the coefficient integrals kc of its micro-cells (the incident ones that lie in the cell); DIAG: W[0] only.
template < typename T, bool DIAG >
__device__ __forceinline__ void elemWeights( const ElemCell& ec, const T* __restrict__ kc, int N, int px, int py, int pz, T W[NDIR] )
{
   const int SN = S3i( N );
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
}

// One thread per point of the closed macro-cell. DIAG: the diagonal entry instead of the product. BND: the points
// on the cell's faces, edges and vertices only, listed in bnd[0 .. nBnd) by their index in the cell array (the
// surrogate operator does the interior points itself). STW: the point's stencil weights are read from wst
// ([cell][direction][point], the blended operator) instead of assembled from the coefficient integrals.
template < typename T, int MODE, bool DIAG, bool BND = false, bool STW = false >
__device__ __forceinline__ void bElemApply( int bid, int L, int tilesPerCell, i64 stride, const ElemCell* __restrict__ cells,
                                                        const T* __restrict__ kb, const CellInfo* __restrict__ cinfo,
                                                        const CellFaceSlot* __restrict__ cfs, const LowerPrim* __restrict__ prims,
                                                        const LowerEntry* __restrict__ entries, T* __restrict__ part,
                                                        T* __restrict__ lowPart, const T* __restrict__ dinv, int flag,
                                                        const T* __restrict__ x, const T* __restrict__ in1,
                                                        T* __restrict__ out0, T* __restrict__ out1, T a,
                                                        const int* __restrict__ bnd = nullptr, int nBnd = 0,
                                                        const T* __restrict__ wst = nullptr )
{
   const int c    = bid / tilesPerCell;
   const int tile = bid - c * tilesPerCell;
   const int N = 1 << L, w = N + 1, NP = N + 1;
   int       s = tile * kBlock + int( threadIdx.x );
   if constexpr ( BND )
   {
      if ( s >= nBnd )
         return;
      s = bnd[s];
   }
   else if ( s >= S3i( w ) )
      return;
   int px, py, pz;
   tetXYZ( w, s, px, py, pz );
   const int slot = classifySlot( N, px, py, pz );
   if ( slot == 0 && !( flag & 1 ) )
      return;
   T W[NDIR];
   if constexpr ( STW )
   {
      const T* wc = wst + i64( c ) * NDIR * S3i( w ) + s;
#pragma unroll
      for ( int d = 0; d < NDIR; ++d )
         W[d] = ( DIAG && d > 0 ) ? T( 0 ) : wc[i64( d ) * S3i( w )];
   }
   else
      elemWeights< T, DIAG >( cells[c], kb + i64( c ) * 6 * S3i( N ), N, px, py, pz, W );
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

template < typename T, int MODE, bool DIAG, bool BND = false, bool STW = false >
__global__ __launch_bounds__( kBlock ) void kElemApply( int L, int tilesPerCell, i64 stride, const ElemCell* __restrict__ cells,
                                                        const T* __restrict__ kb, const CellInfo* __restrict__ cinfo,
                                                        const CellFaceSlot* __restrict__ cfs, const LowerPrim* __restrict__ prims,
                                                        const LowerEntry* __restrict__ entries, T* __restrict__ part,
                                                        T* __restrict__ lowPart, const T* __restrict__ dinv, int flag,
                                                        const T* __restrict__ x, const T* __restrict__ in1,
                                                        T* __restrict__ out0, T* __restrict__ out1, T a,
                                                        const int* __restrict__ bnd = nullptr, int nBnd = 0,
                                                        const T* __restrict__ wst = nullptr, int nb = 0 )
{
   // nb > 0: the grid is padded to a multiple of 8 and each XCD takes a contiguous range of cells (the hardware deals
   // workgroups out round-robin over the XCDs), so a cell's coefficients and source stay in one XCD's L2
   const int b = nb > 0 ? remapXcd( int( blockIdx.x ), nb ) : int( blockIdx.x );
   if ( nb > 0 && b >= nb )
      return;
   bElemApply< T, MODE, DIAG, BND, STW >( b, L, tilesPerCell, stride, cells, kb, cinfo, cfs, prims, entries, part, lowPart, dinv, flag, x,
                                          in1, out0, out1, a, bnd, nBnd, wst );
}

} // namespace hyg
