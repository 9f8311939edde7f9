// Shared by the P2 kernels of p2.hip and the on-the-fly variable-coefficient kernels (p2kq.hpp): class
// sub-lattices of a cell's points, the tile tables that map blocks to them, and the assembly of a point's stencil
// from the coefficient values at the quadrature points of its micro-cells.
#pragma once
#include "face_combine.hpp"

#include <utility>

namespace hyg {

__host__ __device__ constexpr int triCount( int n )
{
   return n > 0 ? n * ( n + 1 ) / 2 : 0;
}

// number of points of one (ku, kv) sub-lattice inside a cell face (face coordinates u = 2 i + 2 - ku, v = 2 j + 2 - kv)
__host__ __device__ constexpr int p2FaceSubCount( int N, int ku, int kv )
{
   return N - 5 + ku + kv < 0 ? 0 : triCount( ( N - 5 + ku + kv ) / 2 + 1 );
}
struct P2FaceTiles
{
   int start[5]; // first tile of each (ku, kv) = (s & 1, s >> 1) sub-lattice within a face; start[4] = per face
};

// parity class of the points of face F whose face coordinates have parities (KU, KV)
__host__ __device__ constexpr int p2FaceClass( int F, int KU, int KV )
{
   return ( F == 2 ? 0 : KU ) | ( ( F == 0 ? KV : ( F == 2 ? KU : ( F == 3 ? ( KU + KV ) & 1 : 0 ) ) ) << 1 ) | ( ( F == 0 ? 0 : KV ) << 2 );
}

// interior points of a lattice with N intervals of parity class cls: p = 2 q + cls, q' = q - (1 - cls), |q'| <= M
__host__ __device__ constexpr int p2InteriorExtent( int N, int cls )
{
   const int k = ( cls & 1 ) + ( ( cls >> 1 ) & 1 ) + ( ( cls >> 2 ) & 1 );
   return ( N - 1 - k ) / 2 - ( 3 - k );
}
// all points of parity class cls of a cell lattice with N intervals: p = 2 q + cls, |q| <= M; M = -1 when the class
// has no point (N = 2, cls = 7: (1, 1, 1) lies outside the cell; integer division would round -1/2 up to 0)
__host__ __device__ constexpr int p2ClassExtent( int N, int cls )
{
   const int r = N - ( cls & 1 ) - ( ( cls >> 1 ) & 1 ) - ( ( cls >> 2 ) & 1 );
   return r < 0 ? -1 : r / 2;
}
struct P2ClassTiles
{
   int start[65]; // first tile of each class within a cell; start[nClasses] = tiles per cell
};

inline P2ClassTiles p2InteriorTiles( int N )
{
   P2ClassTiles ct;
   ct.start[0] = 0;
   for ( int k = 0; k < 8; ++k )
   {
      const int M     = p2InteriorExtent( N, k );
      ct.start[k + 1] = ct.start[k] + ( M >= 0 ? ceilDiv( S3i( M + 1 ), kBlock ) : 0 );
   }
   return ct;
}

// up to three functions handled by one stored-stencil pass (the components of a vector function)
template < typename T >
struct P2Comps
{
   const T* x[3];
   const T* in1[3];
   T*       out0[3];
   T*       out1[3];
};

// Stored stencils (setup) and the on-the-fly apply: the stencil of a point is the sum, over the micro-cells that have
// it as node l (on the cell's side of face F for a face point), of row l of sum_q k_q M_{t,q}, keyed by the offset to
// the other node.
constexpr int p2OffsetIndex( const P2StencilOffsets& o, int dx, int dy, int dz )
{
   for ( int e = 0; e < o.n; ++e )
      if ( o.d[e][0] == dx && o.d[e][1] == dy && o.d[e][2] == dz )
         return e;
   return -1;
}
// offset index of node j seen from node l of micro-cell type t, for points of class CLS (face F)
struct P2OffsetTable
{
   int e[6][10][10] = {};
};
constexpr P2OffsetTable p2OffsetTable( int cls, int f )
{
   P2OffsetTable          tb;
   const P2StencilOffsets off = p2StencilOffsets( cls, f );
   for ( int t = 0; t < 6; ++t )
      for ( int l = 0; l < 10; ++l )
         for ( int j = 0; j < 10; ++j )
            tb.e[t][l][j] = p2OffsetIndex( off, p2NodeOffset( t, j, 0 ) - p2NodeOffset( t, l, 0 ), p2NodeOffset( t, j, 1 ) - p2NodeOffset( t, l, 1 ),
                                           p2NodeOffset( t, j, 2 ) - p2NodeOffset( t, l, 2 ) );
   return tb;
}
// The (type, node) pairs of parity class cls in (t, l) order: the element-matrix rows a point of that class needs.
constexpr int p2PairOrdinal( int cls, int t, int l )
{
   int r = 0;
   for ( int tt = 0; tt < 6; ++tt )
      for ( int ll = 0; ll < 10; ++ll )
         if ( p2NodeClass( tt, ll ) == cls )
         {
            if ( tt == t && ll == l )
               return r;
            ++r;
         }
   return -1;
}
constexpr int kP2MaxPairs = 24; // vertex nodes: 6 types x 4 vertices; edge-midpoint classes have fewer
constexpr int kP2RowPad   = 12; // quadrature-point stride of a staged row (11 points, padded to 16-byte pairs)

// Stages the rows M_{t,q}[l, j] of the class's (t, l) pairs of cell c into LDS, Ms[(r * 10 + j) * kP2RowPad + q]
// (r = p2PairOrdinal), so that the assembly reads them with wave-uniform LDS loads. Every thread of the block calls
// it (it ends with a barrier).
template < typename T, int CLS >
__device__ __forceinline__ void p2StageRows( const T* __restrict__ el, int c, T* __restrict__ Ms )
{
   const T* ec = el + i64( c ) * 6600;
#pragma unroll
   for ( int t = 0; t < 6; ++t )
#pragma unroll
      for ( int l = 0; l < 10; ++l )
      {
         if ( p2NodeClass( t, l ) != CLS )
            continue;
         const int r = p2PairOrdinal( CLS, t, l );
         for ( int o = int( threadIdx.x ); o < 110; o += int( blockDim.x ) )
         {
            const int j = o / 11, q = o - 11 * j;
            Ms[( r * 10 + j ) * kP2RowPad + q] = ec[1100 * t + 100 * q + 10 * l + j];
         }
      }
   __syncthreads();
}

// The micro-cells of a point of class cls (on the cell's side of face f for a face point, f = -1 for interior
// points): type, node, staged-row ordinal, the node's offset from 2 * anchor, and per row j the offset of node j
// from the point ((dz + 2) * 5 + dy + 2 in bits 0..7, dx + 2 in bits 8..10) and its slot in the point's stencil.
struct P2KQEntry
{
   int t = 0, l = 0, ord = 0, mcmax = 0, nd[3] = {}, node[10] = {}, slot[10] = {}; // slot: index of node j in p2StencilOffsets(cls, f)
};
struct P2KQEntries
{
   int       n = 0;
   P2KQEntry e[kP2MaxPairs];
};
struct P2KQEntryTables
{
   P2KQEntries c[8][5]; // [class][face + 1]
};
constexpr P2KQEntryTables p2KQEntryTables()
{
   P2KQEntryTables tb;
   for ( int cls = 0; cls < 8; ++cls )
      for ( int f = -1; f < 4; ++f )
      {
         P2KQEntries&           L  = tb.c[cls][f + 1];
         const P2StencilOffsets so = p2StencilOffsets( cls, f );
         for ( int t = 0; t < 6; ++t )
            for ( int l = 0; l < 10; ++l )
            {
               if ( p2NodeClass( t, l ) != cls || !p2OnCellSide( f, t, l ) )
                  continue;
               P2KQEntry& e = L.e[L.n++];
               e.t          = t;
               e.mcmax      = p2CellMaxSum( t );
               e.l          = l;
               e.ord        = p2PairOrdinal( cls, t, l );
               for ( int k = 0; k < 3; ++k )
                  e.nd[k] = p2NodeOffset( t, l, k );
               for ( int jj = 0; jj < 10; ++jj )
               {
                  const int dx = p2NodeOffset( t, jj, 0 ) - e.nd[0], dy = p2NodeOffset( t, jj, 1 ) - e.nd[1],
                            dz = p2NodeOffset( t, jj, 2 ) - e.nd[2];
                  e.node[jj] = ( ( dz + 2 ) * 5 + dy + 2 ) | ( ( dx + 2 ) << 8 );
                  e.slot[jj] = p2OffsetIndex( so, dx, dy, dz );
               }
            }
      }
   return tb;
}
static __constant__ P2KQEntryTables kP2KQTab = p2KQEntryTables();

// cell-interior and cell-face points of the variable-coefficient operator assembled on the fly (p2kq.hpp; defined for
// double and float in p2kq_f64.hip and p2kq_f32.hip)
template < typename T >
void launchP2KQPoints( int L, int nC, i64 stride, const T* el, const T* kq, i64 kqStride, const T* dinv, const P2Comps< T >& cp, int nc,
                       int mode, T a, const CellFaceSlot* cfs, int flag, T* part, i64 partStride, hipStream_t s );

} // namespace hyg
