// SOR in the reference's order (SORSmoothable::smooth_sor of the constant P1 operator, P1Operator.hpp:343-416).
//
// The reference smooths macro-vertices, -edges, -faces and -cells in turn (backwards: cells first), each primitive
// lexicographically (edges i = 1..N-1; faces j outer, i inner, VertexDoFMacroFace.hpp:751-810; cells z, y, x,
// generatedKernels/sor_3D_macrocell_P1.cpp), reading its neighbours through its own memory: its own values and those
// of its closure are current, the ghost layers hold what the last communication brought. In the replica layout the
// same results follow from one rule: a neighbour q of a point of primitive P is read from the function itself when q
// lies in P's closure or in a primitive of higher dimension, and from a snapshot taken before the sweep otherwise.
// With q's position in a cell, both tests are local: q is in P's closure when its nonzero barycentric coordinates are
// all at P's cell-local vertices, of higher dimension when it has more nonzero coordinates than P has vertices.
//
// Parallel form of the lexicographic order:
//   * vertices: independent (they read each other only from the snapshot), one thread each;
//   * edges: one wave per edge walks i = 1..N-1 (or back), its lanes sharing the terms of each point's sum, so the
//     update of i-1 precedes the read of it;
//   * faces: one block per face; in face coordinates the earlier neighbours of (i, j) are (i-1, j), (i, j-1) and
//     (i+1, j-1), so the points of one anti-diagonal tau = i + 2j are independent and depend only on smaller tau:
//     the block walks tau with a barrier between steps;
//   * cells: the earlier neighbours of (x, y, z) (W, S, SE, BC, BN, BE, BNW) all have a smaller x + 2y + 3z, so all
//     points of one hyperplane tau = x + 2y + 3z of every cell are independent: one launch per tau.
// Update: x_p = (1 - relax) x_p + relax (b_p - sum_{q != p} a_pq x_q) / a_pp.
#include "face_combine.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <stdexcept>
#include <vector>

namespace hyg {

namespace {

// neighbour sum (centre excluded) of point (i, j) of a lower-dimensional primitive with nv vertices, per the rule;
// with MAXC > 0 the primitive has at most MAXC incident cells and the loops are unrolled, so that every load of the
// sum is issued before the first one is waited for
template < typename T, int MAXC = 0 >
__device__ __forceinline__ T ordNeighbourSum( const LowerPrim& p, const LowerEntry* __restrict__ entries, const T* __restrict__ entryW,
                                              int N, i64 stride, int i, int j, int nv, const T* __restrict__ x,
                                              const T* __restrict__ xo, T& centre )
{
   const int w   = N + 1;
   T         sum = T( 0 );
   centre        = T( 0 );
   const int nk  = MAXC > 0 ? MAXC : p.nCells;
#pragma unroll
   for ( int k = 0; k < nk; ++k )
   {
      if ( MAXC > 0 && k >= p.nCells )
         break;
      const LowerEntry e = entries[p.entry0 + k];
      int              qx, qy, qz;
      entryPos( e, N, i, j, qx, qy, qz );
      const T*       wv    = entryW + i64( p.entry0 + k ) * 16;
      const unsigned pmask = ( 1u << e.l0 ) | ( nv > 1 ? 1u << e.l1 : 0u ) | ( nv > 2 ? 1u << e.l2 : 0u );
      const i64      base  = i64( e.cell ) * stride;
      centre += wv[0];
#pragma unroll
      for ( int d = 1; d < NDIR; ++d )
      {
         const T wd = wv[d];
         const int ax = qx + dirX( d ), ay = qy + dirY( d ), az = qz + dirZ( d );
         if ( wd == T( 0 ) || !inCell( N, ax, ay, az ) )
            continue;
         const unsigned nz = ( N - ax - ay - az != 0 ? 1u : 0u ) | ( ax != 0 ? 2u : 0u ) | ( ay != 0 ? 4u : 0u ) | ( az != 0 ? 8u : 0u );
         const bool     cur = __popc( nz ) > nv || ( nz & ~pmask ) == 0u;
         sum += wd * ( cur ? x : xo )[base + cidx( w, ax, ay, az )];
      }
   }
   return sum;
}

// new value of point (i, j) of primitive p, written to every replica
template < typename T, int MAXC = 0 >
__device__ __forceinline__ void ordUpdate( const LowerPrim& p, const LowerEntry* __restrict__ entries, const T* __restrict__ entryW,
                                           int N, i64 stride, int i, int j, int nv, T* __restrict__ x, const T* __restrict__ xo,
                                           const T* __restrict__ b, T relax )
{
   const int w = N + 1;
   T         centre;
   const T   s = ordNeighbourSum< T, MAXC >( p, entries, entryW, N, stride, i, j, nv, x, xo, centre );
   int       qx, qy, qz;
   entryPos( entries[p.entry0], N, i, j, qx, qy, qz );
   const i64 o0 = i64( entries[p.entry0].cell ) * stride + cidx( w, qx, qy, qz );
   const T   v  = ( T( 1 ) - relax ) * x[o0] + relax * ( b[o0] - s ) * ( T( 1 ) / centre );
   for ( int k = 0; k < p.nCells; ++k )
   {
      const LowerEntry e = entries[p.entry0 + k];
      entryPos( e, N, i, j, qx, qy, qz );
      x[i64( e.cell ) * stride + cidx( w, qx, qy, qz )] = v;
   }
}

template < typename T >
__global__ __launch_bounds__( 64 ) void kSorOrdVertices( const int* __restrict__ list, int n, const LowerPrim* __restrict__ prims,
                                                          const LowerEntry* __restrict__ entries, const T* __restrict__ entryW, int L,
                                                          i64 stride, int flag, T* __restrict__ x, const T* __restrict__ xo,
                                                          const T* __restrict__ b, T relax )
{
   const int t = int( blockIdx.x ) * 64 + int( threadIdx.x );
   if ( t >= n )
      return;
   const LowerPrim p = prims[list[t]];
   if ( p.bc & flag )
      ordUpdate( p, entries, entryW, 1 << L, stride, 0, 0, 1, x, xo, b, relax );
}

// one wave per edge; at every point of the edge lane 16 k' + d takes term d (0: the centre weight) of the incident
// cells k = k', k' + 4, ...; the wave's sums give the new value, which the centre lanes write to their cells' replicas
// before the wave moves on (every step reads what the step before wrote)
template < typename T >
__device__ __forceinline__ T waveSum( T v )
{
#pragma unroll
   for ( int m = 32; m >= 1; m >>= 1 )
      v += __shfl_xor( v, m, 64 );
   return v;
}
template < typename T >
__global__ __launch_bounds__( 64 ) void kSorOrdEdges( const int* __restrict__ list, int n, const LowerPrim* __restrict__ prims,
                                                       const LowerEntry* __restrict__ entries, const T* __restrict__ entryW, int L,
                                                       i64 stride, int flag, int backwards, T* __restrict__ x, const T* __restrict__ xo,
                                                       const T* __restrict__ b, T relax )
{
   const int t = int( blockIdx.x );
   if ( t >= n )
      return;
   const LowerPrim p = prims[list[t]];
   if ( !( p.bc & flag ) )
      return;
   const int      N = 1 << L, w = N + 1;
   const int      lane = int( threadIdx.x ), k0 = lane >> 4, d = lane & 15;
   constexpr int  nv   = 2;
   for ( int s = 1; s <= N - 1; ++s )
   {
      const int i   = backwards ? N - s : s;
      T         sum = T( 0 ), cen = T( 0 );
      for ( int k = k0; k < p.nCells; k += 4 )
      {
         const LowerEntry e = entries[p.entry0 + k];
         int              qx, qy, qz;
         entryPos( e, N, i, 0, qx, qy, qz );
         const T* wv = entryW + i64( p.entry0 + k ) * 16;
         if ( d == 0 )
            cen += wv[0];
         else if ( d < NDIR )
         {
            const T   wd = wv[d];
            const int ax = qx + dirX( d ), ay = qy + dirY( d ), az = qz + dirZ( d );
            if ( wd != T( 0 ) && inCell( N, ax, ay, az ) )
            {
               const unsigned pmask = ( 1u << e.l0 ) | ( 1u << e.l1 );
               const unsigned nz    = ( N - ax - ay - az != 0 ? 1u : 0u ) | ( ax != 0 ? 2u : 0u ) | ( ay != 0 ? 4u : 0u ) | ( az != 0 ? 8u : 0u );
               const bool     cur   = __popc( nz ) > nv || ( nz & ~pmask ) == 0u;
               sum += wd * ( cur ? x : xo )[i64( e.cell ) * stride + cidx( w, ax, ay, az )];
            }
         }
      }
      sum = waveSum( sum );
      cen = waveSum( cen );
      int qx, qy, qz;
      entryPos( entries[p.entry0], N, i, 0, qx, qy, qz );
      const i64 o0 = i64( entries[p.entry0].cell ) * stride + cidx( w, qx, qy, qz );
      const T   v  = ( T( 1 ) - relax ) * x[o0] + relax * ( b[o0] - sum ) * ( T( 1 ) / cen );
      if ( d == 0 )
         for ( int k = k0; k < p.nCells; k += 4 )
         {
            const LowerEntry e = entries[p.entry0 + k];
            entryPos( e, N, i, 0, qx, qy, qz );
            x[i64( e.cell ) * stride + cidx( w, qx, qy, qz )] = v;
         }
      // the next point reads these replicas
      __threadfence_block();
   }
}

// one block per face; pts[off[tau] .. off[tau+1]) = the face points (i | j << 16) with i + 2j = tau
template < typename T >
__global__ __launch_bounds__( 256 ) void kSorOrdFaces( const int* __restrict__ list, const LowerPrim* __restrict__ prims,
                                                        const LowerEntry* __restrict__ entries, const T* __restrict__ entryW, int L,
                                                        i64 stride, int flag, int backwards, const int* __restrict__ pts,
                                                        const int* __restrict__ off, int nTau, T* __restrict__ x,
                                                        const T* __restrict__ xo, const T* __restrict__ b, T relax )
{
   const LowerPrim p = prims[list[blockIdx.x]];
   if ( !( p.bc & flag ) )
      return;
   const int N = 1 << L, w = N + 1;
   // the face's (at most two) cells and their stencils, loaded once for the whole sweep of the face
   const int  nc = p.nCells < 2 ? p.nCells : 2;
   LowerEntry E[2];
   T          Wt[2][NDIR];
#pragma unroll
   for ( int k = 0; k < 2; ++k )
   {
      E[k] = entries[p.entry0 + ( k < nc ? k : 0 )];
#pragma unroll
      for ( int d = 0; d < NDIR; ++d )
         Wt[k][d] = entryW[i64( p.entry0 + ( k < nc ? k : 0 ) ) * 16 + d];
   }
   for ( int s = 0; s < nTau; ++s )
   {
      const int tau = backwards ? nTau - 1 - s : s;
      for ( int kk = off[tau] + int( threadIdx.x ); kk < off[tau + 1]; kk += 256 )
      {
         const int ij = pts[kk], i = ij & 0xffff, j = ij >> 16;
         // the same terms in the same order as ordNeighbourSum / ordUpdate
         T sum = T( 0 ), centre = T( 0 );
#pragma unroll
         for ( int k = 0; k < 2; ++k )
         {
            if ( k >= nc )
               break;
            int qx, qy, qz;
            entryPos( E[k], N, i, j, qx, qy, qz );
            const unsigned pmask = ( 1u << E[k].l0 ) | ( 1u << E[k].l1 ) | ( 1u << E[k].l2 );
            const i64      base  = i64( E[k].cell ) * stride;
            centre += Wt[k][0];
#pragma unroll
            for ( int d = 1; d < NDIR; ++d )
            {
               const int ax = qx + dirX( d ), ay = qy + dirY( d ), az = qz + dirZ( d );
               if ( Wt[k][d] == T( 0 ) || !inCell( N, ax, ay, az ) )
                  continue;
               const unsigned nz  = ( N - ax - ay - az != 0 ? 1u : 0u ) | ( ax != 0 ? 2u : 0u ) | ( ay != 0 ? 4u : 0u ) | ( az != 0 ? 8u : 0u );
               const bool     cur = __popc( nz ) > 3 || ( nz & ~pmask ) == 0u;
               sum += Wt[k][d] * ( cur ? x : xo )[base + cidx( w, ax, ay, az )];
            }
         }
         int qx, qy, qz;
         entryPos( E[0], N, i, j, qx, qy, qz );
         const i64 o0 = i64( E[0].cell ) * stride + cidx( w, qx, qy, qz );
         const T   v  = ( T( 1 ) - relax ) * x[o0] + relax * ( b[o0] - sum ) * ( T( 1 ) / centre );
#pragma unroll
         for ( int k = 0; k < 2; ++k )
         {
            if ( k >= nc )
               break;
            entryPos( E[k], N, i, j, qx, qy, qz );
            x[i64( E[k].cell ) * stride + cidx( w, qx, qy, qz )] = v;
         }
      }
      __syncthreads();
   }
}

// the cell-interior points of one hyperplane (cell-array indices pts[0 .. n)) of every cell
template < typename T >
__global__ __launch_bounds__( 256 ) void kSorOrdCells( const int* __restrict__ pts, int n, int tiles, const T* __restrict__ cellSt, int L,
                                                        i64 stride, T* __restrict__ x, const T* __restrict__ b, T relax )
{
   const int c = int( blockIdx.x ) / tiles;
   const int k = ( int( blockIdx.x ) - c * tiles ) * 256 + int( threadIdx.x );
   if ( k >= n )
      return;
   const int w = ( 1 << L ) + 1;
   int       px, py, pz;
   tetXYZ( w, pts[k], px, py, pz );
   const int Wz = w - pz, tc = T2i( Wz ) - py, bc = T2i( Wz + 1 ) - py, nn = Wz - py;
   const int off[NDIR] = { 0, -1, 1, nn, -( nn + 1 ), nn - 1, -nn, tc, tc - 1, tc - nn, tc - nn + 1, -bc, -bc + nn + 1, -bc + 1, -bc + nn };
   const T*  st = cellSt + i64( c ) * 16;
   const i64 q  = i64( c ) * stride + pts[k];
   T         s  = T( 0 );
#pragma unroll
   for ( int d = 1; d < NDIR; ++d )
      s += st[d] * x[q + off[d]];
   x[q] = ( T( 1 ) - relax ) * x[q] + relax * ( b[q] - s ) * ( T( 1 ) / st[0] );
}

template < typename T >
T* devVector( const std::vector< T >& h )
{
   void* d = nullptr;
   if ( hipMalloc( &d, std::max< size_t >( h.size(), 1 ) * sizeof( T ) ) != hipSuccess )
      throw std::runtime_error( "hipMalloc failed for the ordered SOR tables" );
   if ( !h.empty() && hipMemcpy( d, h.data(), h.size() * sizeof( T ), hipMemcpyHostToDevice ) != hipSuccess )
      throw std::runtime_error( "hipMemcpy failed for the ordered SOR tables" );
   return static_cast< T* >( d );
}

} // namespace

static SorOrderTables& sorOrderTables( const MeshCtx& m, hyg_op_s::Level& ol, int L )
{
   if ( ol.sorOrder )
      return *ol.sorOrder;
   auto               t = std::make_unique< SorOrderTables >();
   std::vector< int > lists[3];
   for ( int k = 0; k < int( m.primsHost.size() ); ++k )
      lists[m.primsHost[k].kind].push_back( k );
   t->nVertices = int( lists[0].size() );
   t->nEdges    = int( lists[1].size() );
   t->nFaces    = int( lists[2].size() );
   t->vertices  = devVector( lists[0] );
   t->edges     = devVector( lists[1] );
   t->faces     = devVector( lists[2] );
   const int N = 1 << L, w = N + 1;
   // face points by anti-diagonal tau = i + 2j (tau < 2N)
   std::vector< std::vector< int > > byTau( 2 * N + 1 );
   for ( int j = 1; j <= N - 2; ++j )
      for ( int i = 1; i + j <= N - 1; ++i )
         byTau[i + 2 * j].push_back( i | ( j << 16 ) );
   std::vector< int > fp, fo( 1, 0 );
   for ( auto& v : byTau )
   {
      fp.insert( fp.end(), v.begin(), v.end() );
      fo.push_back( int( fp.size() ) );
   }
   t->nFaceTau = int( byTau.size() );
   t->facePts  = devVector( fp );
   t->faceOff  = devVector( fo );
   // cell-interior points by hyperplane tau = x + 2y + 3z (tau < 3N)
   std::vector< std::vector< int > > byH( 3 * N + 1 );
   for ( int z = 1; z <= N - 3; ++z )
      for ( int y = 1; y + z <= N - 2; ++y )
         for ( int x = 1; x + y + z <= N - 1; ++x )
            byH[x + 2 * y + 3 * z].push_back( cidx( w, x, y, z ) );
   std::vector< int > cp;
   t->cellOff.assign( 1, 0 );
   for ( auto& v : byH )
   {
      cp.insert( cp.end(), v.begin(), v.end() );
      t->cellOff.push_back( int( cp.size() ) );
   }
   t->cellPts = devVector( cp );
   ol.sorOrder = std::move( t );
   return *ol.sorOrder;
}

template < typename T >
static void sorOrderedT( const MeshCtx& m, const LevelGeom& g, hyg_op_s::Level& ol, const SorOrderTables& t, int L, T* x, const T* b,
                         T* xo, T relax, int flag, bool backwards, hipStream_t s )
{
   constexpr bool dbl    = std::is_same< T, double >::value;
   const T*       entryW = dbl ? (const T*) ol.entryW : (const T*) ol.entryWF;
   const T*       cellSt = dbl ? (const T*) ol.cellSt : (const T*) ol.cellStF;
   const i64      n      = i64( m.mesh->nC() ) * g.stride + kTailSlack;
   launchCopy( n, dbl ? 8 : 4, xo, x, s );
   auto vertices = [&]() {
      if ( t.nVertices > 0 )
         hipLaunchKernelGGL( ( kSorOrdVertices< T > ), dim3( ceilDiv( t.nVertices, 64 ) ), dim3( 64 ), 0, s, t.vertices, t.nVertices, m.prims,
                             m.entries, entryW, L, g.stride, flag, x, (const T*) xo, b, relax );
   };
   auto edges = [&]() {
      if ( L >= 1 && t.nEdges > 0 )
         hipLaunchKernelGGL( ( kSorOrdEdges< T > ), dim3( t.nEdges ), dim3( 64 ), 0, s, t.edges, t.nEdges, m.prims, m.entries,
                             entryW, L, g.stride, flag, int( backwards ), x, (const T*) xo, b, relax );
   };
   auto faces = [&]() {
      if ( L >= 2 && t.nFaces > 0 )
         hipLaunchKernelGGL( ( kSorOrdFaces< T > ), dim3( t.nFaces ), dim3( 256 ), 0, s, t.faces, m.prims, m.entries, entryW, L, g.stride,
                             flag, int( backwards ), t.facePts, t.faceOff, t.nFaceTau, x, (const T*) xo, b, relax );
   };
   auto cells = [&]() {
      if ( L < 2 || !( flag & 1 ) )
         return;
      const int nH = int( t.cellOff.size() ) - 1;
      for ( int k = 0; k < nH; ++k )
      {
         const int h  = backwards ? nH - 1 - k : k;
         const int np = t.cellOff[h + 1] - t.cellOff[h];
         if ( np == 0 )
            continue;
         const int tiles = ceilDiv( np, 256 );
         hipLaunchKernelGGL( ( kSorOrdCells< T > ), dim3( m.mesh->nC() * tiles ), dim3( 256 ), 0, s, t.cellPts + t.cellOff[h], np, tiles,
                             cellSt, L, g.stride, x, b, relax );
      }
   };
   if ( backwards )
   {
      cells();
      faces();
      edges();
      vertices();
   }
   else
   {
      vertices();
      edges();
      faces();
      cells();
   }
}

void prepareSorOrdered( const MeshCtx& m, const LevelGeom& g, hyg_op_s::Level& ol, int L, int dtype )
{
   SorOrderTables& t = sorOrderTables( m, ol, L );
   const size_t    n = size_t( m.mesh->nC() ) * size_t( g.stride ) + size_t( kTailSlack );
   void**          p = dtype == 8 ? &t.snap64 : &t.snap32;
   if ( !*p && hipMalloc( p, n * ( dtype == 8 ? sizeof( double ) : sizeof( float ) ) ) != hipSuccess )
      throw std::runtime_error( "hipMalloc failed for the SOR snapshot" );
}

void launchSorOrdered( const MeshCtx& m, const LevelGeom& g, hyg_op_s::Level& ol, int L, int dtype, void* x, const void* b, double relax,
                       int flag, bool backwards, hipStream_t s )
{
   prepareSorOrdered( m, g, ol, L, dtype );
   SorOrderTables& t = *ol.sorOrder;
   if ( dtype == 8 )
      sorOrderedT< double >( m, g, ol, t, L, (double*) x, (const double*) b, (double*) t.snap64, relax, flag, backwards, s );
   else
      sorOrderedT< float >( m, g, ol, t, L, (float*) x, (const float*) b, (float*) t.snap32, float( relax ), flag, backwards, s );
}

} // namespace hyg

SorOrderTables::~SorOrderTables()
{
   (void) hipFree( vertices );
   (void) hipFree( edges );
   (void) hipFree( faces );
   (void) hipFree( facePts );
   (void) hipFree( faceOff );
   (void) hipFree( cellPts );
   (void) hipFree( snap64 );
   (void) hipFree( snap32 );
}
