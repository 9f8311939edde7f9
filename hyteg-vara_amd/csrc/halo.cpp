// Halo exchange in the reference's layout (host-side tables).
//
// Functions here keep every value of a macro-face, -edge or -vertex as a replica in each incident cell array, so
// no ghost layer is needed on one GPU. A caller that talks to code using the reference's per-primitive memory, or
// a multi-GPU extension that ships halos, needs the reference's layout: every lower-dimensional primitive's own
// values followed by its ghost layers (p1functionspace/VertexDoFMemory.hpp:44-64), filled as the reference's
// communicate chain fills them (p1functionspace/VertexDoFPackInfo.hpp: E->V :204-245, F->E :295-408 with the
// face's cell ghost layers, C->F :508-660). The tables below map every element of that layout to a point of a
// cell array (pack: a gather) and every own value of a lower-dimensional primitive to all its replicas (unpack,
// F->C / E->F / V->E downward copies: a scatter), so the device work is one gather or scatter kernel.
//
// Layout of one level, primitives in storage order (vertices, edges, faces); w = N + 1 points per macro-edge:
//   vertex: own value, then for each incident edge (increasing edge index) the edge point next to the vertex;
//   edge:   own values i = 0..N from its first (lower-id) vertex, then for each incident face (increasing face
//           index) the N points one row into the face (barycentric 1 at the face's third vertex), then for each
//           incident cell (increasing cell index) the N-1 points one step towards both other cell vertices;
//   face:   own values (x, y) of the face's vertex order, x fastest, then for each incident cell (cell 0: the
//           reference's VERTEX_TC ghost layer, cell 1: VERTEX_BC) the N(N+1)/2 points (x, y), x + y <= N - 1, one
//           layer into that cell.
#include "kernels.hpp"

#include <algorithm>
#include <map>
#include <memory>
#include <stdexcept>

namespace hyg {

namespace {

// offset of the point with barycentric lattice coordinates b[] over the cell-local vertices 0..3 in cell c's array
i64 cellPoint( i64 stride, int c, int N, const int b[4] ) { return i64( c ) * stride + cidx( N + 1, b[1], b[2], b[3] ); }

int localOf( const Mesh& m, int c, int v )
{
   for ( int l = 0; l < 4; ++l )
      if ( m.Cv[c][l] == v )
         return l;
   throw std::logic_error( "vertex not in cell" );
}

template < typename T >
T* toDevice( const std::vector< T >& h )
{
   if ( h.empty() )
      return nullptr;
   void* d = nullptr;
   if ( hipMalloc( &d, h.size() * sizeof( T ) ) != hipSuccess )
      throw std::runtime_error( "hipMalloc failed for a halo table" );
   if ( hipMemcpy( d, h.data(), h.size() * sizeof( T ), hipMemcpyHostToDevice ) != hipSuccess )
      throw std::runtime_error( "hipMemcpy failed for a halo table" );
   return static_cast< T* >( d );
}

} // namespace

std::vector< i64 > haloLayout( const Mesh& m, int L )
{
   const i64          N = i64( 1 ) << L, w = N + 1;
   std::vector< i64 > off( 1, 0 );
   std::vector< int > edgesAtVertex( m.nV(), 0 ), facesAtEdge( m.nE(), 0 );
   for ( const auto& e : m.E )
   {
      ++edgesAtVertex[e[0]];
      ++edgesAtVertex[e[1]];
   }
   std::map< std::array< int, 2 >, int > eid;
   for ( int e = 0; e < m.nE(); ++e )
      eid[m.E[e]] = e;
   for ( const auto& f : m.F )
   {
      ++facesAtEdge[eid[{ f[0], f[1] }]];
      ++facesAtEdge[eid[{ f[0], f[2] }]];
      ++facesAtEdge[eid[{ f[1], f[2] }]];
   }
   for ( int v = 0; v < m.nV(); ++v )
      off.push_back( off.back() + 1 + edgesAtVertex[v] );
   for ( int e = 0; e < m.nE(); ++e )
      off.push_back( off.back() + w + facesAtEdge[e] * ( w - 1 ) + i64( m.eC[e].size() ) * ( w - 2 ) );
   for ( int f = 0; f < m.nF(); ++f )
      off.push_back( off.back() + w * ( w + 1 ) / 2 + i64( m.fC[f].size() ) * ( N * ( N + 1 ) / 2 ) );
   return off;
}

HaloMaps::~HaloMaps()
{
   (void) hipFree( gather );
   (void) hipFree( ownRow );
   (void) hipFree( ownSrc );
   (void) hipFree( replica );
   (void) hipFree( ownPrimBc );
   (void) hipFree( sumOwn );
}

const HaloMaps& MeshCtx::halo( int L )
{
   if ( (int) haloMaps.size() <= L )
      haloMaps.resize( L + 1 );
   if ( haloMaps[L] )
      return *haloMaps[L];
   const Mesh& m = *mesh;
   const int   N = 1 << L;
   const i64   stride = cellStride( L );
   auto        hm = std::make_unique< HaloMaps >();
   std::vector< i64 > gather, replica;
   std::vector< int > ownRow( 1, 0 ), ownBc;
   std::vector< i64 > ownSrc;
   const auto         off = haloLayout( m, L );
   gather.reserve( size_t( off.back() ) );
   std::map< std::array< int, 2 >, int > eid;
   for ( int e = 0; e < m.nE(); ++e )
      eid[m.E[e]] = e;
   std::vector< std::vector< int > > edgesAt( m.nV() ), facesAt( m.nE() );
   for ( int e = 0; e < m.nE(); ++e )
   {
      edgesAt[m.E[e][0]].push_back( e );
      edgesAt[m.E[e][1]].push_back( e );
   }
   for ( int f = 0; f < m.nF(); ++f )
   {
      const auto& t = m.F[f];
      for ( auto key : { std::array< int, 2 >{ t[0], t[1] }, std::array< int, 2 >{ t[0], t[2] }, std::array< int, 2 >{ t[1], t[2] } } )
         facesAt[eid[key]].push_back( f );
   }
   // point of a primitive given by global vertex ids pv[0..nv-1] with barycentric weights bw (sum N), in cell c
   auto at = [&]( int c, const int* pv, const int* bw, int nv, const int* extra, int nExtra ) {
      int b[4] = { 0, 0, 0, 0 };
      for ( int k = 0; k < nv; ++k )
         b[localOf( m, c, pv[k] )] += bw[k];
      for ( int k = 0; k < nExtra; ++k )
         b[localOf( m, c, extra[k] )] += 1;
      return cellPoint( stride, c, N, b );
   };
   // own value of a point: gathered from the first incident cell, scattered to all of them; `interior`: the point
   // belongs to this primitive (not to a vertex or edge on its boundary), so the replica sum visits it once
   std::vector< int > sumOwn;
   auto own = [&]( const std::vector< int >& cells, const int* pv, const int* bw, int nv, int bc, bool interior ) {
      gather.push_back( at( cells.front(), pv, bw, nv, nullptr, 0 ) );
      ownSrc.push_back( i64( gather.size() ) - 1 );
      if ( interior )
         sumOwn.push_back( int( ownBc.size() ) );
      for ( int c : cells )
         replica.push_back( at( c, pv, bw, nv, nullptr, 0 ) );
      ownRow.push_back( int( replica.size() ) );
      ownBc.push_back( bc );
   };
   for ( int v = 0; v < m.nV(); ++v )
   {
      const int bw1[1] = { N };
      own( m.vC[v], &v, bw1, 1, bcOfFlag( m.vB[v] ), true );
      for ( int e : edgesAt[v] )
      {
         // the edge point next to v: one step from v towards the edge's other vertex
         const int pv[2] = { v, m.E[e][0] == v ? m.E[e][1] : m.E[e][0] };
         const int bw[2] = { N - 1, 1 };
         gather.push_back( at( m.eC[e].front(), pv, bw, 2, nullptr, 0 ) );
      }
   }
   for ( int e = 0; e < m.nE(); ++e )
   {
      const int* ev = m.E[e].data();
      for ( int i = 0; i <= N; ++i )
      {
         const int bw[2] = { N - i, i };
         own( m.eC[e], ev, bw, 2, bcOfFlag( m.eB[e] ), i > 0 && i < N );
      }
      for ( int f : facesAt[e] )
      {
         const auto& t  = m.F[f];
         const int   op = t[0] != ev[0] && t[0] != ev[1] ? t[0] : ( t[1] != ev[0] && t[1] != ev[1] ? t[1] : t[2] );
         for ( int i = 0; i < N; ++i )
         {
            const int pv[3] = { ev[0], ev[1], op };
            const int bw[3] = { N - 1 - i, i, 1 };
            gather.push_back( at( m.fC[f].front(), pv, bw, 3, nullptr, 0 ) );
         }
      }
      for ( int c : m.eC[e] )
      {
         int others[2], k = 0;
         for ( int l = 0; l < 4; ++l )
            if ( m.Cv[c][l] != ev[0] && m.Cv[c][l] != ev[1] )
               others[k++] = m.Cv[c][l];
         for ( int i = 0; i + 1 < N; ++i )
         {
            const int bw[2] = { N - 2 - i, i };
            gather.push_back( at( c, ev, bw, 2, others, 2 ) );
         }
      }
   }
   for ( int f = 0; f < m.nF(); ++f )
   {
      const int* fv = m.F[f].data();
      for ( int y = 0; y <= N; ++y )
         for ( int x = 0; x + y <= N; ++x )
         {
            const int bw[3] = { N - x - y, x, y };
            own( m.fC[f], fv, bw, 3, bcOfFlag( m.fB[f] ), x > 0 && y > 0 && x + y < N );
         }
      for ( int c : m.fC[f] )
      {
         int fourth = -1;
         for ( int l = 0; l < 4; ++l )
            if ( m.Cv[c][l] != fv[0] && m.Cv[c][l] != fv[1] && m.Cv[c][l] != fv[2] )
               fourth = m.Cv[c][l];
         for ( int y = 0; y < N; ++y )
            for ( int x = 0; x + y < N; ++x )
            {
               const int bw[3] = { N - 1 - x - y, x, y };
               gather.push_back( at( c, fv, bw, 3, &fourth, 1 ) );
            }
      }
   }
   if ( i64( gather.size() ) != off.back() )
      throw std::logic_error( "halo layout size mismatch" );
   hm->size      = off.back();
   hm->nOwn      = int( ownSrc.size() );
   hm->gather    = toDevice( gather );
   hm->ownRow    = toDevice( ownRow );
   hm->ownSrc    = toDevice( ownSrc );
   hm->replica   = toDevice( replica );
   hm->ownPrimBc = toDevice( ownBc );
   hm->nSum      = int( sumOwn.size() );
   hm->sumOwn    = toDevice( sumOwn );
   haloMaps[L]   = std::move( hm );
   return *haloMaps[L];
}

} // namespace hyg
