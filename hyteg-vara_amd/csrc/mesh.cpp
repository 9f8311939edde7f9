// Macro-mesh input and topology (host side).
//
// Reads Gmsh v2.2 tetrahedral meshes and builds the macro-primitives in the reference's
// storage order: vertices by Gmsh node id, edges and faces by their sorted vertex ids,
// cells by their vertex list in Gmsh element order, cell-local vertex order = element
// order (src/hyteg/mesh/MeshGenFromGmshFile.cpp:32-200;
// src/hyteg/primitivestorage/SetupPrimitiveStorage.cpp:40-365). The cell-local vertex order
// fixes the micro-tetrahedra of the uniform refinement, so it must follow the file exactly.
// Boundary primitives follow setMeshBoundaryFlagsOnBoundary(1, 0, true)
// (SetupPrimitiveStorage.cpp:781-792, 895-945): faces with one incident cell, and every
// edge and vertex of such a face; cells are always inner. With keepTags the mesh boundary flags
// are the file's instead (MeshGenFromGmshFile.cpp:100-116, 160-200): a line or triangle element
// gives its edge or face the element's first (physical) tag, every other primitive gets 0.
#include "hyg_internal.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <sstream>

namespace hyg {

static const int kCellFaceVerts[4][3] = { { 0, 1, 2 }, { 0, 1, 3 }, { 0, 2, 3 }, { 1, 2, 3 } };
static const int kCellEdgeVerts[6][2] = { { 0, 1 }, { 0, 2 }, { 1, 2 }, { 0, 3 }, { 1, 3 }, { 2, 3 } };

std::unique_ptr< Mesh > buildMesh( const std::map< long, std::array< double, 3 > >& nodes,
                                   const std::map< std::vector< long >, int >& cellKeys, const std::string& what, std::string& err );

// mesh boundary flags of the file's tagged edges and faces (node id lists as written in the file)
struct GmshTags
{
   std::map< std::vector< long >, int > edges, faces;
};

std::unique_ptr< Mesh > readGmsh( const std::string& path, std::string& err, bool keepTags )
{
   std::ifstream in( path );
   if ( !in )
   {
      err = "cannot open mesh file " + path;
      return nullptr;
   }
   std::string tok;
   in >> tok;
   if ( tok != "$MeshFormat" )
   {
      err = "missing $MeshFormat in " + path;
      return nullptr;
   }
   std::string version;
   in >> version >> tok >> tok >> tok;
   if ( version != "2.2" )
   {
      err = "mesh file version must be 2.2: " + path;
      return nullptr;
   }
   in >> tok;
   if ( tok != "$Nodes" )
   {
      err = "missing $Nodes in " + path;
      return nullptr;
   }
   long nn = 0;
   in >> nn;
   std::map< long, std::array< double, 3 > > nodes;
   for ( long k = 0; k < nn; ++k )
   {
      long                    id;
      std::array< double, 3 > x;
      in >> id >> x[0] >> x[1] >> x[2];
      nodes[id] = x;
   }
   in >> tok >> tok;
   if ( tok != "$Elements" || !in )
   {
      err = "malformed $Nodes/$Elements section in " + path;
      return nullptr;
   }
   long ne = 0;
   in >> ne;
   std::map< std::vector< long >, int > cellKeys; // ordered like the reference's CellContainer
   GmshTags                             tags;
   for ( long k = 0; k < ne; ++k )
   {
      long id, type, ntags, first = 0;
      in >> id >> type >> ntags;
      for ( long t = 0; t < ntags; ++t )
      {
         long tag;
         in >> tag;
         if ( t == 0 )
            first = tag;
      }
      int nv = type == 15 ? 1 : type == 1 ? 2 : type == 2 ? 3 : type == 4 ? 4 : -1;
      if ( nv < 0 || !in )
      {
         err = "unsupported element type in " + path;
         return nullptr;
      }
      std::vector< long > vs( nv );
      for ( auto& v : vs )
         in >> v;
      if ( type == 4 )
      {
         for ( long v : vs )
            if ( !nodes.count( v ) )
            {
               err = "element references unknown node in " + path;
               return nullptr;
            }
         cellKeys[vs] = 0;
      }
      else if ( type == 1 )
         tags.edges[vs] = int( first );
      else if ( type == 2 )
         tags.faces[vs] = int( first );
   }
   auto m = buildMesh( nodes, cellKeys, path, err );
   if ( m && keepTags )
   {
      // the file's flags replace the geometric rule; like MeshInfo::addEdge / addFace the first entry of a
      // sorted vertex set (in the order of the parsed containers) keeps its flag
      std::map< long, int > vid; // vertex index = rank of the node id, as in buildMesh
      for ( auto& kv : nodes )
      {
         const int idx = int( vid.size() );
         vid[kv.first] = idx;
      }
      std::map< std::array< int, 2 >, int > eIdx;
      std::map< std::array< int, 3 >, int > fIdx;
      for ( int e = 0; e < m->nE(); ++e )
         eIdx[m->E[e]] = e;
      for ( int f = 0; f < m->nF(); ++f )
         fIdx[m->F[f]] = f;
      m->vB.assign( m->nV(), 0 );
      m->eB.assign( m->nE(), 0 );
      m->fB.assign( m->nF(), 0 );
      std::vector< bool > eSet( m->nE(), false ), fSet( m->nF(), false );
      for ( auto& kv : tags.edges )
      {
         if ( !vid.count( kv.first[0] ) || !vid.count( kv.first[1] ) )
            continue;
         std::array< int, 2 > key{ vid[kv.first[0]], vid[kv.first[1]] };
         std::sort( key.begin(), key.end() );
         auto it = eIdx.find( key );
         if ( it != eIdx.end() && !eSet[it->second] )
         {
            m->eB[it->second] = uint8_t( kv.second );
            eSet[it->second]  = true;
         }
      }
      for ( auto& kv : tags.faces )
      {
         if ( !vid.count( kv.first[0] ) || !vid.count( kv.first[1] ) || !vid.count( kv.first[2] ) )
            continue;
         std::array< int, 3 > key{ vid[kv.first[0]], vid[kv.first[1]], vid[kv.first[2]] };
         std::sort( key.begin(), key.end() );
         auto it = fIdx.find( key );
         if ( it != fIdx.end() && !fSet[it->second] )
         {
            m->fB[it->second] = uint8_t( kv.second );
            fSet[it->second]  = true;
         }
      }
   }
   return m;
}

// Primitives of a tetrahedral mesh given by vertex IDs -> coordinates and cells (vertex ID lists in the
// cell's local order), numbered as SetupPrimitiveStorage numbers the containers of a MeshInfo
// (vertices by ID, cells by their vertex list, edges and faces by their sorted vertex lists); boundary
// flags by setMeshBoundaryFlagsOnBoundary (faces with one cell, and their edges and vertices).
std::unique_ptr< Mesh > buildMesh( const std::map< long, std::array< double, 3 > >& nodes,
                                   const std::map< std::vector< long >, int >& cellKeys, const std::string& what, std::string& err )
{
   const std::string path = what;
   if ( cellKeys.empty() )
   {
      err = "mesh has no tetrahedra: " + path;
      return nullptr;
   }
   auto                m = std::make_unique< Mesh >();
   std::map< long, int > vid;
   for ( auto& kv : nodes )
   {
      vid[kv.first] = m->nV();
      m->X.push_back( kv.second );
   }
   std::map< std::array< int, 2 >, int > eid;
   std::map< std::array< int, 3 >, int > fid;
   for ( auto& kv : cellKeys )
   {
      std::array< int, 4 > cv{ vid[kv.first[0]], vid[kv.first[1]], vid[kv.first[2]], vid[kv.first[3]] };
      m->Cv.push_back( cv );
      for ( auto& e : kCellEdgeVerts )
      {
         std::array< int, 2 > key{ cv[e[0]], cv[e[1]] };
         std::sort( key.begin(), key.end() );
         eid[key] = 0;
      }
      for ( auto& f : kCellFaceVerts )
      {
         std::array< int, 3 > key{ cv[f[0]], cv[f[1]], cv[f[2]] };
         std::sort( key.begin(), key.end() );
         fid[key] = 0;
      }
   }
   for ( auto& kv : eid )
   {
      kv.second = m->nE();
      m->E.push_back( kv.first );
   }
   for ( auto& kv : fid )
   {
      kv.second = m->nF();
      m->F.push_back( kv.first );
   }
   m->vC.resize( m->nV() );
   m->eC.resize( m->nE() );
   m->fC.resize( m->nF() );
   for ( int c = 0; c < m->nC(); ++c )
   {
      const auto&          cv = m->Cv[c];
      std::array< int, 4 > cf;
      std::array< int, 6 > ce;
      for ( int k = 0; k < 4; ++k )
      {
         std::array< int, 3 > key{ cv[kCellFaceVerts[k][0]], cv[kCellFaceVerts[k][1]], cv[kCellFaceVerts[k][2]] };
         std::sort( key.begin(), key.end() );
         cf[k] = fid[key];
         m->fC[cf[k]].push_back( c );
      }
      for ( int k = 0; k < 6; ++k )
      {
         std::array< int, 2 > key{ cv[kCellEdgeVerts[k][0]], cv[kCellEdgeVerts[k][1]] };
         std::sort( key.begin(), key.end() );
         ce[k] = eid[key];
         m->eC[ce[k]].push_back( c );
      }
      for ( int v : cv )
         m->vC[v].push_back( c );
      m->cF.push_back( cf );
      m->cE.push_back( ce );
   }
   m->vB.assign( m->nV(), 0 );
   m->eB.assign( m->nE(), 0 );
   m->fB.assign( m->nF(), 0 );
   for ( int f = 0; f < m->nF(); ++f )
   {
      if ( m->fC[f].size() > 2 )
      {
         err = "face with more than two cells in " + path;
         return nullptr;
      }
      if ( m->fC[f].size() != 1 )
         continue;
      m->fB[f] = 1;
      const auto& t = m->F[f];
      m->eB[eid[{ t[0], t[1] }]] = 1;
      m->eB[eid[{ t[0], t[2] }]] = 1;
      m->eB[eid[{ t[1], t[2] }]] = 1;
      for ( int v : t )
         m->vB[v] = 1;
   }
   return m;
}

// MeshInfo::meshSphericalShell( ntan, nrad, rmin, rmax ) with the default SHELLMESH_CLASSIC meshing
// (mesh/MeshGenSphericalShell.cpp): the icosahedron's ten diamonds on the unit sphere, refined by arc
// midpoints (ntan - 1 a power of two), scaled to nrad equidistant radii; each diamond cell between two
// layers is split into six tetrahedra (northern and southern hemispheres use mirrored splittings).
std::unique_ptr< Mesh > meshSphericalShell( int ntan, int nrad, double rmin, double rmax, std::string& err )
{
   const int nt = ntan - 1;
   if ( ntan < 2 || ( nt & ( nt - 1 ) ) != 0 || nrad < 2 || !( rmin > 0.0 ) || !( rmin < rmax ) )
   {
      err = "spherical shell needs ntan - 1 a power of two, nrad >= 2 and 0 < rmin < rmax";
      return nullptr;
   }
   using P = std::array< double, 3 >;
   // the twelve icosahedral nodes: poles, upper ring, lower ring (anti-clockwise)
   const double fifthpi = 0.4 * std::asin( 1.0 );
   const double w       = 2.0 * std::acos( 1.0 / ( 2.0 * std::sin( fifthpi ) ) );
   const double cosw = std::cos( w ), sinw = std::sin( w );
   P            ico[12];
   ico[0]  = { 0.0, 0.0, 1.0 };
   ico[11] = { 0.0, 0.0, -1.0 };
   for ( int k = 1; k <= 5; ++k )
   {
      const double pu = 2.0 * ( double( k ) - 0.5 ) * fifthpi, pl = 2.0 * ( double( k ) - 1.0 ) * fifthpi;
      ico[k]     = { sinw * std::cos( pu ), sinw * std::sin( pu ), cosw };
      ico[k + 5] = { sinw * std::cos( pl ), sinw * std::sin( pl ), -cosw };
   }
   // diamonds: pole, then the other three nodes anti-clockwise
   static const int dia[10][4] = { { 0, 5, 6, 1 },   { 0, 1, 7, 2 },  { 0, 2, 8, 3 },  { 0, 3, 9, 4 },  { 0, 4, 10, 5 },
                                   { 11, 7, 1, 6 }, { 11, 8, 2, 7 }, { 11, 9, 3, 8 }, { 11, 10, 4, 9 }, { 11, 6, 5, 10 } };
   // unit-sphere nodes of every diamond, [is1][is2][id], by repeated arc-midpoint refinement
   std::vector< P > uc( size_t( ntan ) * ntan * 10 );
   auto             U = [&]( int a, int b, int d ) -> P& { return uc[( size_t( a ) * ntan + b ) * 10 + d]; };
   auto             mid = []( const P& a, const P& b ) {
      const double x = a[0] + b[0], y = a[1] + b[1], z = a[2] + b[2], s = std::sqrt( x * x + y * y + z * z );
      return P{ x / s, y / s, z / s };
   };
   for ( int d = 0; d < 10; ++d )
   {
      const int l = d < 5 ? 1 : 3, r = d < 5 ? 3 : 1;
      U( 0, 0, d )     = ico[dia[d][0]];
      U( nt, 0, d )    = ico[dia[d][l]];
      U( nt, nt, d )   = ico[dia[d][2]];
      U( 0, nt, d )    = ico[dia[d][r]];
      for ( int m = 1; m < nt; m *= 2 )
      {
         const int len = nt / m, h = len / 2;
         for ( int j1 = 0; j1 <= m; ++j1 ) // rows
            for ( int j2 = 0; j2 < m; ++j2 )
               U( j1 * len, j2 * len + h, d ) = mid( U( j1 * len, j2 * len, d ), U( j1 * len, j2 * len + len, d ) );
         for ( int j1 = 0; j1 <= m; ++j1 ) // columns
            for ( int j2 = 0; j2 < m; ++j2 )
               U( j2 * len + h, j1 * len, d ) = mid( U( j2 * len, j1 * len, d ), U( j2 * len + len, j1 * len, d ) );
         for ( int j1 = 0; j1 < m; ++j1 ) // diagonals
            for ( int j2 = 0; j2 < m; ++j2 )
               U( j1 * len + h, j2 * len + h, d ) = mid( U( j1 * len, j2 * len + len, d ), U( j1 * len + len, j2 * len, d ) );
      }
   }
   std::vector< double > radius( nrad );
   for ( int k = 0; k < nrad; ++k )
      radius[k] = rmin + ( ( rmax - rmin ) / double( nrad - 1 ) ) * double( k );
   // vertex IDs: north poles (one per layer), south poles, then per layer, diamond, row and column
   auto vertexId = [&]( int s1, int s2, int d, int r ) -> long {
      if ( s1 == 0 && s2 == 0 )
         return d < 5 ? r : nrad + r;
      if ( s2 == 0 ) // upper left edge belongs to the neighbouring diamond
      {
         s2 = s1;
         s1 = 0;
         d  = d < 5 ? ( d + 4 ) % 5 : ( d + 4 ) % 5 + 5;
      }
      else if ( s1 == nt ) // lower left edge
      {
         s1 = nt - s2;
         s2 = nt;
         d  = d < 5 ? ( d + 4 ) % 5 + 5 : d % 5;
      }
      return 2L * nrad + long( r ) * nt * nt * 10 + long( d ) * nt * nt + long( s2 - 1 ) * nt + s1;
   };
   std::map< long, P > nodes;
   const long          nVerts = ( 10L * nt * nt + 2 ) * nrad;
   for ( long v = 0; v < nVerts; ++v )
   {
      int s1 = 0, s2 = 0, d = 0, r = 0;
      if ( v >= 2L * nrad )
      {
         long i = v - 2L * nrad;
         r      = int( i / ( 10L * nt * nt ) );
         i -= long( r ) * 10 * nt * nt;
         d = int( i / ( long( nt ) * nt ) );
         i -= long( d ) * nt * nt;
         s2 = int( i / nt ) + 1;
         s1 = int( i - long( s2 - 1 ) * nt );
      }
      else if ( v < nrad )
         r = int( v );
      else
      {
         d = 5;
         r = int( v - nrad );
      }
      const P& u = U( s1, s2, d );
      nodes[v]   = { u[0] * radius[r], u[1] * radius[r], u[2] * radius[r] };
   }
   // the six tetrahedra of a diamond cell (local vertices 0-3 on the outer layer, 4-7 on the inner)
   static const int tN[6][4]  = { { 0, 1, 3, 7 }, { 4, 7, 5, 0 }, { 0, 5, 1, 7 }, { 1, 2, 3, 6 }, { 5, 7, 6, 1 }, { 3, 6, 7, 1 } };
   static const int tS[6][4]  = { { 4, 7, 5, 1 }, { 0, 1, 3, 4 }, { 4, 3, 7, 1 }, { 7, 6, 5, 2 }, { 3, 1, 2, 7 }, { 5, 2, 1, 7 } };
   static const int off[8][3] = { { 0, 0, 1 }, { 1, 0, 1 }, { 1, 1, 1 }, { 0, 1, 1 }, { 0, 0, 0 }, { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 } };
   std::map< std::vector< long >, int > cellKeys;
   const long                           nElems = long( nt ) * nt * ( nrad - 1 ) * 10 * 6;
   for ( long e = 0; e < nElems; ++e )
   {
      long      i  = e;
      const int r  = int( i / ( 60L * nt * nt ) );
      i -= long( r ) * 60 * nt * nt;
      const int d = int( i / ( 6L * nt * nt ) );
      i -= long( d ) * 6 * nt * nt;
      const int s2 = int( i / ( 6L * nt ) );
      i -= long( s2 ) * 6 * nt;
      const int s1 = int( i / 6 );
      const int t  = int( i - 6L * s1 );
      long      vi[8];
      for ( int k = 0; k < 8; ++k )
         vi[k] = d <= 4 ? vertexId( s1 + off[k][0], s2 + off[k][1], d, r + off[k][2] )
                        : vertexId( s1 + off[k][1], s2 + off[k][0], d, r + off[k][2] );
      const int( &tl )[4] = d <= 4 ? tN[t] : tS[t];
      cellKeys[{ vi[tl[0]], vi[tl[1]], vi[tl[2]], vi[tl[3]] }] = 0;
   }
   return buildMesh( nodes, cellKeys, "spherical shell", err );
}

// MeshInfo::meshCuboid( lowerLeftFront, upperRightBack, nx, ny, nz ) (mesh/MeshGenCuboid.cpp:35-174): the
// (nx+1)(ny+1)(nz+1) grid points, numbered x fastest, then y, then z, each cube of the grid split into the same six
// tetrahedra (the reference's tNode table over the cube corners 0..7 = (0,0,0) (1,0,0) (1,1,0) (0,1,0) and the
// same four one layer up), cell vertices in that table's order.
std::unique_ptr< Mesh > meshCuboid( const double lo[3], const double hi[3], int nx, int ny, int nz, std::string& err )
{
   if ( nx < 1 || ny < 1 || nz < 1 )
   {
      err = "cuboid needs at least one cube in every direction";
      return nullptr;
   }
   static const int corner[8][3] = { { 0, 0, 0 }, { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 },
                                     { 0, 0, 1 }, { 1, 0, 1 }, { 1, 1, 1 }, { 0, 1, 1 } };
   static const int split[6][4]  = { { 4, 5, 7, 3 }, { 0, 3, 1, 4 }, { 4, 1, 5, 3 }, { 5, 6, 7, 2 }, { 1, 3, 2, 5 }, { 7, 2, 3, 5 } };
   const double     h[3] = { ( hi[0] - lo[0] ) / double( nx ), ( hi[1] - lo[1] ) / double( ny ), ( hi[2] - lo[2] ) / double( nz ) };
   auto             gid  = [&]( int i, int j, int k ) { return ( long( k ) * ( ny + 1 ) + j ) * ( nx + 1 ) + i; };
   std::map< long, std::array< double, 3 > > nodes;
   for ( int k = 0; k <= nz; ++k )
      for ( int j = 0; j <= ny; ++j )
         for ( int i = 0; i <= nx; ++i )
            nodes[gid( i, j, k )] = { lo[0] + double( i ) * h[0], lo[1] + double( j ) * h[1], lo[2] + double( k ) * h[2] };
   std::map< std::vector< long >, int > cellKeys;
   for ( int k = 0; k < nz; ++k )
      for ( int j = 0; j < ny; ++j )
         for ( int i = 0; i < nx; ++i )
            for ( const auto& t : split )
            {
               std::vector< long > key;
               for ( int c : t )
                  key.push_back( gid( i + corner[c][0], j + corner[c][1], k + corner[c][2] ) );
               cellKeys[key] = 0;
            }
   return buildMesh( nodes, cellKeys, "cuboid", err );
}

// MeshInfo::meshCubedDomain( cubeCoordinates, cubeType ) (mesh/MeshGenCubedDomain.cpp:37-383): unit cubes at the
// given integer positions, each split into 6 tetrahedra (type 0, the cuboid's corner table) or 24 around its face
// centres (type 1, the symmetric cuboid's table). Vertex IDs follow the reference's numbering over the bounding
// box of the cubes: the corner grid x fastest, then (type 1) the centres of the xy-, xz- and yz-faces, each family
// over its own grid.
std::unique_ptr< Mesh > meshCubedDomain( const int* cubes, int nCubes, int cubeType, std::string& err )
{
   if ( nCubes < 1 || ( cubeType != 0 && cubeType != 1 ) )
   {
      err = "cubed domain needs at least one cube and cube type 0 or 1";
      return nullptr;
   }
   int lo[3] = { cubes[0], cubes[1], cubes[2] }, hi[3] = { cubes[0], cubes[1], cubes[2] };
   for ( int q = 0; q < nCubes; ++q )
      for ( int c = 0; c < 3; ++c )
      {
         lo[c] = std::min( lo[c], cubes[3 * q + c] );
         hi[c] = std::max( hi[c], cubes[3 * q + c] );
      }
   const long NX = hi[0] + 1 - lo[0], NY = hi[1] + 1 - lo[1], NZ = hi[2] + 1 - lo[2]; // cubes per direction
   const long nCorner = ( NX + 1 ) * ( NY + 1 ) * ( NZ + 1 ), nXY = NX * NY * ( NZ + 1 ), nXZ = NX * NZ * ( NY + 1 );
   // ID of a point given in half-cube units (h = 2 x for a grid coordinate x)
   auto idOfHalf = [&]( int hx, int hy, int hz ) -> long {
      const long x = hx >> 1, y = hy >> 1, z = hz >> 1; // lower grid coordinate
      const long i = x - lo[0], j = y - lo[1], k = z - lo[2];
      const int  odd = ( hx & 1 ) | ( ( hy & 1 ) << 1 ) | ( ( hz & 1 ) << 2 );
      switch ( odd )
      {
      case 0: return ( k * ( NY + 1 ) + j ) * ( NX + 1 ) + i;                       // corner
      case 3: return nCorner + ( k * NY + j ) * NX + i;                             // xy-face centre
      case 5: return nCorner + nXY + ( k * ( NY + 1 ) + j ) * NX + i;                // xz-face centre
      default: return nCorner + nXY + nXZ + ( k * NY + j ) * ( NX + 1 ) + i;        // yz-face centre (odd == 6)
      }
   };
   static const int corner[8][3] = { { 0, 0, 0 }, { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 },
                                     { 0, 0, 1 }, { 1, 0, 1 }, { 1, 1, 1 }, { 0, 1, 1 } };
   static const int split6[6][4] = { { 4, 5, 7, 3 }, { 0, 3, 1, 4 }, { 4, 1, 5, 3 }, { 5, 6, 7, 2 }, { 1, 3, 2, 5 }, { 7, 2, 3, 5 } };
   static const int lp[14][3]    = { { 0, 0, 0 }, { 2, 0, 0 }, { 1, 1, 0 }, { 0, 2, 0 }, { 2, 2, 0 }, { 1, 0, 1 }, { 0, 1, 1 },
                                     { 2, 1, 1 }, { 1, 2, 1 }, { 0, 0, 2 }, { 2, 0, 2 }, { 1, 1, 2 }, { 0, 2, 2 }, { 2, 2, 2 } };
   static const int split24[24][4] = { { 0, 1, 2, 5 },    { 0, 3, 2, 6 },    { 1, 2, 4, 7 },    { 2, 4, 3, 8 },   { 2, 0, 5, 6 },
                                       { 2, 1, 5, 7 },    { 2, 3, 6, 8 },    { 2, 4, 7, 8 },    { 0, 9, 5, 6 },   { 1, 10, 5, 7 },
                                       { 3, 12, 6, 8 },   { 4, 13, 7, 8 },   { 2, 5, 6, 8 },    { 2, 5, 7, 8 },   { 11, 5, 6, 8 },
                                       { 11, 5, 7, 8 },   { 11, 9, 5, 6 },   { 11, 10, 5, 7 },  { 11, 12, 6, 8 }, { 11, 13, 7, 8 },
                                       { 11, 9, 10, 5 },  { 11, 9, 12, 6 },  { 11, 10, 13, 7 }, { 11, 12, 13, 8 } };
   std::map< long, std::array< double, 3 > > nodes;
   std::map< std::vector< long >, int >      cellKeys;
   for ( int q = 0; q < nCubes; ++q )
   {
      const int x = cubes[3 * q], y = cubes[3 * q + 1], z = cubes[3 * q + 2];
      auto      vert = [&]( int hx, int hy, int hz ) {
         const long id = idOfHalf( 2 * x + hx, 2 * y + hy, 2 * z + hz );
         nodes[id]     = { double( x ) + 0.5 * hx, double( y ) + 0.5 * hy, double( z ) + 0.5 * hz };
         return id;
      };
      if ( cubeType == 0 )
         for ( const auto& t : split6 )
         {
            std::vector< long > key;
            for ( int c : t )
               key.push_back( vert( 2 * corner[c][0], 2 * corner[c][1], 2 * corner[c][2] ) );
            cellKeys[key] = 0;
         }
      else
         for ( const auto& t : split24 )
         {
            std::vector< long > key;
            for ( int c : t )
               key.push_back( vert( lp[c][0], lp[c][1], lp[c][2] ) );
            cellKeys[key] = 0;
         }
   }
   return buildMesh( nodes, cellKeys, "cubed domain", err );
}

// MeshInfo::meshSymmetricCuboid( lowerLeftFront, upperRightBack, nx, ny, nz ) (mesh/MeshGenSymmetricCuboid.cpp:35-253):
// nx x ny x nz sub-cubes of 24 tetrahedra around the centres of their six faces. Sub-cubes are visited z-major,
// then y, then x; a sub-cube's 14 vertices (8 corners, 6 face centres) take the ID of the same point in an
// earlier sub-cube, and the points seen for the first time get new IDs in the order of their local index.
// Every sub-cube writes the coordinates of all its vertices, so a shared vertex keeps the value computed in
// the last sub-cube that contains it.
std::unique_ptr< Mesh > meshSymmetricCuboid( const double lo[3], const double hi[3], int nx, int ny, int nz, std::string& err )
{
   if ( nx < 1 || ny < 1 || nz < 1 )
   {
      err = "symmetric cuboid needs at least one sub-cube in every direction";
      return nullptr;
   }
   const int    n[3] = { nx, ny, nz };
   double       w[3];
   for ( int c = 0; c < 3; ++c )
      w[c] = ( hi[c] - lo[c] ) / double( n[c] );
   // local vertices in units of half a sub-cube: corners of the bottom face, its centre, the four side-face
   // centres, then the top face
   static const int lp[14][3] = { { 0, 0, 0 }, { 2, 0, 0 }, { 1, 1, 0 }, { 0, 2, 0 }, { 2, 2, 0 }, { 1, 0, 1 }, { 0, 1, 1 },
                                  { 2, 1, 1 }, { 1, 2, 1 }, { 0, 0, 2 }, { 2, 0, 2 }, { 1, 1, 2 }, { 0, 2, 2 }, { 2, 2, 2 } };
   static const int tets[24][4] = { { 0, 1, 2, 5 },    { 0, 3, 2, 6 },    { 1, 2, 4, 7 },    { 2, 4, 3, 8 },   { 2, 0, 5, 6 },
                                    { 2, 1, 5, 7 },    { 2, 3, 6, 8 },    { 2, 4, 7, 8 },    { 0, 9, 5, 6 },   { 1, 10, 5, 7 },
                                    { 3, 12, 6, 8 },   { 4, 13, 7, 8 },   { 2, 5, 6, 8 },    { 2, 5, 7, 8 },   { 11, 5, 6, 8 },
                                    { 11, 5, 7, 8 },   { 11, 9, 5, 6 },   { 11, 10, 5, 7 },  { 11, 12, 6, 8 }, { 11, 13, 7, 8 },
                                    { 11, 9, 10, 5 },  { 11, 9, 12, 6 },  { 11, 10, 13, 7 }, { 11, 12, 13, 8 } };
   std::map< std::array< int, 3 >, long > idOf;
   std::map< long, std::array< double, 3 > > nodes;
   std::map< std::vector< long >, int >      cellKeys;
   for ( int z = 0; z < nz; ++z )
      for ( int y = 0; y < ny; ++y )
         for ( int x = 0; x < nx; ++x )
         {
            const int q[3] = { x, y, z };
            long      id[14];
            for ( int i = 0; i < 14; ++i )
            {
               const std::array< int, 3 > key{ 2 * x + lp[i][0], 2 * y + lp[i][1], 2 * z + lp[i][2] };
               auto                       it = idOf.find( key );
               if ( it == idOf.end() )
               {
                  const long fresh = long( idOf.size() );
                  it               = idOf.emplace( key, fresh ).first;
               }
               id[i] = it->second;
               std::array< double, 3 > X;
               for ( int c = 0; c < 3; ++c )
               {
                  // offset of the sub-cube plus the local coordinate (0, w/2 or w), in that order
                  const double loc = lp[i][c] == 0 ? 0.0 : ( lp[i][c] == 1 ? w[c] / 2 : w[c] );
                  X[c]             = ( lo[c] + w[c] * double( q[c] ) ) + loc;
               }
               nodes[id[i]] = X;
            }
            for ( auto& t : tets )
               cellKeys[{ id[t[0]], id[t[1]], id[t[2]], id[t[3]] }] = 0;
         }
   return buildMesh( nodes, cellKeys, "symmetric cuboid", err );
}

i64 numDofs( const Mesh& m, int L )
{
   const i64 N = i64( 1 ) << L;
   return i64( m.nV() ) + i64( m.nE() ) * ( N - 1 ) + i64( m.nF() ) * ( ( N - 1 ) * ( N - 2 ) / 2 ) +
          i64( m.nC() ) * ( ( N - 1 ) * ( N - 2 ) * ( N - 3 ) / 6 );
}

} // namespace hyg
