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
// edge and vertex of such a face; cells are always inner.
#include "hyg_internal.hpp"

#include <algorithm>
#include <fstream>
#include <map>
#include <sstream>

namespace hyg {

static const int kCellFaceVerts[4][3] = { { 0, 1, 2 }, { 0, 1, 3 }, { 0, 2, 3 }, { 1, 2, 3 } };
static const int kCellEdgeVerts[6][2] = { { 0, 1 }, { 0, 2 }, { 1, 2 }, { 0, 3 }, { 1, 3 }, { 2, 3 } };

std::unique_ptr< Mesh > readGmsh( const std::string& path, std::string& err )
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
   for ( long k = 0; k < ne; ++k )
   {
      long id, type, ntags;
      in >> id >> type >> ntags;
      for ( long t = 0; t < ntags; ++t )
      {
         long tag;
         in >> tag;
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
   }
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

i64 numDofs( const Mesh& m, int L )
{
   const i64 N = i64( 1 ) << L;
   return i64( m.nV() ) + i64( m.nE() ) * ( N - 1 ) + i64( m.nF() ) * ( ( N - 1 ) * ( N - 2 ) / 2 ) +
          i64( m.nC() ) * ( ( N - 1 ) * ( N - 2 ) * ( N - 3 ) / 6 );
}

} // namespace hyg
