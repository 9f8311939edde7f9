// ORACLE — CPU restatement of HyTeG's P1 matrix-free geometric-multigrid path.
//
// This file is TEST INFRASTRUCTURE. It is built into oracle/liboracle.so and is only
// loaded by tests/, by __graft_entry__.smoke() (as the checker) and by bench.py's
// cpu_baseline leg. The product library never links, loads or calls it.
//
// It follows the reference's algorithms, data layout and loop order function by
// function (citations are relative to the reference repository root):
//   * macro-mesh construction  src/hyteg/mesh/MeshGenFromGmshFile.cpp:32-200,
//     src/hyteg/primitivestorage/SetupPrimitiveStorage.cpp:40-365 (primitive order,
//     face vertex order from sorted edges, cell vertex order = Gmsh element order),
//     boundary flags SetupPrimitiveStorage.cpp:781-792 + onBoundary :895-945;
//   * per-primitive arrays    src/hyteg/indexing/MacroCellIndexing.hpp:40-53,
//     MacroFaceIndexing.hpp:39-48 (owned data incl. boundary copies; the ghost layers
//     of faces/edges/vertices are replaced by direct reads of the neighbour cell arrays,
//     which hold identical values after the downward communication);
//   * downward halo refresh   P1Operator.hpp:205-211 (V->E, E->F, F->C chain);
//   * stencil assembly        P1Elements.hpp:89-143 (24 micro-tets at a vertex),
//     :463-527 calculateStencilInMacroCellForm_new, FEniCS p1_tet_diffusion
//     tabulate_tensor (src/hyteg/forms/form_fenics_generated/p1_tet_diffusion.h:4113);
//   * apply                   P1Operator.hpp:196-324, generated cell kernel
//     p1functionspace/generatedKernels/apply_3D_macrocell_vertexdof_to_vertexdof_replace.cpp:32-87;
//   * smooth_jac              P1Operator.hpp:433-451, WeightedJacobiSmoother.hpp:42-62;
//   * inverse diagonal        P1Operator.hpp:645-915, VertexDoFFunction.cpp:1524-1666;
//   * restriction             P1toP1LinearRestriction.cpp:168-345 with the leaf-location
//     scaling of generatedKernels/restrict_3D_macrocell_P1_pull_additive.cpp;
//   * prolongation            P1toP1LinearProlongation.cpp:194-398;
//   * additive communication  VertexDoFFunction.hpp:322-415;
//   * vector ops              VertexDoFFunction.cpp:1107-1199, 1366-1523, 1668-1753;
//   * Chebyshev smoother      solvers/ChebyshevSmoother.hpp:48-190, power iteration
//     numerictools/SpectrumEstimation.hpp:57-80;
//   * CG                      solvers/CGSolver.hpp:66-134, 237-248;
//   * GMG V-cycle             solvers/GeometricMultigridSolver.hpp:127-235;
//   * mixed-precision IR      tests/hyteg/mixedPrecision/mixedPrecisionIterativeRefinement.cpp:162-172;
//   * P2 (quadratic) path      P2ConstantOperator.cpp:99-113 (restated element by element, as
//     P2ElementwiseOperator.cpp:100-310), p2_tet_diffusion, P2toP2QuadraticProlongation.cpp:217-424,
//     P2toP2QuadraticRestriction.cpp:131-286, on the level-(L+1) vertex lattice.
//
// How each function is pinned (tests/test_oracle_*.py): the apply, smoother, transfer
// and V-cycle against an independent scipy global assembly of the FE matrices on the
// refined meshes (brute force), the stencils against A*1 = 0 / A*linear = 0, row sums
// and h-scaling, the transfer against exactness on linear functions and R = P^T,
// the Chebyshev coefficients against the closed-form Chebyshev polynomial, CG against
// scipy's direct solve. Parity unpinned: none of the functions exported here.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <fstream>
#include <climits>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

using idx = int64_t;

// ---------------------------------------------------------------- indexing
inline idx S3( idx w ) { return w * ( w + 1 ) * ( w + 2 ) / 6; }
inline idx T2( idx w ) { return w * ( w + 1 ) / 2; }
// MacroCellIndexing.hpp:46-53
inline idx cidx( idx w, idx x, idx y, idx z ) { return S3( w ) - S3( w - z ) + y * ( w - z + 1 ) - y * ( y + 1 ) / 2 + x; }
// MacroFaceIndexing.hpp:39-48
inline idx fidx( idx w, idx x, idx y ) { return y * ( w + 1 ) - y * ( y + 1 ) / 2 + x; }

enum DoFType : int { None = 0, Inner = 1, DirichletBoundary = 2, NeumannBoundary = 4, FreeslipBoundary = 8, All = 15 };

using I3 = std::array< idx, 3 >;
using P3 = std::array< double, 3 >;

// Stencil directions, VertexDoFIndexing.cpp:484-520
enum Dir { V_C, V_W, V_E, V_N, V_S, V_NW, V_SE, V_TC, V_TW, V_TS, V_TSE, V_BC, V_BN, V_BE, V_BNW, NDIR };
static const int DOFF[NDIR][3] = { { 0, 0, 0 },
                                   { -1, 0, 0 },
                                   { 1, 0, 0 },
                                   { 0, 1, 0 },
                                   { 0, -1, 0 },
                                   { -1, 1, 0 },
                                   { 1, -1, 0 },
                                   { 0, 0, 1 },
                                   { -1, 0, 1 },
                                   { 0, -1, 1 },
                                   { 1, -1, 1 },
                                   { 0, 0, -1 },
                                   { 0, 1, -1 },
                                   { 1, 0, -1 },
                                   { -1, 1, -1 } };

// The 24 micro-tets around an inner micro-vertex, P1Elements.hpp:93-143 (each starts with C).
static const int CELLS_AT_VERTEX[24][4] = {
    // white up
    { V_C, V_BC, V_BE, V_BN },
    { V_C, V_S, V_SE, V_TS },
    { V_C, V_W, V_NW, V_TW },
    { V_C, V_N, V_E, V_TC },
    // white down
    { V_C, V_W, V_BC, V_S },
    { V_C, V_E, V_SE, V_BE },
    { V_C, V_N, V_NW, V_BN },
    { V_C, V_TS, V_TC, V_TW },
    // blue up
    { V_C, V_BC, V_BN, V_BNW },
    { V_C, V_W, V_S, V_TS },
    { V_C, V_E, V_SE, V_TSE },
    { V_C, V_NW, V_N, V_TC },
    // blue down
    { V_C, V_BC, V_S, V_SE },
    { V_C, V_W, V_NW, V_BNW },
    { V_C, V_E, V_BN, V_N },
    { V_C, V_TC, V_TS, V_TSE },
    // green up
    { V_C, V_W, V_BC, V_BNW },
    { V_C, V_E, V_BE, V_BN },
    { V_C, V_TC, V_TW, V_NW },
    { V_C, V_SE, V_TS, V_TSE },
    // green down
    { V_C, V_BC, V_BE, V_SE },
    { V_C, V_BN, V_BNW, V_NW },
    { V_C, V_E, V_TSE, V_TC },
    { V_C, V_W, V_TS, V_TW } };

// ---------------------------------------------------------------- mesh / storage
struct Mesh
{
   std::vector< P3 >                   X;     // vertices ordered by Gmsh id
   std::vector< std::array< int, 2 > > E;     // sorted vertex pairs, map order
   std::vector< std::array< int, 3 > > F;     // sorted vertex triples, map order
   std::vector< std::array< int, 4 > > Cl;    // cell vertices in Gmsh element order, map order
   std::vector< std::array< int, 4 > > cF;    // cell-local face k -> face id
   std::vector< std::array< int, 6 > > cE;    // cell-local edge k -> edge id
   std::vector< std::vector< int > >   vC, eC, fC, vE, vF, eF;
   std::vector< int >                  vB, eB, fB; // 1 = on boundary (Dirichlet), 0 = inner
   // IcosahedralShellMap of every cell (empty: affine cells), see shellMapOf
   struct ShellMap
   {
      std::array< double, 3 > ref, ray, thr, fo;
      double                  rRef, rRay;
   };
   std::vector< ShellMap > shell;
};

static const int CELL_FACE_VERTS[4][3] = { { 0, 1, 2 }, { 0, 1, 3 }, { 0, 2, 3 }, { 1, 2, 3 } };
static const int CELL_EDGE_VERTS[6][2] = { { 0, 1 }, { 0, 2 }, { 1, 2 }, { 0, 3 }, { 1, 3 }, { 2, 3 } };

static std::unique_ptr< Mesh > buildMesh( const std::map< idx, P3 >& nodes, const std::set< std::vector< idx > >& cellsGmsh );

// keepTags: MeshInfo::fromGmshFile's own flags (MeshGenFromGmshFile.cpp:100-116, 160-200) instead of
// setMeshBoundaryFlagsOnBoundary: the parsed line and triangle elements carry their first tag into the
// EdgeContainer / FaceContainer (addEdge / addFace keep the first entry of a sorted vertex set), all else 0.
static std::unique_ptr< Mesh > loadGmsh( const std::string& path, bool keepTags = false )
{
   std::ifstream in( path );
   if ( !in )
      return nullptr;
   std::string tok;
   in >> tok >> tok >> tok >> tok >> tok >> tok; // $MeshFormat 2.2 0 8 $EndMeshFormat $Nodes
   if ( tok != "$Nodes" )
      return nullptr;
   idx nv;
   in >> nv;
   std::map< idx, P3 > nodes;
   for ( idx i = 0; i < nv; ++i )
   {
      idx id;
      P3  x;
      in >> id >> x[0] >> x[1] >> x[2];
      nodes[id] = x;
   }
   in >> tok >> tok; // $EndNodes $Elements
   idx ne;
   in >> ne;
   std::set< std::vector< idx > > cellsGmsh; // keyed like MeshInfo::CellContainer
   std::vector< std::vector< idx > > cellList;
   std::map< std::vector< idx >, idx > parsedEdges, parsedFaces; // file-order vertex lists -> physical tag
   for ( idx i = 0; i < ne; ++i )
   {
      idx id, type, ntags;
      in >> id >> type >> ntags;
      std::vector< idx > tags( ntags );
      for ( auto& t : tags )
         in >> t;
      int nn = type == 15 ? 1 : type == 1 ? 2 : type == 2 ? 3 : type == 4 ? 4 : -1;
      if ( nn < 0 )
         return nullptr;
      std::vector< idx > vs( nn );
      for ( auto& v : vs )
         in >> v;
      if ( type == 4 )
         cellsGmsh.insert( vs );
      else if ( type == 1 )
         parsedEdges[vs] = ntags > 0 ? tags[0] : 0;
      else if ( type == 2 )
         parsedFaces[vs] = ntags > 0 ? tags[0] : 0;
   }
   auto m = buildMesh( nodes, cellsGmsh );
   if ( m && keepTags )
   {
      std::map< idx, int > rank;
      int                  r = 0;
      for ( auto& kv : nodes )
         rank[kv.first] = r++;
      std::fill( m->vB.begin(), m->vB.end(), 0 );
      std::fill( m->eB.begin(), m->eB.end(), 0 );
      std::fill( m->fB.begin(), m->fB.end(), 0 );
      std::set< size_t > doneE, doneF;
      for ( auto& [vs, tag] : parsedEdges )
      {
         std::array< int, 2 > key{ rank.at( vs[0] ), rank.at( vs[1] ) };
         std::sort( key.begin(), key.end() );
         for ( size_t e = 0; e < m->E.size(); ++e )
            if ( m->E[e][0] == key[0] && m->E[e][1] == key[1] && doneE.insert( e ).second )
               m->eB[e] = int( tag );
      }
      for ( auto& [vs, tag] : parsedFaces )
      {
         std::array< int, 3 > key{ rank.at( vs[0] ), rank.at( vs[1] ), rank.at( vs[2] ) };
         std::sort( key.begin(), key.end() );
         for ( size_t f = 0; f < m->F.size(); ++f )
            if ( m->F[f][0] == key[0] && m->F[f][1] == key[1] && m->F[f][2] == key[2] && doneF.insert( f ).second )
               m->fB[f] = int( tag );
      }
   }
   return m;
}

// primitives of a MeshInfo given by its vertex container (ID -> coordinates) and cell container (vertex lists)
static std::unique_ptr< Mesh > buildMesh( const std::map< idx, P3 >& nodes, const std::set< std::vector< idx > >& cellsGmsh )
{
   auto m = std::make_unique< Mesh >();
   std::map< idx, int > vid;
   for ( auto& [id, x] : nodes )
   {
      vid[id] = (int) m->X.size();
      m->X.push_back( x );
   }
   std::set< std::array< int, 2 > > es;
   std::set< std::array< int, 3 > > fs;
   for ( auto& c : cellsGmsh )
   {
      std::array< int, 4 > cv{ vid[c[0]], vid[c[1]], vid[c[2]], vid[c[3]] };
      m->Cl.push_back( cv );
      for ( auto& e : CELL_EDGE_VERTS )
      {
         std::array< int, 2 > k{ cv[e[0]], cv[e[1]] };
         std::sort( k.begin(), k.end() );
         es.insert( k );
      }
      for ( auto& f : CELL_FACE_VERTS )
      {
         std::array< int, 3 > k{ cv[f[0]], cv[f[1]], cv[f[2]] };
         std::sort( k.begin(), k.end() );
         fs.insert( k );
      }
   }
   m->E.assign( es.begin(), es.end() );
   m->F.assign( fs.begin(), fs.end() );
   std::map< std::array< int, 2 >, int > eid;
   std::map< std::array< int, 3 >, int > fid;
   for ( size_t i = 0; i < m->E.size(); ++i )
      eid[m->E[i]] = (int) i;
   for ( size_t i = 0; i < m->F.size(); ++i )
      fid[m->F[i]] = (int) i;
   size_t nV = m->X.size(), nE = m->E.size(), nF = m->F.size(), nC = m->Cl.size();
   m->vC.resize( nV );
   m->eC.resize( nE );
   m->fC.resize( nF );
   m->vE.resize( nV );
   m->vF.resize( nV );
   m->eF.resize( nE );
   for ( size_t e = 0; e < nE; ++e )
      for ( int v : m->E[e] )
         m->vE[v].push_back( (int) e );
   for ( size_t f = 0; f < nF; ++f )
   {
      auto& t = m->F[f];
      for ( int v : t )
         m->vF[v].push_back( (int) f );
      m->eF[eid[{ t[0], t[1] }]].push_back( (int) f );
      m->eF[eid[{ t[0], t[2] }]].push_back( (int) f );
      m->eF[eid[{ t[1], t[2] }]].push_back( (int) f );
   }
   m->cF.resize( nC );
   m->cE.resize( nC );
   for ( size_t c = 0; c < nC; ++c )
   {
      auto& cv = m->Cl[c];
      for ( int k = 0; k < 4; ++k )
      {
         std::array< int, 3 > key{ cv[CELL_FACE_VERTS[k][0]], cv[CELL_FACE_VERTS[k][1]], cv[CELL_FACE_VERTS[k][2]] };
         std::sort( key.begin(), key.end() );
         m->cF[c][k] = fid[key];
         m->fC[fid[key]].push_back( (int) c );
      }
      for ( int k = 0; k < 6; ++k )
      {
         std::array< int, 2 > key{ cv[CELL_EDGE_VERTS[k][0]], cv[CELL_EDGE_VERTS[k][1]] };
         std::sort( key.begin(), key.end() );
         m->cE[c][k] = eid[key];
         m->eC[eid[key]].push_back( (int) c );
      }
      for ( int v : cv )
         m->vC[v].push_back( (int) c );
   }
   // onBoundary (SetupPrimitiveStorage.cpp:895-945): face with one cell; edge/vertex touching such a face.
   m->fB.assign( nF, 0 );
   m->eB.assign( nE, 0 );
   m->vB.assign( nV, 0 );
   for ( size_t f = 0; f < nF; ++f )
      if ( m->fC[f].size() == 1 )
      {
         m->fB[f] = 1;
         auto& t  = m->F[f];
         m->eB[eid[{ t[0], t[1] }]] = 1;
         m->eB[eid[{ t[0], t[2] }]] = 1;
         m->eB[eid[{ t[1], t[2] }]] = 1;
         for ( int v : t )
            m->vB[v] = 1;
      }
   return m;
}

// BoundaryCondition::create0123BC (boundary/BoundaryConditions.cpp:26-33): mesh flag 1 Dirichlet, 2 Neumann,
// 3 free-slip; any other flag takes the default, Inner
inline int bcType( int meshFlag )
{
   switch ( meshFlag )
   {
   case 1: return DirichletBoundary;
   case 2: return NeumannBoundary;
   case 3: return FreeslipBoundary;
   default: return Inner;
   }
}

// Cell coordinates of a point of a lower-dimensional primitive given by global vertex ids pv[0..nv-1]
// and barycentric weights bw[0..nv-1] (sum N); VertexDoFMacroCell / Cell local vertex maps.
inline I3 cellCoords( const Mesh& m, int c, const int* pv, const idx* bw, int nv )
{
   idx b[4] = { 0, 0, 0, 0 };
   for ( int k = 0; k < nv; ++k )
      for ( int l = 0; l < 4; ++l )
         if ( m.Cl[c][l] == pv[k] )
            b[l] = bw[k];
   return { b[1], b[2], b[3] };
}

inline bool inCell( idx N, idx x, idx y, idx z ) { return x >= 0 && y >= 0 && z >= 0 && x + y + z <= N; }

// ---------------------------------------------------------------- function memory
template < typename T >
struct Function
{
   const Mesh* m;
   int         minL, maxL;
   // [level][primitive] -> owned array (incl. boundary copies for edges/faces/cells)
   std::vector< std::vector< std::vector< T > > > v, e, f, c;
   Function( const Mesh* mesh, int lo, int hi )
   : m( mesh )
   , minL( lo )
   , maxL( hi )
   {
      v.resize( hi + 1 );
      e.resize( hi + 1 );
      f.resize( hi + 1 );
      c.resize( hi + 1 );
      for ( int L = lo; L <= hi; ++L )
      {
         idx w = ( idx( 1 ) << L ) + 1;
         v[L].assign( m->X.size(), std::vector< T >( 1, T( 0 ) ) );
         e[L].assign( m->E.size(), std::vector< T >( w, T( 0 ) ) );
         f[L].assign( m->F.size(), std::vector< T >( T2( w ), T( 0 ) ) );
         c[L].assign( m->Cl.size(), std::vector< T >( S3( w ), T( 0 ) ) );
      }
   }
   // communicate<Vertex,Edge>, <Edge,Face>, <Face,Cell> (VertexDoFPackInfo.hpp:161-201, 250-292, 411-505)
   void commDown( int L ) const
   {
      auto&     self = const_cast< Function& >( *this );
      const idx N    = idx( 1 ) << L;
      const idx w    = N + 1;
      for ( size_t ei = 0; ei < m->E.size(); ++ei )
      {
         self.e[L][ei][0] = v[L][m->E[ei][0]][0];
         self.e[L][ei][N] = v[L][m->E[ei][1]][0];
      }
      for ( size_t fi = 0; fi < m->F.size(); ++fi )
      {
         auto& t = m->F[fi];
         // face edges: (t0,t1) y=0 ; (t0,t2) x=0 ; (t1,t2) x+y=N
         for ( idx i = 0; i <= N; ++i )
         {
            self.f[L][fi][fidx( w, i, 0 )]     = edgeValue( L, t[0], t[1], i );
            self.f[L][fi][fidx( w, 0, i )]     = edgeValue( L, t[0], t[2], i );
            self.f[L][fi][fidx( w, N - i, i )] = edgeValue( L, t[1], t[2], i );
         }
      }
#pragma omp parallel for schedule( dynamic )
      for ( idx ci = 0; ci < (idx) m->Cl.size(); ++ci )
      {
         for ( int k = 0; k < 4; ++k )
         {
            int  fi = m->cF[ci][k];
            auto tv = m->F[fi];
            int  pv[3] = { tv[0], tv[1], tv[2] };
            for ( idx j = 0; j <= N; ++j )
               for ( idx i = 0; i + j <= N; ++i )
               {
                  idx bw[3] = { N - i - j, i, j };
                  I3  q     = cellCoords( *m, (int) ci, pv, bw, 3 );
                  self.c[L][ci][cidx( w, q[0], q[1], q[2] )] = f[L][fi][fidx( w, i, j )];
               }
         }
      }
   }
   // value at edge-local position i measured from vertex a towards vertex b
   T edgeValue( int L, int a, int b, idx i ) const
   {
      const idx N = idx( 1 ) << L;
      for ( size_t ei = 0; ei < m->E.size(); ++ei )
      {
         if ( m->E[ei][0] == a && m->E[ei][1] == b )
            return e[L][ei][i];
         if ( m->E[ei][0] == b && m->E[ei][1] == a )
            return e[L][ei][N - i];
      }
      return T( 0 );
   }
};

// canonical DoF order: vertices, edge interiors, face interiors, cell interiors
inline idx numDofs( const Mesh& m, int L )
{
   idx N = idx( 1 ) << L;
   return idx( m.X.size() ) + idx( m.E.size() ) * ( N - 1 ) + idx( m.F.size() ) * ( ( N - 1 ) * ( N - 2 ) / 2 ) +
          idx( m.Cl.size() ) * ( ( N - 1 ) * ( N - 2 ) * ( N - 3 ) / 6 );
}

template < typename T, typename Fn >
void forEachDof( const Mesh& m, int L, Fn&& fn )
{
   // fn(kind, primitive, i, j, k, canonicalIndex)
   const idx N = idx( 1 ) << L;
   idx       k = 0;
   for ( size_t vi = 0; vi < m.X.size(); ++vi )
      fn( 0, (int) vi, 0, 0, 0, k++ );
   for ( size_t ei = 0; ei < m.E.size(); ++ei )
      for ( idx i = 1; i <= N - 1; ++i )
         fn( 1, (int) ei, i, 0, 0, k++ );
   for ( size_t fi = 0; fi < m.F.size(); ++fi )
      for ( idx j = 1; j <= N - 2; ++j )
         for ( idx i = 1; i + j <= N - 1; ++i )
            fn( 2, (int) fi, i, j, 0, k++ );
   for ( size_t ci = 0; ci < m.Cl.size(); ++ci )
      for ( idx z = 1; z <= N - 3; ++z )
         for ( idx y = 1; y + z <= N - 2; ++y )
            for ( idx x = 1; x + y + z <= N - 1; ++x )
               fn( 3, (int) ci, x, y, z, k++ );
}

// ---------------------------------------------------------------- local form (FEniCS p1_tet_diffusion)
// Row 0 of the P1 Laplace element matrix: A_0j = |det J|/6 * (J^-T g0).(J^-T gj), g = reference gradients.
inline void p1DiffusionRow0( const P3* x, double* row )
{
   double J[3][3];
   for ( int r = 0; r < 3; ++r )
      for ( int k = 0; k < 3; ++k )
         J[r][k] = x[k + 1][r] - x[0][r];
   double det = J[0][0] * ( J[1][1] * J[2][2] - J[1][2] * J[2][1] ) - J[0][1] * ( J[1][0] * J[2][2] - J[1][2] * J[2][0] ) +
                J[0][2] * ( J[1][0] * J[2][1] - J[1][1] * J[2][0] );
   double K[3][3]; // K = J^-1
   K[0][0] = ( J[1][1] * J[2][2] - J[1][2] * J[2][1] ) / det;
   K[0][1] = ( J[0][2] * J[2][1] - J[0][1] * J[2][2] ) / det;
   K[0][2] = ( J[0][1] * J[1][2] - J[0][2] * J[1][1] ) / det;
   K[1][0] = ( J[1][2] * J[2][0] - J[1][0] * J[2][2] ) / det;
   K[1][1] = ( J[0][0] * J[2][2] - J[0][2] * J[2][0] ) / det;
   K[1][2] = ( J[0][2] * J[1][0] - J[0][0] * J[1][2] ) / det;
   K[2][0] = ( J[1][0] * J[2][1] - J[1][1] * J[2][0] ) / det;
   K[2][1] = ( J[0][1] * J[2][0] - J[0][0] * J[2][1] ) / det;
   K[2][2] = ( J[0][0] * J[1][1] - J[0][1] * J[1][0] ) / det;
   const double g[4][3] = { { -1, -1, -1 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
   double       G[4][3];
   for ( int i = 0; i < 4; ++i )
      for ( int r = 0; r < 3; ++r )
         G[i][r] = K[0][r] * g[i][0] + K[1][r] * g[i][1] + K[2][r] * g[i][2];
   const double vol = std::fabs( det ) / 6.0;
   for ( int j = 0; j < 4; ++j )
      row[j] = vol * ( G[0][0] * G[j][0] + G[0][1] * G[j][1] + G[0][2] * G[j][2] );
}

// VertexDoFMacroCell.hpp:68-75
inline P3 microCoord( const Mesh& m, int c, int L, idx x, idx y, idx z )
{
   const double h  = 1.0 / double( idx( 1 ) << L );
   const P3&    v0 = m.X[m.Cl[c][0]];
   const P3&    v1 = m.X[m.Cl[c][1]];
   const P3&    v2 = m.X[m.Cl[c][2]];
   const P3&    v3 = m.X[m.Cl[c][3]];
   P3           r;
   for ( int k = 0; k < 3; ++k )
      r[k] = v0[k] + ( v1[k] - v0[k] ) * h * double( x ) + ( v2[k] - v0[k] ) * h * double( y ) + ( v3[k] - v0[k] ) * h * double( z );
   return r;
}

// ---------------------------------------------------------------- spherical-shell blending
// IcosahedralShellMap (geometry/IcosahedralShellMap.hpp): classifyTet :281-317, classifyVertices :346-470,
// findNonSkewTetInPrism :473-515, evalF :92-128. Two vertices of a shell cell lie on one ray from the origin; with
// one vertex on one sphere and three on the other the lone one is `ref`, its ray partner `ray`, the others `thr` (the
// later one in local order) and `fo`; a cell with two vertices on each sphere takes the map of the cell across its
// first local face that does not contain both ray vertices.
inline Mesh::ShellMap shellMapOf( const Mesh& m, int c, int depth = 0 )
{
   std::array< P3, 4 >     v;
   std::array< double, 4 > rad;
   double                  lo = 1e300, hi = 0.0;
   for ( int k = 0; k < 4; ++k )
   {
      v[k]   = m.X[m.Cl[c][k]];
      rad[k] = std::sqrt( v[k][0] * v[k][0] + v[k][1] * v[k][1] + v[k][2] * v[k][2] );
      lo     = std::min( lo, rad[k] );
      hi     = std::max( hi, rad[k] );
   }
   int nOuter = 0;
   for ( int k = 0; k < 4; ++k )
      nOuter += ( hi - rad[k] ) < ( hi - lo ) * 0.001 ? 1 : 0;
   int a = -1, b = -1;
   for ( int k = 0; k < 4 && a < 0; ++k )
      for ( int j = k + 1; j < 4 && a < 0; ++j )
      {
         const double cx = v[k][1] * v[j][2] - v[k][2] * v[j][1], cy = v[k][2] * v[j][0] - v[k][0] * v[j][2],
                      cz = v[k][0] * v[j][1] - v[k][1] * v[j][0], s = rad[k] * rad[j];
         if ( std::fabs( cx ) / s <= 1e-14 && std::fabs( cy ) / s <= 1e-14 && std::fabs( cz ) / s <= 1e-14 )
         {
            a = k;
            b = j;
         }
      }
   if ( a < 0 || nOuter < 1 || nOuter > 3 || depth > 3 )
      throw std::runtime_error( "cell is not a spherical-shell cell" );
   if ( nOuter == 2 )
   {
      for ( int f = 0; f < 4; ++f )
      {
         bool hasA = false, hasB = false;
         for ( int t = 0; t < 3; ++t )
         {
            hasA = hasA || CELL_FACE_VERTS[f][t] == a;
            hasB = hasB || CELL_FACE_VERTS[f][t] == b;
         }
         if ( hasA && hasB )
            continue;
         for ( int o : m.fC[m.cF[c][f]] )
            if ( o != c )
               return shellMapOf( m, o, depth + 1 );
      }
      throw std::runtime_error( "no non-skew cell next to a skew shell cell" );
   }
   int ref = a, ray = b;
   if ( ( nOuter == 1 && rad[ref] < rad[ray] ) || ( nOuter == 3 && rad[ref] > rad[ray] ) )
      std::swap( ref, ray );
   int thr = -1, fo = -1;
   for ( int k = 0; k < 4; ++k )
      if ( k != ref && k != ray )
         thr = k;
   for ( int k = 0; k < 4; ++k )
      if ( k != ref && k != ray && k != thr )
         fo = k;
   return { v[ref], v[ray], v[thr], v[fo], rad[ref], rad[ray] };
}

// lambda(x): fraction of the way from the plane (ray, thr, fo) to ref; and its gradient (with the sign of |.|)
inline double shellLambda( const Mesh::ShellMap& s, const P3& x, double g[3] )
{
   double a[3], b[3], d[3], n[3];
   for ( int k = 0; k < 3; ++k )
   {
      a[k] = s.thr[k] - s.ray[k];
      b[k] = s.fo[k] - s.ray[k];
      d[k] = s.ref[k] - s.ray[k];
   }
   n[0]            = a[1] * b[2] - a[2] * b[1];
   n[1]            = a[2] * b[0] - a[0] * b[2];
   n[2]            = a[0] * b[1] - a[1] * b[0];
   const double vt = d[0] * n[0] + d[1] * n[1] + d[2] * n[2];
   const double vx = ( x[0] - s.ray[0] ) * n[0] + ( x[1] - s.ray[1] ) * n[1] + ( x[2] - s.ray[2] ) * n[2];
   const double sg = vx / vt < 0.0 ? -1.0 : 1.0;
   for ( int k = 0; k < 3; ++k )
      g[k] = sg * n[k] / vt;
   return std::fabs( vx / vt );
}

// evalF: x scaled to the radius r_ray + lambda (r_ref - r_ray)
inline P3 shellF( const Mesh::ShellMap& s, const P3& x )
{
   double       g[3];
   const double lam = shellLambda( s, x, g );
   const double r   = s.rRay + lam * ( s.rRef - s.rRay );
   const double n   = std::sqrt( x[0] * x[0] + x[1] * x[1] + x[2] * x[2] );
   return { x[0] * r / n, x[1] * r / n, x[2] * r / n };
}

// Jacobian of shellF, written out from the formula: DF = (r/|x|) I + x (x) ((r_ref - r_ray) grad lambda / |x| - r x / |x|^3)
inline void shellDF( const Mesh::ShellMap& s, const P3& x, double D[3][3] )
{
   double       g[3];
   const double lam = shellLambda( s, x, g );
   const double r   = s.rRay + lam * ( s.rRef - s.rRay );
   const double n   = std::sqrt( x[0] * x[0] + x[1] * x[1] + x[2] * x[2] );
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j )
         D[i][j] = ( i == j ? r / n : 0.0 ) + x[i] * ( ( s.rRef - s.rRay ) * g[j] / n - r * x[j] / ( n * n * n ) );
}

// p1_diffusion_blending_q3::integrateAll (forms/form_hyteg_generated/p1/p1_diffusion_blending_q3.cpp:287-...):
// a_ij = |det J| sum_q w_q |det DF_q| (DF_q^-T grad lambda_i) . (DF_q^-T grad lambda_j), DF at the affine quadrature
// points of the 5-point degree-3 rule (centroid -2/15, points (1/2, 1/6, 1/6, 1/6) 3/40)
inline double baryGrads( const P3* x, double G[4][3] );
inline void blendingLaplaceElement( const P3* x, const Mesh::ShellMap& s, double a[4][4] )
{
   double       G[4][3];
   const double det = baryGrads( x, G );
   const double lam[5][4] = { { 0.25, 0.25, 0.25, 0.25 },
                              { 1.0 / 6, 1.0 / 6, 1.0 / 6, 0.5 },
                              { 1.0 / 6, 1.0 / 6, 0.5, 1.0 / 6 },
                              { 1.0 / 6, 0.5, 1.0 / 6, 1.0 / 6 },
                              { 0.5, 1.0 / 6, 1.0 / 6, 1.0 / 6 } };
   const double wq[5]     = { -2.0 / 15.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0 };
   for ( int i = 0; i < 4; ++i )
      for ( int j = 0; j < 4; ++j )
         a[i][j] = 0.0;
   for ( int q = 0; q < 5; ++q )
   {
      P3 xq{ 0, 0, 0 };
      for ( int v = 0; v < 4; ++v )
         for ( int c = 0; c < 3; ++c )
            xq[c] += lam[q][v] * x[v][c];
      double D[3][3];
      shellDF( s, xq, D );
      // inverse of D by cofactors
      const double dd = D[0][0] * ( D[1][1] * D[2][2] - D[1][2] * D[2][1] ) - D[0][1] * ( D[1][0] * D[2][2] - D[1][2] * D[2][0] ) +
                        D[0][2] * ( D[1][0] * D[2][1] - D[1][1] * D[2][0] );
      double Di[3][3];
      Di[0][0] = ( D[1][1] * D[2][2] - D[1][2] * D[2][1] ) / dd;
      Di[0][1] = ( D[0][2] * D[2][1] - D[0][1] * D[2][2] ) / dd;
      Di[0][2] = ( D[0][1] * D[1][2] - D[0][2] * D[1][1] ) / dd;
      Di[1][0] = ( D[1][2] * D[2][0] - D[1][0] * D[2][2] ) / dd;
      Di[1][1] = ( D[0][0] * D[2][2] - D[0][2] * D[2][0] ) / dd;
      Di[1][2] = ( D[0][2] * D[1][0] - D[0][0] * D[1][2] ) / dd;
      Di[2][0] = ( D[1][0] * D[2][1] - D[1][1] * D[2][0] ) / dd;
      Di[2][1] = ( D[0][1] * D[2][0] - D[0][0] * D[2][1] ) / dd;
      Di[2][2] = ( D[0][0] * D[1][1] - D[0][1] * D[1][0] ) / dd;
      // physical gradients: row vector g_i^T Di
      double P[4][3];
      for ( int i = 0; i < 4; ++i )
         for ( int c = 0; c < 3; ++c )
            P[i][c] = G[i][0] * Di[0][c] + G[i][1] * Di[1][c] + G[i][2] * Di[2][c];
      const double s_ = wq[q] * std::fabs( dd );
      for ( int i = 0; i < 4; ++i )
         for ( int j = 0; j < 4; ++j )
            a[i][j] += s_ * ( P[i][0] * P[j][0] + P[i][1] * P[j][1] + P[i][2] * P[j][2] );
   }
   for ( int i = 0; i < 4; ++i )
      for ( int j = 0; j < 4; ++j )
         a[i][j] *= std::fabs( det );
}

// calculateStencilInMacroCellForm_new (P1Elements.hpp:463-527): sum row 0 over the micro-tets at q inside cell c.
inline std::array< double, NDIR > stencilInCell( const Mesh& m, int c, int L, const I3& q )
{
   std::array< double, NDIR > s{};
   const idx                  N = idx( 1 ) << L;
   for ( auto& el : CELLS_AT_VERTEX )
   {
      bool inside = true;
      P3   xs[4];
      for ( int k = 0; k < 4; ++k )
      {
         idx x = q[0] + DOFF[el[k]][0], y = q[1] + DOFF[el[k]][1], z = q[2] + DOFF[el[k]][2];
         if ( !inCell( N, x, y, z ) )
            inside = false;
         else
            xs[k] = microCoord( m, c, L, x, y, z );
      }
      if ( !inside )
         continue;
      double row[4];
      p1DiffusionRow0( xs, row );
      for ( int k = 0; k < 4; ++k )
         s[el[k]] += row[k];
   }
   return s;
}

// ---------------------------------------------------------------- P1 constant operator
struct LowerStencil
{
   int                                       nCells;
   std::vector< int >                        cells;
   std::vector< std::array< double, NDIR > > w; // per neighbour cell
};

// Operator<Src, Dst>::apply (operators/Operator.hpp:73-80) plus the inverse diagonal used by the smoothers
template < typename T >
struct OperatorBase
{
   const Mesh*                      m = nullptr;
   int                              minL = 0, maxL = 0;
   std::unique_ptr< Function< T > > invDiag;
   virtual ~OperatorBase() = default;
   virtual void apply( const Function< T >& src, Function< T >& dst, int L, int flag, bool add ) const = 0;
   // grid transfers of the operator's function space (P1 linear by default, defined after the transfers)
   virtual void restrictOp( Function< T >& f, int srcL, int flag ) const;
   virtual void prolongateOp( Function< T >& f, int srcL, int flag, bool add ) const;
};

template < typename T >
struct P1ConstantOperator : OperatorBase< T >
{
   using OperatorBase< T >::m;
   using OperatorBase< T >::invDiag;
   // [level]
   std::vector< std::vector< std::array< double, NDIR > > > cellSt;            // per cell
   std::vector< std::vector< LowerStencil > >               vSt, eSt, fSt;     // per primitive

   P1ConstantOperator( const Mesh* mesh, int lo, int hi )
   {
      this->m    = mesh;
      this->minL = lo;
      this->maxL = hi;
      cellSt.resize( hi + 1 );
      vSt.resize( hi + 1 );
      eSt.resize( hi + 1 );
      fSt.resize( hi + 1 );
      for ( int L = lo; L <= hi; ++L )
         assemble( L );
      invDiag = std::make_unique< Function< T > >( m, lo, hi );
      for ( int L = lo; L <= hi; ++L )
         computeInverseDiagonal( L );
   }

   LowerStencil lowerStencil( int L, const std::vector< int >& cells, const int* pv, const idx* bw, int nv ) const
   {
      LowerStencil ls;
      ls.nCells = (int) cells.size();
      ls.cells  = cells;
      for ( int c : cells )
      {
         I3 q = cellCoords( *m, c, pv, bw, nv );
         ls.w.push_back( stencilInCell( *m, c, L, q ) );
      }
      return ls;
   }

   // P1ConstantOperator::assembleStencils (P1ConstantOperator.cpp:602-655), P1Operator.hpp:510-552, 1792-1950, 2071-2161
   void assemble( int L )
   {
      const idx N = idx( 1 ) << L;
      cellSt[L].resize( m->Cl.size() );
      if ( L >= 2 )
         for ( size_t c = 0; c < m->Cl.size(); ++c )
            cellSt[L][c] = stencilInCell( *m, (int) c, L, { 1, 1, 1 } );
      for ( size_t vi = 0; vi < m->X.size(); ++vi )
      {
         int pv[1] = { (int) vi };
         idx bw[1] = { N };
         vSt[L].push_back( lowerStencil( L, m->vC[vi], pv, bw, 1 ) );
      }
      for ( size_t ei = 0; ei < m->E.size(); ++ei )
      {
         int pv[2] = { m->E[ei][0], m->E[ei][1] };
         idx bw[2] = { N - 1, 1 };
         eSt[L].push_back( L >= 1 ? lowerStencil( L, m->eC[ei], pv, bw, 2 ) : LowerStencil{} );
      }
      for ( size_t fi = 0; fi < m->F.size(); ++fi )
      {
         int pv[3] = { m->F[fi][0], m->F[fi][1], m->F[fi][2] };
         idx bw[3] = { N - 2, 1, 1 };
         fSt[L].push_back( L >= 2 ? lowerStencil( L, m->fC[fi], pv, bw, 3 ) : LowerStencil{} );
      }
   }

   // value of stencil sum at a lower-dim point (reads the neighbour cells' arrays)
   T lowerApply( const Function< T >& src, int L, const LowerStencil& ls, const int* pv, const idx* bw, int nv ) const
   {
      const idx w   = ( idx( 1 ) << L ) + 1;
      T         sum = T( 0 );
      for ( int k = 0; k < ls.nCells; ++k )
      {
         int         c   = ls.cells[k];
         I3          q   = cellCoords( *m, c, pv, bw, nv );
         const auto& arr = src.c[L][c];
         const idx   N   = w - 1;
         for ( int d = 0; d < NDIR; ++d )
         {
            if ( ls.w[k][d] == 0.0 )
               continue;
            idx x = q[0] + DOFF[d][0], y = q[1] + DOFF[d][1], z = q[2] + DOFF[d][2];
            if ( !inCell( N, x, y, z ) )
               continue;
            sum += T( ls.w[k][d] ) * arr[cidx( w, x, y, z )];
         }
      }
      return sum;
   }

   void apply( const Function< T >& src, Function< T >& dst, int L, int flag, bool add ) const override
   {
      src.commDown( L );
      const idx N = idx( 1 ) << L;
      const idx w = N + 1;
      for ( size_t vi = 0; vi < m->X.size(); ++vi )
      {
         if ( !( bcType( m->vB[vi] ) & flag ) )
            continue;
         int pv[1] = { (int) vi };
         idx bw[1] = { N };
         T   r     = lowerApply( src, L, vSt[L][vi], pv, bw, 1 );
         dst.v[L][vi][0] = add ? dst.v[L][vi][0] + r : r;
      }
      if ( L >= 1 )
         for ( size_t ei = 0; ei < m->E.size(); ++ei )
         {
            if ( !( bcType( m->eB[ei] ) & flag ) )
               continue;
            int pv[2] = { m->E[ei][0], m->E[ei][1] };
            for ( idx i = 1; i <= N - 1; ++i )
            {
               idx bw[2] = { N - i, i };
               T   r     = lowerApply( src, L, eSt[L][ei], pv, bw, 2 );
               dst.e[L][ei][i] = add ? dst.e[L][ei][i] + r : r;
            }
         }
      if ( L >= 2 )
      {
         for ( size_t fi = 0; fi < m->F.size(); ++fi )
         {
            if ( !( bcType( m->fB[fi] ) & flag ) )
               continue;
            int pv[3] = { m->F[fi][0], m->F[fi][1], m->F[fi][2] };
            for ( idx j = 1; j <= N - 2; ++j )
               for ( idx i = 1; i + j <= N - 1; ++i )
               {
                  idx bw[3] = { N - i - j, i, j };
                  T   r     = lowerApply( src, L, fSt[L][fi], pv, bw, 3 );
                  auto& d   = dst.f[L][fi][fidx( w, i, j )];
                  d         = add ? d + r : r;
               }
         }
         if ( flag & Inner )
         {
            const idx nC = (idx) m->Cl.size();
#pragma omp parallel for collapse( 2 ) schedule( dynamic, 4 )
            for ( idx ci = 0; ci < nC; ++ci )
               for ( idx z = 1; z <= N - 3; ++z )
                  applyCellPlane( src.c[L][ci].data(), dst.c[L][ci].data(), cellSt[L][ci], L, z, add );
         }
      }
   }

   // SORSmoothable::smooth_sor( dst, rhs, relax, level, flag, backwards ) in the reference's order
   // (P1Operator.hpp:343-416): after the six communications every copy of dst is consistent; then (forward)
   // macro-vertices, edges, faces and cells are smoothed in turn, each primitive lexicographically (edges i = 1..N-1,
   // faces j outer / i inner, cells z, y, x: VertexDoFMacroEdge.hpp, VertexDoFMacroFace.hpp:751-810 smoothSOR3D,
   // generatedKernels/sor_3D_macrocell_P1.cpp), with a downward communication after each phase; backwards the
   // phases and the visiting order are reversed and the communications go upward. A primitive reads its
   // neighbours through its own memory: its own values and those of its closure are the current ones, the ghost
   // layers hold what the last communication left there. Hence the rule used here (dst keeps every copy up to
   // date, `old` is dst before the sweep): neighbour q of a point of primitive P is read from dst when q lies in P's
   // closure or in a primitive of higher dimension than P (which the sweep has not reached forward, and whose
   // updated values the upward communications bring backwards), and from `old` otherwise. Update:
   // x = (1 - relax) x + relax (b - sum_{q != p} a_pq x_q) / a_pp.
   void smoothSorOrdered( Function< T >& dst, const Function< T >& rhs, double relax, int L, int flag, bool backwards ) const
   {
      dst.commDown( L );
      rhs.commDown( L );
      const Function< T > old = dst;
      const idx           N   = idx( 1 ) << L, w = N + 1;
      // sum over the incident cells' one-sided stencils, centre excluded, reading each neighbour per the rule;
      // pv / bw: the point as barycentric weights over the global vertices pv of its primitive (dimension nv - 1)
      auto offSum = [&]( const LowerStencil& ls, const int* pv, const idx* bw, int nv, double& centre ) {
         double sum = 0.0;
         centre     = 0.0;
         for ( int k = 0; k < ls.nCells; ++k )
         {
            const int c = ls.cells[k];
            const I3  q = cellCoords( *m, c, pv, bw, nv );
            centre += ls.w[k][V_C];
            for ( int d = 1; d < NDIR; ++d )
            {
               if ( ls.w[k][d] == 0.0 )
                  continue;
               const idx x = q[0] + DOFF[d][0], y = q[1] + DOFF[d][1], z = q[2] + DOFF[d][2];
               if ( !inCell( N, x, y, z ) )
                  continue;
               const idx b[4] = { N - x - y - z, x, y, z };
               int       nz   = 0;
               bool      sub  = true;
               for ( int l = 0; l < 4; ++l )
                  if ( b[l] != 0 )
                  {
                     ++nz;
                     bool in = false;
                     for ( int k2 = 0; k2 < nv; ++k2 )
                        in = in || m->Cl[c][l] == pv[k2];
                     sub = sub && in;
                  }
               const bool cur = nz > nv || sub;
               sum += ls.w[k][d] * double( ( cur ? dst : old ).c[L][c][cidx( w, x, y, z )] );
            }
         }
         return sum;
      };
      // new value of a lower-dimensional point, written into its owned array slot and every cell copy
      auto updLower = [&]( const LowerStencil& ls, const int* pv, const idx* bw, int nv, T& own, T rhsv ) {
         double     centre;
         const double s  = offSum( ls, pv, bw, nv, centre );
         const T      nv_ = T( ( 1.0 - relax ) * double( own ) + relax * ( double( rhsv ) - s ) * ( 1.0 / centre ) );
         own              = nv_;
         for ( int k = 0; k < ls.nCells; ++k )
         {
            const I3 q = cellCoords( *m, ls.cells[k], pv, bw, nv );
            dst.c[L][ls.cells[k]][cidx( w, q[0], q[1], q[2] )] = nv_;
         }
      };
      auto vertices = [&]() {
         for ( size_t vi = 0; vi < m->X.size(); ++vi )
         {
            if ( !( bcType( m->vB[vi] ) & flag ) )
               continue;
            int pv[1] = { (int) vi };
            idx bw[1] = { N };
            updLower( vSt[L][vi], pv, bw, 1, dst.v[L][vi][0], rhs.v[L][vi][0] );
         }
      };
      auto edges = [&]() {
         if ( L < 1 )
            return;
         for ( size_t ei = 0; ei < m->E.size(); ++ei )
         {
            if ( !( bcType( m->eB[ei] ) & flag ) )
               continue;
            int pv[2] = { m->E[ei][0], m->E[ei][1] };
            for ( idx s = 1; s <= N - 1; ++s )
            {
               const idx i     = backwards ? N - s : s;
               idx       bw[2] = { N - i, i };
               updLower( eSt[L][ei], pv, bw, 2, dst.e[L][ei][i], rhs.e[L][ei][i] );
            }
         }
      };
      auto faces = [&]() {
         if ( L < 2 )
            return;
         std::vector< std::array< idx, 2 > > pts;
         for ( idx j = 1; j <= N - 2; ++j )
            for ( idx i = 1; i + j <= N - 1; ++i )
               pts.push_back( { i, j } );
         if ( backwards )
            std::reverse( pts.begin(), pts.end() );
         for ( size_t fi = 0; fi < m->F.size(); ++fi )
         {
            if ( !( bcType( m->fB[fi] ) & flag ) )
               continue;
            int pv[3] = { m->F[fi][0], m->F[fi][1], m->F[fi][2] };
            for ( const auto& ij : pts )
            {
               idx bw[3] = { N - ij[0] - ij[1], ij[0], ij[1] };
               updLower( fSt[L][fi], pv, bw, 3, dst.f[L][fi][fidx( w, ij[0], ij[1] )], rhs.f[L][fi][fidx( w, ij[0], ij[1] )] );
            }
         }
      };
      auto cells = [&]() {
         if ( L < 2 || !( flag & Inner ) )
            return;
         std::vector< std::array< idx, 3 > > pts;
         for ( idx z = 1; z <= N - 3; ++z )
            for ( idx y = 1; y + z <= N - 2; ++y )
               for ( idx x = 1; x + y + z <= N - 1; ++x )
                  pts.push_back( { x, y, z } );
         if ( backwards )
            std::reverse( pts.begin(), pts.end() );
         for ( size_t ci = 0; ci < m->Cl.size(); ++ci )
         {
            auto&       a  = dst.c[L][ci];
            const auto& r  = rhs.c[L][ci];
            const auto& st = cellSt[L][ci];
            for ( const auto& p : pts )
            {
               double sum = 0.0;
               for ( int d = 1; d < NDIR; ++d )
                  sum += st[d] * double( a[cidx( w, p[0] + DOFF[d][0], p[1] + DOFF[d][1], p[2] + DOFF[d][2] )] );
               const idx q = cidx( w, p[0], p[1], p[2] );
               a[q]        = T( ( 1.0 - relax ) * double( a[q] ) + relax * ( double( r[q] ) - sum ) * ( 1.0 / st[V_C] ) );
            }
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
      dst.commDown( L );
   }

   // generated kernel apply_3D_macrocell_vertexdof_to_vertexdof_{replace,add}.cpp:32-87, one z-plane:
   // the 15 coefficients are hoisted, the row offsets computed once per (y, z) row.
   static void applyCellPlane( const T* s, T* d, const std::array< double, NDIR >& st, int L, idx z, bool add )
   {
      const idx N = idx( 1 ) << L;
      const idx w = N + 1;
      T         c[NDIR];
      for ( int k = 0; k < NDIR; ++k )
         c[k] = T( st[k] );
      for ( idx y = 1; y + z <= N - 2; ++y )
      {
         const T* rB0 = s + cidx( w, 0, y, z - 1 );     // (y,   z-1)
         const T* rB1 = s + cidx( w, 0, y + 1, z - 1 ); // (y+1, z-1)
         const T* rS  = s + cidx( w, 0, y - 1, z );     // (y-1, z)
         const T* rC  = s + cidx( w, 0, y, z );         // (y,   z)
         const T* rN  = s + cidx( w, 0, y + 1, z );     // (y+1, z)
         const T* rTS = s + cidx( w, 0, y - 1, z + 1 ); // (y-1, z+1)
         const T* rT  = s + cidx( w, 0, y, z + 1 );     // (y,   z+1)
         T*       o   = d + cidx( w, 0, y, z );
         for ( idx x = 1; x + y + z <= N - 1; ++x )
         {
            T r = c[V_C] * rC[x] + c[V_W] * rC[x - 1] + c[V_E] * rC[x + 1] + c[V_N] * rN[x] + c[V_S] * rS[x] + c[V_NW] * rN[x - 1] +
                  c[V_SE] * rS[x + 1] + c[V_TC] * rT[x] + c[V_TW] * rT[x - 1] + c[V_TS] * rTS[x] + c[V_TSE] * rTS[x + 1] +
                  c[V_BC] * rB0[x] + c[V_BN] * rB1[x] + c[V_BE] * rB0[x + 1] + c[V_BNW] * rB1[x - 1];
            o[x] = add ? o[x] + r : r;
         }
      }
   }

   // computeDiagonalOperatorValues(invert=true), P1Operator.hpp:645-915 + invertElementwise
   void computeInverseDiagonal( int L )
   {
      const idx N = idx( 1 ) << L;
      const idx w = N + 1;
      auto&     D = *invDiag;
      for ( size_t vi = 0; vi < m->X.size(); ++vi )
      {
         double s = 0;
         for ( auto& ww : vSt[L][vi].w )
            s += ww[V_C];
         D.v[L][vi][0] = T( 1.0 / s );
      }
      if ( L >= 1 )
         for ( size_t ei = 0; ei < m->E.size(); ++ei )
         {
            double s = 0;
            for ( auto& ww : eSt[L][ei].w )
               s += ww[V_C];
            for ( idx i = 1; i <= N - 1; ++i )
               D.e[L][ei][i] = T( 1.0 / s );
         }
      if ( L >= 2 )
      {
         for ( size_t fi = 0; fi < m->F.size(); ++fi )
         {
            double s = 0;
            for ( auto& ww : fSt[L][fi].w )
               s += ww[V_C];
            for ( idx j = 1; j <= N - 2; ++j )
               for ( idx i = 1; i + j <= N - 1; ++i )
                  D.f[L][fi][fidx( w, i, j )] = T( 1.0 / s );
         }
         for ( size_t ci = 0; ci < m->Cl.size(); ++ci )
            for ( idx z = 1; z <= N - 3; ++z )
               for ( idx y = 1; y + z <= N - 2; ++y )
                  for ( idx x = 1; x + y + z <= N - 1; ++x )
                     D.c[L][ci][cidx( w, x, y, z )] = T( 1.0 / cellSt[L][ci][V_C] );
      }
   }
};

// ---------------------------------------------------------------- vector operations (owned points, flag-filtered)
template < typename T, typename Fn >
void forOwned( const Mesh& m, int L, int flag, Fn&& fn )
{
   // fn(kind, prim, arrayIndex)
   const idx N = idx( 1 ) << L;
   const idx w = N + 1;
   for ( size_t vi = 0; vi < m.X.size(); ++vi )
      if ( bcType( m.vB[vi] ) & flag )
         fn( 0, (int) vi, 0 );
   for ( size_t ei = 0; ei < m.E.size(); ++ei )
      if ( bcType( m.eB[ei] ) & flag )
         for ( idx i = 1; i <= N - 1; ++i )
            fn( 1, (int) ei, i );
   for ( size_t fi = 0; fi < m.F.size(); ++fi )
      if ( bcType( m.fB[fi] ) & flag )
         for ( idx j = 1; j <= N - 2; ++j )
            for ( idx i = 1; i + j <= N - 1; ++i )
               fn( 2, (int) fi, fidx( w, i, j ) );
   if ( flag & Inner )
      for ( size_t ci = 0; ci < m.Cl.size(); ++ci )
         for ( idx z = 1; z <= N - 3; ++z )
            for ( idx y = 1; y + z <= N - 2; ++y )
               for ( idx x = 1; x + y + z <= N - 1; ++x )
                  fn( 3, (int) ci, cidx( w, x, y, z ) );
}

template < typename T >
inline T& at( Function< T >& f, int L, int kind, int p, idx i )
{
   switch ( kind )
   {
   case 0:
      return f.v[L][p][i];
   case 1:
      return f.e[L][p][i];
   case 2:
      return f.f[L][p][i];
   default:
      return f.c[L][p][i];
   }
}

template < typename T >
void assign( Function< T >& dst, const std::vector< double >& s, const std::vector< const Function< T >* >& src, int L, int flag )
{
   forOwned< T >( *dst.m, L, flag, [&]( int kind, int p, idx i ) {
      T r = T( 0 );
      for ( size_t k = 0; k < s.size(); ++k )
         r += T( s[k] ) * at( const_cast< Function< T >& >( *src[k] ), L, kind, p, i );
      at( dst, L, kind, p, i ) = r;
   } );
}

template < typename T >
void addScaled( Function< T >& dst, double s, const Function< T >& src, int L, int flag )
{
   forOwned< T >( *dst.m, L, flag, [&]( int kind, int p, idx i ) {
      at( dst, L, kind, p, i ) += T( s ) * at( const_cast< Function< T >& >( src ), L, kind, p, i );
   } );
}

template < typename T >
void multElementwise( Function< T >& dst, const Function< T >& a, const Function< T >& b, int L, int flag )
{
   forOwned< T >( *dst.m, L, flag, [&]( int kind, int p, idx i ) {
      at( dst, L, kind, p, i ) = at( const_cast< Function< T >& >( a ), L, kind, p, i ) *
                                 at( const_cast< Function< T >& >( b ), L, kind, p, i );
   } );
}

template < typename T >
void interpolate( Function< T >& dst, double val, int L, int flag )
{
   forOwned< T >( *dst.m, L, flag, [&]( int kind, int p, idx i ) { at( dst, L, kind, p, i ) = T( val ); } );
}

template < typename T >
double dot( const Function< T >& a, const Function< T >& b, int L, int flag )
{
   T s = T( 0 );
   forOwned< T >( *a.m, L, flag, [&]( int kind, int p, idx i ) {
      s += at( const_cast< Function< T >& >( a ), L, kind, p, i ) * at( const_cast< Function< T >& >( b ), L, kind, p, i );
   } );
   return double( s );
}

template < typename T >
void copyFrom( Function< T >& dst, const Function< double >& src, int L )
{
   forOwned< T >( *dst.m, L, All, [&]( int kind, int p, idx i ) {
      at( dst, L, kind, p, i ) = T( at( const_cast< Function< double >& >( src ), L, kind, p, i ) );
   } );
}

template < typename T >
void copyTo( Function< double >& dst, const Function< T >& src, int L )
{
   forOwned< T >( *dst.m, L, All, [&]( int kind, int p, idx i ) {
      at( dst, L, kind, p, i ) = double( at( const_cast< Function< T >& >( src ), L, kind, p, i ) );
   } );
}

// ---------------------------------------------------------------- grid transfer
// classification of a cell-local point: returns (kind, local id): kind 3 interior, 2 face, 1 edge, 0 vertex
inline std::pair< int, int > classify( idx N, idx x, idx y, idx z )
{
   idx b[4] = { N - x - y - z, x, y, z };
   int nz = 0, zmask = 0;
   for ( int k = 0; k < 4; ++k )
      if ( b[k] == 0 )
      {
         ++nz;
         zmask |= 1 << k;
      }
   if ( nz == 0 )
      return { 3, 0 };
   if ( nz == 1 )
   {
      // the face opposite to the zero coordinate: b0=0 -> face3 (1,2,3), x=0 -> face2 (0,2,3), y=0 -> face1, z=0 -> face0
      int zk = __builtin_ctz( zmask );
      return { 2, 3 - zk };
   }
   if ( nz == 2 )
   {
      int nzm = ( ~zmask ) & 15;
      int a   = __builtin_ctz( nzm );
      int bb  = 31 - __builtin_clz( nzm );
      for ( int k = 0; k < 6; ++k )
         if ( CELL_EDGE_VERTS[k][0] == a && CELL_EDGE_VERTS[k][1] == bb )
            return { 1, k };
   }
   int nzm = ( ~zmask ) & 15;
   return { 0, __builtin_ctz( nzm ) };
}

inline double invNumCells( const Mesh& m, int c, const std::pair< int, int >& cl )
{
   switch ( cl.first )
   {
   case 3:
      return 1.0;
   case 2:
      return 1.0 / double( m.fC[m.cF[c][cl.second]].size() );
   case 1:
      return 1.0 / double( m.eC[m.cE[c][cl.second]].size() );
   default:
      return 1.0 / double( m.vC[m.Cl[c][cl.second]].size() );
   }
}

// additive communication Cell -> {Vertex, Edge, Face} (VertexDoFAdditivePackInfo.hpp:378-571):
// lower-dim owners receive the sum of the cells' boundary values; excluded primitives are skipped.
template < typename T >
void commAdditiveCellToLower( Function< T >& fn, int L, int excludeFlag, bool zero )
{
   const Mesh& m = *fn.m;
   const idx   N = idx( 1 ) << L;
   const idx   w = N + 1;
   for ( size_t vi = 0; vi < m.X.size(); ++vi )
   {
      if ( bcType( m.vB[vi] ) & excludeFlag )
         continue;
      T   s     = zero ? T( 0 ) : fn.v[L][vi][0];
      int pv[1] = { (int) vi };
      idx bw[1] = { N };
      for ( int c : m.vC[vi] )
      {
         I3 q = cellCoords( m, c, pv, bw, 1 );
         s += fn.c[L][c][cidx( w, q[0], q[1], q[2] )];
      }
      fn.v[L][vi][0] = s;
   }
   for ( size_t ei = 0; ei < m.E.size(); ++ei )
   {
      if ( bcType( m.eB[ei] ) & excludeFlag )
         continue;
      int pv[2] = { m.E[ei][0], m.E[ei][1] };
      for ( idx i = 1; i <= N - 1; ++i )
      {
         T   s     = zero ? T( 0 ) : fn.e[L][ei][i];
         idx bw[2] = { N - i, i };
         for ( int c : m.eC[ei] )
         {
            I3 q = cellCoords( m, c, pv, bw, 2 );
            s += fn.c[L][c][cidx( w, q[0], q[1], q[2] )];
         }
         fn.e[L][ei][i] = s;
      }
   }
   for ( size_t fi = 0; fi < m.F.size(); ++fi )
   {
      if ( bcType( m.fB[fi] ) & excludeFlag )
         continue;
      int pv[3] = { m.F[fi][0], m.F[fi][1], m.F[fi][2] };
      for ( idx j = 1; j <= N - 2; ++j )
         for ( idx i = 1; i + j <= N - 1; ++i )
         {
            T&  d     = fn.f[L][fi][fidx( w, i, j )];
            T   s     = zero ? T( 0 ) : d;
            idx bw[3] = { N - i - j, i, j };
            for ( int c : m.fC[fi] )
            {
               I3 q = cellCoords( m, c, pv, bw, 3 );
               s += fn.c[L][c][cidx( w, q[0], q[1], q[2] )];
            }
            d = s;
         }
   }
}

// ---------------------------------------------------------------- variable-coefficient elementwise operator
// MonomialBasis3D (polynomial/MonomialBasis3D.hpp): monomials x^a y^b z^c ordered by total degree, within a degree
// by decreasing a, then decreasing b. Exponents of basis function i.
inline int monomialCount3( int deg ) { return ( deg + 1 ) * ( deg + 2 ) * ( deg + 3 ) / 6; }
inline std::array< int, 3 > monomialExponents3( int i )
{
   for ( int d = 0;; ++d )
      for ( int a = d; a >= 0; --a )
         for ( int b = d - a; b >= 0; --b )
            if ( i-- == 0 )
               return { a, b, d - a - b };
}
inline double monomial3( int i, const P3& x )
{
   const auto e = monomialExponents3( i );
   return std::pow( x[0], e[0] ) * std::pow( x[1], e[1] ) * std::pow( x[2], e[2] );
}

// Coefficient k(x) of -div(k grad u): kind 0: k0 exp( sum_m a_m cos(2 pi kappa_m . x + phi_m) ),
// kind 1: tanh(alpha (x - 1/2)) + 2 (the coefficient of ElementwiseDivKGradCGConvergenceTest.cpp:60-62),
// kind 2: a polynomial sum_i poly_i m_i(x) in the monomial basis above (the coefficients k1..k4 of
// tests/hyteg/operators/P1SurrogateOperatorTest.cpp:76-112 are of this kind).
struct KField
{
   int                                   kind  = 0;
   double                                k0    = 1.0;
   double                                alpha = 2.0;
   std::vector< std::array< double, 5 > > modes; // kappa_x, kappa_y, kappa_z, phi, a
   std::vector< double >                 poly;
   double operator()( const P3& x ) const
   {
      if ( kind == 1 )
         return std::tanh( alpha * ( x[0] - 0.5 ) ) + 2.0;
      if ( kind == 2 )
      {
         double v = 0.0;
         for ( size_t i = 0; i < poly.size(); ++i )
            v += poly[i] * monomial3( int( i ), x );
         return v;
      }
      double g = 0.0;
      for ( auto& md : modes )
         g += md[4] * std::cos( 2.0 * M_PI * ( md[0] * x[0] + md[1] * x[1] + md[2] * x[2] ) + md[3] );
      return k0 * std::exp( g );
   }
};

// The six micro-tetrahedron types of a refined macro-cell (WHITE_UP, BLUE_UP, GREEN_UP, WHITE_DOWN,
// BLUE_DOWN, GREEN_DOWN), lattice offsets of their vertices from the micro-cell index
// (volumedofspace/CellDoFIndexing.hpp:155-198, vertexdof getVertexDoFDataIndicesFromMicroCell).
static const int MICRO_CELL[6][4][3] = {
    { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } }, { { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 }, { 1, 0, 1 } },
    { { 1, 0, 0 }, { 0, 1, 0 }, { 1, 0, 1 }, { 0, 0, 1 } }, { { 1, 1, 0 }, { 1, 1, 1 }, { 0, 1, 1 }, { 1, 0, 1 } },
    { { 1, 0, 1 }, { 0, 1, 1 }, { 0, 0, 1 }, { 0, 1, 0 } }, { { 0, 1, 0 }, { 1, 1, 0 }, { 1, 0, 1 }, { 0, 1, 1 } } };

// p1_div_k_grad_affine_q3::integrateAll (3D) (forms/form_hyteg_generated/p1/p1_div_k_grad_affine_q3.cpp:179-381):
// a_ij = sum_q w_q k(x_q) |det J| grad phi_i . grad phi_j with the 5-point degree-3 rule: the centroid with
// weight -2/15 and the four points with barycentric coordinates (1/2, 1/6, 1/6, 1/6) with weight 3/40.
inline double baryGrads( const P3* x, double G[4][3] );
inline void divKGradElement( const P3* x, const KField& k, double a[4][4] )
{
   double       G[4][3];
   const double det = baryGrads( x, G );
   const double lam[5][4] = { { 0.25, 0.25, 0.25, 0.25 },
                              { 1.0 / 6, 1.0 / 6, 1.0 / 6, 0.5 },
                              { 1.0 / 6, 1.0 / 6, 0.5, 1.0 / 6 },
                              { 1.0 / 6, 0.5, 1.0 / 6, 1.0 / 6 },
                              { 0.5, 1.0 / 6, 1.0 / 6, 1.0 / 6 } };
   const double wq[5]     = { -2.0 / 15.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0 };
   double       kbar      = 0.0;
   for ( int q = 0; q < 5; ++q )
   {
      P3 xq{ 0, 0, 0 };
      for ( int v = 0; v < 4; ++v )
         for ( int c = 0; c < 3; ++c )
            xq[c] += lam[q][v] * x[v][c];
      kbar += wq[q] * k( xq );
   }
   const double s = kbar * std::fabs( det );
   for ( int i = 0; i < 4; ++i )
      for ( int j = 0; j < 4; ++j )
         a[i][j] = s * ( G[i][0] * G[j][0] + G[i][1] * G[j][1] + G[i][2] * G[j][2] );
}

// gradients of the barycentric coordinates of tetrahedron x (rows of J^-1, J = [x1-x0 x2-x0 x3-x0]); returns det J
inline double baryGrads( const P3* x, double G[4][3] )
{
   double J[3][3];
   for ( int r = 0; r < 3; ++r )
      for ( int c = 0; c < 3; ++c )
         J[r][c] = x[c + 1][r] - x[0][r];
   const double det = J[0][0] * ( J[1][1] * J[2][2] - J[1][2] * J[2][1] ) - J[0][1] * ( J[1][0] * J[2][2] - J[1][2] * J[2][0] ) +
                      J[0][2] * ( J[1][0] * J[2][1] - J[1][1] * J[2][0] );
   // rows of J^-1 are the gradients of the barycentric coordinates 1..3
   double Ji[3][3];
   Ji[0][0] = ( J[1][1] * J[2][2] - J[1][2] * J[2][1] ) / det;
   Ji[0][1] = ( J[0][2] * J[2][1] - J[0][1] * J[2][2] ) / det;
   Ji[0][2] = ( J[0][1] * J[1][2] - J[0][2] * J[1][1] ) / det;
   Ji[1][0] = ( J[1][2] * J[2][0] - J[1][0] * J[2][2] ) / det;
   Ji[1][1] = ( J[0][0] * J[2][2] - J[0][2] * J[2][0] ) / det;
   Ji[1][2] = ( J[0][2] * J[1][0] - J[0][0] * J[1][2] ) / det;
   Ji[2][0] = ( J[1][0] * J[2][1] - J[1][1] * J[2][0] ) / det;
   Ji[2][1] = ( J[0][1] * J[2][0] - J[0][0] * J[2][1] ) / det;
   Ji[2][2] = ( J[0][0] * J[1][1] - J[0][1] * J[1][0] ) / det;
   for ( int c = 0; c < 3; ++c )
   {
      G[1][c] = Ji[0][c];
      G[2][c] = Ji[1][c];
      G[3][c] = Ji[2][c];
      G[0][c] = -( G[1][c] + G[2][c] + G[3][c] );
   }
   return det;
}

// P1ElementwiseOperator<p1_div_k_grad_affine_q3> (elementwiseoperators/P1ElementwiseOperator.cpp:64-220, 275-301,
// computeDiagonalOperatorValues :304-415): element loops over the micro-cells of every macro-cell accumulate into
// the cell arrays, whose boundary values are then summed into the lower-dimensional primitives.
template < typename T >
struct P1ElementwiseDivKGrad : OperatorBase< T >
{
   using OperatorBase< T >::m;
   using OperatorBase< T >::invDiag;
   KField k;

   // element matrix of a micro-cell of macro-cell c with (affine) vertices xc
   virtual void element( int c, const P3* xc, double el[4][4] ) const
   {
      (void) c;
      divKGradElement( xc, k, el );
   }

   P1ElementwiseDivKGrad( const Mesh* mesh, int lo, int hi, const KField& kf )
   : k( kf )
   {
      this->m    = mesh;
      this->minL = lo;
      this->maxL = hi;
      invDiag    = std::make_unique< Function< T > >( m, lo, hi );
      for ( int L = lo; L <= hi; ++L )
         computeInverseDiagonal( L );
   }

   // fn( cellIndex, vertex array indices, element matrix ) for every micro-cell of cell c
   template < typename Fn >
   void forElements( int c, int L, Fn&& fn ) const
   {
      const idx N = idx( 1 ) << L;
      const idx w = N + 1;
      for ( int t = 0; t < 6; ++t )
         for ( idx z = 0; z <= N; ++z )
            for ( idx y = 0; y + z <= N; ++y )
               for ( idx x = 0; x + y + z <= N; ++x )
               {
                  idx vi[4];
                  P3  xc[4];
                  bool inside = true;
                  for ( int v = 0; v < 4; ++v )
                  {
                     const idx a = x + MICRO_CELL[t][v][0], b = y + MICRO_CELL[t][v][1], d = z + MICRO_CELL[t][v][2];
                     if ( !inCell( N, a, b, d ) )
                     {
                        inside = false;
                        break;
                     }
                     vi[v] = cidx( w, a, b, d );
                     xc[v] = microCoord( *m, c, L, a, b, d );
                  }
                  if ( !inside )
                     continue;
                  double el[4][4];
                  element( c, xc, el );
                  fn( vi, el );
               }
   }

   void apply( const Function< T >& src, Function< T >& dst, int L, int flag, bool add ) const override
   {
      src.commDown( L );
      if ( !add )
         interpolate< T >( dst, 0.0, L, flag );
      const idx N = idx( 1 ) << L;
      const idx w = N + 1;
#pragma omp parallel for schedule( dynamic )
      for ( idx c = 0; c < (idx) m->Cl.size(); ++c )
      {
         auto&       dc = dst.c[L][c];
         const auto& sc = src.c[L][c];
         // zero the cell's copies of face, edge and vertex values only
         for ( idx z = 0; z <= N; ++z )
            for ( idx y = 0; y + z <= N; ++y )
               for ( idx x = 0; x + y + z <= N; ++x )
                  if ( x == 0 || y == 0 || z == 0 || x + y + z == N )
                     dc[cidx( w, x, y, z )] = T( 0 );
         forElements( (int) c, L, [&]( const idx* vi, const double( *el )[4] ) {
            for ( int i = 0; i < 4; ++i )
            {
               double s = 0.0;
               for ( int j = 0; j < 4; ++j )
                  s += el[i][j] * double( sc[vi[j]] );
               dc[vi[i]] += T( s );
            }
         } );
      }
      commAdditiveCellToLower( dst, L, All ^ flag, !add );
   }

   void computeInverseDiagonal( int L )
   {
      auto&     D = *invDiag;
      const idx N = idx( 1 ) << L;
      interpolate< T >( D, 0.0, L, All );
      for ( idx c = 0; c < (idx) m->Cl.size(); ++c )
      {
         auto& dc = D.c[L][c];
         std::fill( dc.begin(), dc.end(), T( 0 ) );
         forElements( (int) c, L, [&]( const idx* vi, const double( *el )[4] ) {
            for ( int i = 0; i < 4; ++i )
               dc[vi[i]] += T( el[i][i] );
         } );
      }
      commAdditiveCellToLower( D, L, None, true );
      (void) N;
      forOwned< T >( *m, L, All, [&]( int kind, int p, idx i ) {
         T& v = at( D, L, kind, p, i );
         v    = T( 1.0 ) / v;
      } );
   }
};

// P1ElementwiseBlendingLaplaceOperator = P1ElementwiseOperator< p1_diffusion_blending_q3 >
// (elementwiseoperators/P1ElementwiseOperator.hpp:332) on a mesh with the spherical-shell map: the element loops of
// the elementwise operator with blendingLaplaceElement and each macro-cell's map.
template < typename T >
struct P1ElementwiseBlendingLaplace : P1ElementwiseDivKGrad< T >
{
   P1ElementwiseBlendingLaplace( const Mesh* mesh, int lo, int hi )
   : P1ElementwiseDivKGrad< T >( mesh, lo, hi, KField{} )
   {
      if ( mesh->shell.size() != mesh->Cl.size() )
         throw std::invalid_argument( "the blended operator needs the shell map of every cell" );
      for ( int L = lo; L <= hi; ++L )
         this->computeInverseDiagonal( L ); // the base constructor used its own element matrices
   }
   void element( int c, const P3* xc, double el[4][4] ) const override { blendingLaplaceElement( xc, this->m->shell[c], el ); }
};

// Least squares by Householder QR with column pivoting (the reference solves its fits with Eigen's
// colPivHouseholderQr().solve, polynomial/LSQPInterpolator.hpp:166-180): a is m x n row-major; every right-hand
// side in rhs (length m) is replaced by its solution (length n). Pivot: the remaining column of largest norm (first
// of equals); the numerical rank counts |R_kk| > eps min(m, n) |R_00|; columns beyond it get zero (the basic
// solution of a rank-deficient system).
inline void lsqColPiv( std::vector< double > a, int m, int n, std::vector< std::vector< double > >& rhs )
{
   std::vector< int >    perm( n );
   std::vector< double > tau( n, 0.0 );
   for ( int j = 0; j < n; ++j )
      perm[j] = j;
   const int kmax = std::min( m, n );
   int       rank = 0;
   double    r00  = 0.0;
   for ( int k = 0; k < kmax; ++k )
   {
      int    jb = k;
      double nb = -1.0;
      for ( int j = k; j < n; ++j )
      {
         double s = 0.0;
         for ( int i = k; i < m; ++i )
            s += a[size_t( i ) * n + j] * a[size_t( i ) * n + j];
         if ( s > nb )
         {
            nb = s;
            jb = j;
         }
      }
      if ( jb != k )
      {
         for ( int i = 0; i < m; ++i )
            std::swap( a[size_t( i ) * n + k], a[size_t( i ) * n + jb] );
         std::swap( perm[k], perm[jb] );
      }
      const double nx = std::sqrt( nb );
      if ( k == 0 )
         r00 = nx;
      if ( nx <= std::numeric_limits< double >::epsilon() * kmax * r00 || nx == 0.0 )
         break;
      ++rank;
      const double alpha = a[size_t( k ) * n + k];
      const double beta  = alpha >= 0.0 ? -nx : nx;
      tau[k]             = ( beta - alpha ) / beta;
      const double sc    = 1.0 / ( alpha - beta );
      for ( int i = k + 1; i < m; ++i )
         a[size_t( i ) * n + k] *= sc;
      a[size_t( k ) * n + k] = beta;
      for ( int j = k + 1; j < n; ++j )
      {
         double wv = a[size_t( k ) * n + j];
         for ( int i = k + 1; i < m; ++i )
            wv += a[size_t( i ) * n + k] * a[size_t( i ) * n + j];
         wv *= tau[k];
         a[size_t( k ) * n + j] -= wv;
         for ( int i = k + 1; i < m; ++i )
            a[size_t( i ) * n + j] -= wv * a[size_t( i ) * n + k];
      }
   }
   for ( auto& b : rhs )
   {
      for ( int k = 0; k < rank; ++k )
      {
         double wv = b[k];
         for ( int i = k + 1; i < m; ++i )
            wv += a[size_t( i ) * n + k] * b[i];
         wv *= tau[k];
         b[k] -= wv;
         for ( int i = k + 1; i < m; ++i )
            b[i] -= wv * a[size_t( i ) * n + k];
      }
      std::vector< double > z( n, 0.0 );
      for ( int k = rank - 1; k >= 0; --k )
      {
         double v = b[k];
         for ( int j = k + 1; j < rank; ++j )
            v -= a[size_t( k ) * n + j] * z[j];
         z[k] = v / a[size_t( k ) * n + k];
      }
      b.assign( n, 0.0 );
      for ( int j = 0; j < n; ++j )
         b[perm[j]] = z[j];
   }
}

// P1SurrogateOperator< p1_div_k_grad_affine_q3 > (p1functionspace/P1SurrogateOperator.hpp): after
// interpolateStencils( degree, interpolationLevel ) (:232-335) every stencil direction of the interior points of a
// macro-cell is a polynomial of the given degree in the cell-local coordinates (i, j, k) / N, fitted by least
// squares to the variable stencils (P1Elements3D::assembleP1LocalStencilNew_new, P1Operator.hpp:2134-2161) at the
// interior points of the interpolation level (the level's own points scaled by 2^(L - interpolation level)).
// Applies, the inverse diagonal (P1Operator.hpp:645-915 with assemble_stencil_cell) and everything built on them use
// the polynomials at the interior points and the variable stencils at face, edge and vertex points (:404-469).
template < typename T >
struct P1SurrogateDivKGrad : P1ElementwiseDivKGrad< T >
{
   using P1ElementwiseDivKGrad< T >::P1ElementwiseDivKGrad;
   using P1ElementwiseDivKGrad< T >::k;
   int degree = -1;
   // [level][cell]: NDIR x monomialCount3(degree) coefficients, direction-major
   std::vector< std::vector< std::vector< double > > > poly;

   // variable stencil of lattice point (x, y, z) of cell c: rows of the element matrices of the incident micro-cells
   void variableStencil( int c, int L, idx x, idx y, idx z, double st[NDIR] ) const
   {
      const idx N = idx( 1 ) << L;
      for ( int d = 0; d < NDIR; ++d )
         st[d] = 0.0;
      for ( int t = 0; t < 6; ++t )
         for ( int v = 0; v < 4; ++v )
         {
            const idx ax = x - MICRO_CELL[t][v][0], ay = y - MICRO_CELL[t][v][1], az = z - MICRO_CELL[t][v][2];
            P3        xc[4];
            bool      inside = true;
            for ( int u = 0; u < 4; ++u )
            {
               const idx a = ax + MICRO_CELL[t][u][0], b = ay + MICRO_CELL[t][u][1], e = az + MICRO_CELL[t][u][2];
               if ( !inCell( N, a, b, e ) )
               {
                  inside = false;
                  break;
               }
               xc[u] = microCoord( *this->m, c, L, a, b, e );
            }
            if ( !inside )
               continue;
            double el[4][4];
            divKGradElement( xc, k, el );
            for ( int u = 0; u < 4; ++u )
            {
               const int o[3] = { MICRO_CELL[t][u][0] - MICRO_CELL[t][v][0], MICRO_CELL[t][u][1] - MICRO_CELL[t][v][1],
                                  MICRO_CELL[t][u][2] - MICRO_CELL[t][v][2] };
               for ( int d = 0; d < NDIR; ++d )
                  if ( DOFF[d][0] == o[0] && DOFF[d][1] == o[1] && DOFF[d][2] == o[2] )
                     st[d] += el[v][u];
            }
         }
   }

   void interpolateStencils( int deg, int interpolationLevel )
   {
      degree      = deg;
      const int nc = monomialCount3( deg );
      poly.assign( this->maxL + 1, {} );
      for ( int L = std::max( this->minL, 1 ); L <= this->maxL; ++L )
      {
         const int    Li = std::min( L, interpolationLevel ), sh = L - Li;
         const idx    Ni = idx( 1 ) << Li;
         const double hi = 1.0 / double( Ni );
         std::vector< std::array< idx, 3 > > pts;
         for ( idx z = 1; z + 3 <= Ni; ++z )
            for ( idx y = 1; y + z + 2 <= Ni; ++y )
               for ( idx x = 1; x + y + z + 1 <= Ni; ++x )
                  pts.push_back( { x, y, z } );
         const int             np = int( pts.size() );
         std::vector< double > a( size_t( np ) * nc );
         for ( int r = 0; r < np; ++r )
            for ( int i = 0; i < nc; ++i )
               a[size_t( r ) * nc + i] = monomial3( i, P3{ pts[r][0] * hi, pts[r][1] * hi, pts[r][2] * hi } );
         poly[L].assign( this->m->Cl.size(), std::vector< double >( size_t( NDIR ) * nc, 0.0 ) );
         for ( int c = 0; c < (int) this->m->Cl.size(); ++c )
         {
            std::vector< std::vector< double > > rhs( NDIR, std::vector< double >( np ) );
            for ( int r = 0; r < np; ++r )
            {
               double st[NDIR];
               variableStencil( c, L, pts[r][0] << sh, pts[r][1] << sh, pts[r][2] << sh, st );
               for ( int d = 0; d < NDIR; ++d )
                  rhs[d][r] = st[d];
            }
            if ( np > 0 )
               lsqColPiv( a, np, nc, rhs );
            for ( int d = 0; d < NDIR; ++d )
               for ( int i = 0; i < nc && np > 0; ++i )
                  poly[L][c][size_t( d ) * nc + i] = rhs[d][i];
         }
         surrogateInverseDiagonal( L );
      }
   }

   double evalDir( int L, int c, int d, idx x, idx y, idx z ) const
   {
      const int    nc = monomialCount3( degree );
      const double h  = 1.0 / double( idx( 1 ) << L );
      const P3     p{ x * h, y * h, z * h };
      double       v = 0.0;
      for ( int i = 0; i < nc; ++i )
         v += poly[L][c][size_t( d ) * nc + i] * monomial3( i, p );
      return v;
   }
   bool active( int L ) const { return degree >= 0 && L < (int) poly.size() && !poly[L].empty(); }

   void apply( const Function< T >& src, Function< T >& dst, int L, int flag, bool add ) const override
   {
      if ( !active( L ) )
      {
         P1ElementwiseDivKGrad< T >::apply( src, dst, L, flag, add );
         return;
      }
      const idx N = idx( 1 ) << L, w = N + 1;
      // interior values before the apply (the element loops add into them)
      std::vector< std::vector< T > > keep;
      if ( add && ( flag & Inner ) )
         for ( const auto& dc : dst.c[L] )
            keep.push_back( dc );
      P1ElementwiseDivKGrad< T >::apply( src, dst, L, flag, add );
      if ( !( flag & Inner ) )
         return;
      src.commDown( L );
#pragma omp parallel for schedule( dynamic )
      for ( idx c = 0; c < (idx) this->m->Cl.size(); ++c )
      {
         const auto& sc = src.c[L][c];
         auto&       dc = dst.c[L][c];
         for ( idx z = 1; z + 3 <= N; ++z )
            for ( idx y = 1; y + z + 2 <= N; ++y )
               for ( idx x = 1; x + y + z + 1 <= N; ++x )
               {
                  double s = 0.0;
                  for ( int d = 0; d < NDIR; ++d )
                     s += evalDir( L, (int) c, d, x, y, z ) * double( sc[cidx( w, x + DOFF[d][0], y + DOFF[d][1], z + DOFF[d][2] )] );
                  const idx q = cidx( w, x, y, z );
                  dc[q]       = ( add ? keep[c][q] : T( 0 ) ) + T( s );
               }
      }
   }

   // interior points: 1 / the polynomial centre weight; elsewhere the variable stencil's
   void surrogateInverseDiagonal( int L )
   {
      auto&     D = *this->invDiag;
      const idx N = idx( 1 ) << L, w = N + 1;
      for ( idx c = 0; c < (idx) this->m->Cl.size(); ++c )
         for ( idx z = 1; z + 3 <= N; ++z )
            for ( idx y = 1; y + z + 2 <= N; ++y )
               for ( idx x = 1; x + y + z + 1 <= N; ++x )
                  D.c[L][c][cidx( w, x, y, z )] = T( 1.0 / evalDir( L, (int) c, V_C, x, y, z ) );
   }
};

// P1toP1LinearRestriction::restrict3D with the generated kernel's per-leaf scaling
template < typename T >
void restrict3D( Function< T >& fn, int srcL, int flag )
{
   const Mesh& m  = *fn.m;
   const int   dL = srcL - 1;
   fn.commDown( srcL );
   const idx Nc = idx( 1 ) << dL, wc = Nc + 1;
   const idx Nf = idx( 1 ) << srcL, wf = Nf + 1;
#pragma omp parallel for schedule( dynamic )
   for ( idx ci = 0; ci < (idx) m.Cl.size(); ++ci )
   {
      const auto& src = fn.c[srcL][ci];
      auto&       dst = fn.c[dL][ci];
      for ( idx z = 0; z <= Nc; ++z )
         for ( idx y = 0; y + z <= Nc; ++y )
            for ( idx x = 0; x + y + z <= Nc; ++x )
            {
               T   r  = T( 0 );
               idx fx = 2 * x, fy = 2 * y, fz = 2 * z;
               for ( int d = 0; d < NDIR; ++d )
               {
                  idx lx = fx + DOFF[d][0], ly = fy + DOFF[d][1], lz = fz + DOFF[d][2];
                  if ( !inCell( Nf, lx, ly, lz ) )
                     continue;
                  double wgt = ( d == V_C ? 1.0 : 0.5 ) * invNumCells( m, (int) ci, classify( Nf, lx, ly, lz ) );
                  r += T( wgt ) * src[cidx( wf, lx, ly, lz )];
               }
               dst[cidx( wc, x, y, z )] = r;
            }
   }
   commAdditiveCellToLower( fn, dL, flag ^ All, true );
}

// P1toP1LinearProlongation::prolongate3DAdditively
template < typename T >
void prolongate3D( Function< T >& fn, int srcL, int flag, bool add )
{
   const Mesh& m  = *fn.m;
   const int   dL = srcL + 1;
   fn.commDown( srcL );
   const idx Nc = idx( 1 ) << srcL, wc = Nc + 1;
   const idx Nf = idx( 1 ) << dL, wf = Nf + 1;
#pragma omp parallel for schedule( dynamic )
   for ( idx ci = 0; ci < (idx) m.Cl.size(); ++ci )
   {
      const auto& src = fn.c[srcL][ci];
      auto&       dst = fn.c[dL][ci];
      for ( idx z = 0; z <= Nf; ++z )
         for ( idx y = 0; y + z <= Nf; ++y )
            for ( idx x = 0; x + y + z <= Nf; ++x )
            {
               bool onFace = x == 0 || y == 0 || z == 0 || x + y + z == Nf;
               if ( !add || onFace )
                  dst[cidx( wf, x, y, z )] = T( 0 );
            }
      for ( idx z = 0; z <= Nc; ++z )
         for ( idx y = 0; y + z <= Nc; ++y )
            for ( idx x = 0; x + y + z <= Nc; ++x )
            {
               const T   v  = src[cidx( wc, x, y, z )];
               const idx fx = 2 * x, fy = 2 * y, fz = 2 * z;
               for ( int d = 0; d < NDIR; ++d )
               {
                  idx lx = fx + DOFF[d][0], ly = fy + DOFF[d][1], lz = fz + DOFF[d][2];
                  if ( !inCell( Nf, lx, ly, lz ) )
                     continue;
                  double wgt = ( d == V_C ? 1.0 : 0.5 ) * invNumCells( m, (int) ci, classify( Nf, lx, ly, lz ) );
                  dst[cidx( wf, lx, ly, lz )] += T( wgt ) * v;
               }
            }
   }
   commAdditiveCellToLower( fn, dL, flag ^ All, !add );
}

template < typename T >
void OperatorBase< T >::restrictOp( Function< T >& f, int srcL, int flag ) const
{
   restrict3D< T >( f, srcL, flag );
}
template < typename T >
void OperatorBase< T >::prolongateOp( Function< T >& f, int srcL, int flag, bool add ) const
{
   prolongate3D< T >( f, srcL, flag, add );
}

// ---------------------------------------------------------------- P2 (quadratic Lagrange) path
// A P2Function of level L (p2functionspace/P2Function.hpp) holds vertex DoFs at the vertices of the level-L
// micro-mesh and edge DoFs at the midpoints of its micro-edges (edgedofspace/EdgeDoFIndexing.hpp:919-1000, the
// seven orientations X Y Z XY XZ YZ XYZ). Together these points are exactly the vertices of level L+1, so the
// oracle stores a P2 function of level L as a vertex-DoF function of "lattice" level L+1 (vertex DoF (i,j,k) at
// lattice point (2i,2j,2k), the edge DoF between vertices a and b at a+b). Every P2 routine below takes lattice
// levels and works on the micro-cells of level lattice-1.

// local nodes of a P2 tetrahedron: vertices 0..3, then the edges (2,3) (1,3) (1,2) (0,3) (0,2) (0,1)
// (the FEniCS P2 tetrahedron of form_fenics_generated/p2_tet_diffusion.h)
static const int P2_EDGE[6][2] = { { 2, 3 }, { 1, 3 }, { 1, 2 }, { 0, 3 }, { 0, 2 }, { 0, 1 } };

inline void p2Basis( const double lam[4], double phi[10] )
{
   for ( int v = 0; v < 4; ++v )
      phi[v] = lam[v] * ( 2.0 * lam[v] - 1.0 );
   for ( int e = 0; e < 6; ++e )
      phi[4 + e] = 4.0 * lam[P2_EDGE[e][0]] * lam[P2_EDGE[e][1]];
}

inline void p2Grads( const double lam[4], const double G[4][3], double d[10][3] )
{
   for ( int v = 0; v < 4; ++v )
      for ( int c = 0; c < 3; ++c )
         d[v][c] = ( 4.0 * lam[v] - 1.0 ) * G[v][c];
   for ( int e = 0; e < 6; ++e )
   {
      const int a = P2_EDGE[e][0], b = P2_EDGE[e][1];
      for ( int c = 0; c < 3; ++c )
         d[4 + e][c] = 4.0 * ( lam[a] * G[b][c] + lam[b] * G[a][c] );
   }
}

// p2_tet_diffusion tabulate_tensor: a_ij = integral over the tetrahedron of grad phi_i . grad phi_j. The
// integrand is quadratic, so the 4-point rule with barycentric coordinates (a, b, b, b), a = (5 + 3 sqrt 5) / 20,
// b = (5 - sqrt 5) / 20, and weights vol / 4 is exact.
inline void p2LaplaceElement( const P3* x, double a[10][10] )
{
   double       G[4][3];
   const double vol = std::fabs( baryGrads( x, G ) ) / 6.0;
   const double qa = ( 5.0 + 3.0 * std::sqrt( 5.0 ) ) / 20.0, qb = ( 5.0 - std::sqrt( 5.0 ) ) / 20.0;
   for ( int i = 0; i < 10; ++i )
      for ( int j = 0; j < 10; ++j )
         a[i][j] = 0.0;
   for ( int q = 0; q < 4; ++q )
   {
      double lam[4] = { qb, qb, qb, qb };
      lam[q]        = qa;
      double d[10][3];
      p2Grads( lam, G, d );
      for ( int i = 0; i < 10; ++i )
         for ( int j = 0; j < 10; ++j )
            a[i][j] += 0.25 * vol * ( d[i][0] * d[j][0] + d[i][1] * d[j][1] + d[i][2] * d[j][2] );
   }
}

// lattice coordinates (level lat) of the 10 nodes of micro-cell (t, anchor) of level lat-1
inline void p2Nodes( int t, idx ax, idx ay, idx az, idx out[10][3] )
{
   idx v[4][3];
   for ( int k = 0; k < 4; ++k )
   {
      v[k][0] = ax + MICRO_CELL[t][k][0];
      v[k][1] = ay + MICRO_CELL[t][k][1];
      v[k][2] = az + MICRO_CELL[t][k][2];
      for ( int c = 0; c < 3; ++c )
         out[k][c] = 2 * v[k][c];
   }
   for ( int e = 0; e < 6; ++e )
      for ( int c = 0; c < 3; ++c )
         out[4 + e][c] = v[P2_EDGE[e][0]][c] + v[P2_EDGE[e][1]][c];
}

// fn( t, node lattice coordinates ) for every micro-cell of level lat-1 of a macro-cell
template < typename Fn >
void forP2Cells( int lat, Fn&& fn )
{
   const idx Nc = idx( 1 ) << ( lat - 1 );
   for ( int t = 0; t < 6; ++t )
      for ( idx z = 0; z <= Nc; ++z )
         for ( idx y = 0; y + z <= Nc; ++y )
            for ( idx x = 0; x + y + z <= Nc; ++x )
            {
               bool inside = true;
               for ( int v = 0; v < 4 && inside; ++v )
                  inside = inCell( Nc, x + MICRO_CELL[t][v][0], y + MICRO_CELL[t][v][1], z + MICRO_CELL[t][v][2] );
               if ( !inside )
                  continue;
               idx nd[10][3];
               p2Nodes( t, x, y, z, nd );
               fn( t, nd );
            }
}

// P2ConstantOperator< p2_tet_diffusion > (p2functionspace/P2ConstantOperator.cpp:99-113 with the stencils of
// P2Elements3D, i.e. the sums of element-matrix entries over the micro-cells at a DoF), restated element by
// element: for every micro-cell, y_T += A_t x_T with the element matrix of its type in its macro-cell
// (P2ElementwiseOperator::apply, elementwiseoperators/P2ElementwiseOperator.cpp:100-310, performs the same sums),
// followed by the additive Cell -> Face/Edge/Vertex reduction. Levels are lattice levels.
template < typename T >
struct P2ConstantLaplace : OperatorBase< T >
{
   using OperatorBase< T >::m;
   using OperatorBase< T >::invDiag;
   // [lat][6 * cell + type] element matrices of the micro-cells of level lat-1
   std::vector< std::vector< std::array< std::array< double, 10 >, 10 > > > el;

   P2ConstantLaplace( const Mesh* mesh, int lo, int hi ) // P2 levels
   {
      this->m    = mesh;
      this->minL = lo + 1;
      this->maxL = hi + 1;
      el.resize( hi + 2 );
      for ( int lat = lo + 1; lat <= hi + 1; ++lat )
      {
         el[lat].resize( 6 * m->Cl.size() );
         for ( size_t c = 0; c < m->Cl.size(); ++c )
            for ( int t = 0; t < 6; ++t )
            {
               P3 xc[4];
               for ( int v = 0; v < 4; ++v )
                  xc[v] = microCoord( *m, (int) c, lat - 1, MICRO_CELL[t][v][0], MICRO_CELL[t][v][1], MICRO_CELL[t][v][2] );
               double a[10][10];
               p2LaplaceElement( xc, a );
               for ( int i = 0; i < 10; ++i )
                  for ( int j = 0; j < 10; ++j )
                     el[lat][6 * c + t][i][j] = a[i][j];
            }
      }
      invDiag = std::make_unique< Function< T > >( m, this->minL, this->maxL );
      for ( int lat = this->minL; lat <= this->maxL; ++lat )
         computeInverseDiagonal( lat );
   }

   void apply( const Function< T >& src, Function< T >& dst, int lat, int flag, bool add ) const override
   {
      src.commDown( lat );
      if ( !add )
         interpolate< T >( dst, 0.0, lat, flag );
      const idx N = idx( 1 ) << lat;
      const idx w = N + 1;
#pragma omp parallel for schedule( dynamic )
      for ( idx c = 0; c < (idx) m->Cl.size(); ++c )
      {
         auto&       dc = dst.c[lat][c];
         const auto& sc = src.c[lat][c];
         for ( idx z = 0; z <= N; ++z )
            for ( idx y = 0; y + z <= N; ++y )
               for ( idx x = 0; x + y + z <= N; ++x )
                  if ( x == 0 || y == 0 || z == 0 || x + y + z == N )
                     dc[cidx( w, x, y, z )] = T( 0 );
         forP2Cells( lat, [&]( int t, const idx( *nd )[3] ) {
            const auto& A = el[lat][6 * c + t];
            idx         vi[10];
            for ( int i = 0; i < 10; ++i )
               vi[i] = cidx( w, nd[i][0], nd[i][1], nd[i][2] );
            for ( int i = 0; i < 10; ++i )
            {
               double s = 0.0;
               for ( int j = 0; j < 10; ++j )
                  s += A[i][j] * double( sc[vi[j]] );
               dc[vi[i]] += T( s );
            }
         } );
      }
      commAdditiveCellToLower( dst, lat, All ^ flag, !add );
   }

   // P2ConstantOperator::computeInverseDiagonalOperatorValues: sums of the element diagonals, inverted
   void computeInverseDiagonal( int lat )
   {
      auto&     D = *invDiag;
      const idx w = ( idx( 1 ) << lat ) + 1;
      interpolate< T >( D, 0.0, lat, All );
      for ( idx c = 0; c < (idx) m->Cl.size(); ++c )
      {
         auto& dc = D.c[lat][c];
         std::fill( dc.begin(), dc.end(), T( 0 ) );
         forP2Cells( lat, [&]( int t, const idx( *nd )[3] ) {
            const auto& A = el[lat][6 * c + t];
            for ( int i = 0; i < 10; ++i )
               dc[cidx( w, nd[i][0], nd[i][1], nd[i][2] )] += T( A[i][i] );
         } );
      }
      commAdditiveCellToLower( D, lat, None, true );
      forOwned< T >( *m, lat, All, [&]( int kind, int p, idx i ) {
         T& v = at( D, lat, kind, p, i );
         v    = T( 1.0 ) / v;
      } );
   }

   void restrictOp( Function< T >& f, int srcL, int flag ) const override;
   void prolongateOp( Function< T >& f, int srcL, int flag, bool add ) const override;
};

// A micro-cell of level lat-2 containing lattice point p of level lat (the nodes of the level-(lat-1) lattice
// are those of the P2 function of level lat-2): type, anchor and 4x the barycentric coordinates of p.
inline void p2Locate( idx Nc, const idx p[3], int& t, idx a[3], idx lam4[4] )
{
   t = -1;
   for ( int c = 0; c < 3; ++c )
      a[c] = 0;
   for ( int k = 0; k < 4; ++k )
      lam4[k] = 0;
   const idx q[3] = { p[0] / 4, p[1] / 4, p[2] / 4 };
   for ( int tt = 0; tt < 6; ++tt )
      for ( int o = 0; o < 8; ++o )
      {
         const idx an[3] = { q[0] - ( o & 1 ), q[1] - ( ( o >> 1 ) & 1 ), q[2] - ( ( o >> 2 ) & 1 ) };
         idx       v[4][3];
         bool      inside = true;
         for ( int k = 0; k < 4 && inside; ++k )
         {
            for ( int c = 0; c < 3; ++c )
               v[k][c] = an[c] + MICRO_CELL[tt][k][c];
            inside = inCell( Nc, v[k][0], v[k][1], v[k][2] );
         }
         if ( !inside )
            continue;
         // 4 lambda_1..3 = J^-1 (p - 4 v0), J = [v1-v0 v2-v0 v3-v0] (integer, det +-1)
         long J[3][3];
         for ( int r = 0; r < 3; ++r )
            for ( int c = 0; c < 3; ++c )
               J[r][c] = long( v[c + 1][r] - v[0][r] );
         const long det = J[0][0] * ( J[1][1] * J[2][2] - J[1][2] * J[2][1] ) - J[0][1] * ( J[1][0] * J[2][2] - J[1][2] * J[2][0] ) +
                          J[0][2] * ( J[1][0] * J[2][1] - J[1][1] * J[2][0] );
         const long Ji[3][3] = { { J[1][1] * J[2][2] - J[1][2] * J[2][1], J[0][2] * J[2][1] - J[0][1] * J[2][2],
                                   J[0][1] * J[1][2] - J[0][2] * J[1][1] },
                                 { J[1][2] * J[2][0] - J[1][0] * J[2][2], J[0][0] * J[2][2] - J[0][2] * J[2][0],
                                   J[0][2] * J[1][0] - J[0][0] * J[1][2] },
                                 { J[1][0] * J[2][1] - J[1][1] * J[2][0], J[0][1] * J[2][0] - J[0][0] * J[2][1],
                                   J[0][0] * J[1][1] - J[0][1] * J[1][0] } };
         const long d[3] = { long( p[0] - 4 * v[0][0] ), long( p[1] - 4 * v[0][1] ), long( p[2] - 4 * v[0][2] ) };
         long       l[4];
         l[0] = 4;
         bool ok = true;
         for ( int r = 0; r < 3; ++r )
         {
            l[r + 1] = ( Ji[r][0] * d[0] + Ji[r][1] * d[1] + Ji[r][2] * d[2] ) / det;
            ok       = ok && l[r + 1] >= 0;
            l[0] -= l[r + 1];
         }
         if ( !ok || l[0] < 0 )
            continue;
         t = tt;
         for ( int c = 0; c < 3; ++c )
            a[c] = an[c];
         for ( int k = 0; k < 4; ++k )
            lam4[k] = l[k];
         return;
      }
   throw std::runtime_error( "p2Locate: lattice point outside the macro-cell" );
}

// P2toP2QuadraticProlongation::prolongateAdditively3D (gridtransferoperators/P2toP2QuadraticProlongation.cpp:217-424):
// every DoF of level src+1 gets the value of the quadratic interpolant of the level-src function at its point.
// The reference pushes the coarse DoFs with the weights phi_j(p) (3/8, -1/8, 3/4, 1/4, ...) cell by cell, scaling
// points on macro-faces/-edges/-vertices by 1 / (number of cells sharing them), then sums additively; the oracle
// pulls the same weighted sum per fine point. srcLat is the lattice level of the coarse function.
template < typename T >
void prolongateP2( Function< T >& fn, int srcLat, int flag, bool add )
{
   const Mesh& m  = *fn.m;
   const int   dL = srcLat + 1;
   fn.commDown( srcLat );
   const idx Nf = idx( 1 ) << dL, wf = Nf + 1;
   const idx Nc = idx( 1 ) << ( srcLat - 1 ), wc = ( idx( 1 ) << srcLat ) + 1;
#pragma omp parallel for schedule( dynamic )
   for ( idx ci = 0; ci < (idx) m.Cl.size(); ++ci )
   {
      const auto& src = fn.c[srcLat][ci];
      auto&       dst = fn.c[dL][ci];
      for ( idx z = 0; z <= Nf; ++z )
         for ( idx y = 0; y + z <= Nf; ++y )
            for ( idx x = 0; x + y + z <= Nf; ++x )
            {
               const bool onFace = x == 0 || y == 0 || z == 0 || x + y + z == Nf;
               T&         d      = dst[cidx( wf, x, y, z )];
               if ( !add || onFace )
                  d = T( 0 );
               const idx p[3] = { x, y, z };
               int       t;
               idx       a[3], l4[4];
               p2Locate( Nc, p, t, a, l4 );
               idx nd[10][3];
               p2Nodes( t, a[0], a[1], a[2], nd );
               const double lam[4] = { l4[0] / 4.0, l4[1] / 4.0, l4[2] / 4.0, l4[3] / 4.0 };
               double       phi[10];
               p2Basis( lam, phi );
               double v = 0.0;
               for ( int j = 0; j < 10; ++j )
                  if ( phi[j] != 0.0 )
                     v += phi[j] * double( src[cidx( wc, nd[j][0], nd[j][1], nd[j][2] )] );
               d += T( invNumCells( m, (int) ci, classify( Nf, x, y, z ) ) * v );
            }
   }
   commAdditiveCellToLower( fn, dL, flag ^ All, !add );
}

// P2toP2QuadraticRestriction::restrictAdditively3D (P2toP2QuadraticRestriction.cpp:131-286): the transpose of
// the prolongation, coarse_j = sum over the fine points p of phi_j(p) fine_p (fine points on macro-faces/-edges/
// -vertices scaled by 1 / (number of cells sharing them) per cell), then the additive Cell -> lower reduction.
template < typename T >
void restrictP2( Function< T >& fn, int srcLat, int flag )
{
   const Mesh& m  = *fn.m;
   const int   dL = srcLat - 1;
   fn.commDown( srcLat );
   const idx Nf = idx( 1 ) << srcLat, wf = Nf + 1;
   const idx Nc = idx( 1 ) << ( dL - 1 ), wc = ( idx( 1 ) << dL ) + 1;
#pragma omp parallel for schedule( dynamic )
   for ( idx ci = 0; ci < (idx) m.Cl.size(); ++ci )
   {
      const auto&           src = fn.c[srcLat][ci];
      auto&                 dst = fn.c[dL][ci];
      std::vector< double > acc( dst.size(), 0.0 );
      for ( idx z = 0; z <= Nf; ++z )
         for ( idx y = 0; y + z <= Nf; ++y )
            for ( idx x = 0; x + y + z <= Nf; ++x )
            {
               const idx p[3] = { x, y, z };
               int       t;
               idx       a[3], l4[4];
               p2Locate( Nc, p, t, a, l4 );
               idx nd[10][3];
               p2Nodes( t, a[0], a[1], a[2], nd );
               const double lam[4] = { l4[0] / 4.0, l4[1] / 4.0, l4[2] / 4.0, l4[3] / 4.0 };
               double       phi[10];
               p2Basis( lam, phi );
               const double s = invNumCells( m, (int) ci, classify( Nf, x, y, z ) ) * double( src[cidx( wf, x, y, z )] );
               for ( int j = 0; j < 10; ++j )
                  if ( phi[j] != 0.0 )
                     acc[cidx( wc, nd[j][0], nd[j][1], nd[j][2] )] += phi[j] * s;
            }
      for ( size_t i = 0; i < dst.size(); ++i )
         dst[i] = T( acc[i] );
   }
   commAdditiveCellToLower( fn, dL, flag ^ All, true );
}

template < typename T >
void P2ConstantLaplace< T >::restrictOp( Function< T >& f, int srcL, int flag ) const
{
   restrictP2< T >( f, srcL, flag );
}
template < typename T >
void P2ConstantLaplace< T >::prolongateOp( Function< T >& f, int srcL, int flag, bool add ) const
{
   prolongateP2< T >( f, srcL, flag, add );
}

// ---------------------------------------------------------------- P2 variable coefficient
// Coefficient of -div(k grad u) for the P2 elementwise operator. kind 0: k(x) = k0 exp(alpha (1 - |x|)) (1 + beta xi_e)
// with xi_e in [-1, 1) drawn per micro-element by SplitMix64 from (seed, element id); kind 1: the polynomial
// k(x) = 1 + x/2 + y^2/4 (for exactness checks of the quadrature).
struct P2KField
{
   int      kind  = 0;
   double   k0    = 1.0, alpha = 0.0, beta = 0.0;
   uint64_t seed  = 0;
   double   operator()( const P3& x, uint64_t eid ) const
   {
      if ( kind == 1 )
         return 1.0 + 0.5 * x[0] + 0.25 * x[1] * x[1];
      uint64_t z = seed + ( eid + 1 ) * 0x9E3779B97F4A7C15ull;
      z          = ( z ^ ( z >> 30 ) ) * 0xBF58476D1CE4E5B9ull;
      z          = ( z ^ ( z >> 27 ) ) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      const double xi = 2.0 * ( double( z >> 11 ) * 0x1.0p-53 ) - 1.0;
      return k0 * std::exp( alpha * ( 1.0 - std::sqrt( x[0] * x[0] + x[1] * x[1] + x[2] * x[2] ) ) ) * ( 1.0 + beta * xi );
   }
};

// Xiao-Gimbutas degree-4 rule with 11 points (barycentric coordinates w.r.t. vertices 0..3, weights summing to 1/6),
// the rule of forms/form_hyteg_generated/p2/p2_div_k_grad_affine_q4.cpp (integrateAll, quadrature points and the
// weight factors of |det J|).
static const double XG4_LAMBDA[11][4] = {
    { 0.17469405869723054, 0.040490506727590428, 0.01356070187980288, 0.77125473269537614 },
    { 0.081404918402859225, 0.75250850700965499, 0.068099370938206658, 0.097987203649279112 },
    { 0.30444840243449711, 0.24027666492807259, 0.12680172591539199, 0.32847320672203839 },
    { 0.74122888209362259, 0.067223294893383398, 0.035183929773598722, 0.15636389323939531 },
    { 0.053341239535745266, 0.41926631387951302, 0.047781435559086663, 0.47961101102565512 },
    { 0.43295349048135578, 0.45076587609127677, 0.059456616299433829, 0.056824017127933668 },
    { 0.53800720391618573, 0.12941137378891041, 0.33019041483746447, 0.0023910074574393651 },
    { 0.008991260093335729, 0.1215419913339278, 0.30649398842969028, 0.56297276014304609 },
    { 0.10660417256199362, 0.097204644587583267, 0.68439041545304002, 0.11180076739738309 },
    { 0.32923295974264682, 0.029569495206479609, 0.31790356021339461, 0.32329398483747901 },
    { 0.10384411641099309, 0.43271023904776862, 0.35382323920929709, 0.1096224053319412 } };
static const double XG4_W[11] = { 0.006541848487473325,  0.0092122281926561474, 0.032235135341605742, 0.0092322998119293929,
                                  0.0099888641910932524, 0.011578327656272558,  0.012693785874259723, 0.013237780011337548,
                                  0.017744672359248346,  0.018372372071416277,  0.025829352669374347 };

// p2_div_k_grad_affine_q4::integrateAll: a_ij = sum_q w_q |det J| k(x_q) grad phi_i(x_q) . grad phi_j(x_q)
inline void p2DivKGradElement( const P3* x, const P2KField& k, uint64_t eid, double a[10][10] )
{
   double       G[4][3];
   const double adet = std::fabs( baryGrads( x, G ) );
   for ( int i = 0; i < 10; ++i )
      for ( int j = 0; j < 10; ++j )
         a[i][j] = 0.0;
   for ( int q = 0; q < 11; ++q )
   {
      P3 xq{ 0, 0, 0 };
      for ( int v = 0; v < 4; ++v )
         for ( int c = 0; c < 3; ++c )
            xq[c] += XG4_LAMBDA[q][v] * x[v][c];
      const double s = XG4_W[q] * adet * k( xq, eid );
      double       d[10][3];
      p2Grads( XG4_LAMBDA[q], G, d );
      for ( int i = 0; i < 10; ++i )
         for ( int j = 0; j < 10; ++j )
            a[i][j] += s * ( d[i][0] * d[j][0] + d[i][1] * d[j][1] + d[i][2] * d[j][2] );
   }
}

// P2ElementwiseOperator< p2_div_k_grad_affine_q4 > (elementwiseoperators/P2ElementwiseOperator.cpp:100-310): the
// element matrix of every micro-cell of level lat-1 from the 11-point rule, y_T += A_T x_T, additive reduction to
// the lower-dimensional primitives. Element id of micro-cell (t, anchor a) of cell c: (6 c + t) S3(Nc + 1) +
// index(Nc + 1, a), Nc = 2^(lat-1). Levels are lattice levels.
template < typename T >
struct P2ElementwiseDivKGrad : P2ConstantLaplace< T >
{
   using OperatorBase< T >::m;
   using OperatorBase< T >::invDiag;
   P2KField k;

   P2ElementwiseDivKGrad( const Mesh* mesh, int lo, int hi, const P2KField& kf )
   : P2ConstantLaplace< T >( mesh, lo, hi )
   , k( kf )
   {
      for ( int lat = this->minL; lat <= this->maxL; ++lat )
         computeInverseDiagonalK( lat );
   }

   template < typename Fn >
   void forElements( idx c, int lat, Fn&& fn ) const
   {
      const idx Nc = idx( 1 ) << ( lat - 1 );
      forP2Cells( lat, [&]( int t, const idx( *nd )[3] ) {
         // anchor = (vertex 0 lattice coords) / 2 - type offset of vertex 0
         const idx ax = nd[0][0] / 2 - MICRO_CELL[t][0][0], ay = nd[0][1] / 2 - MICRO_CELL[t][0][1], az = nd[0][2] / 2 - MICRO_CELL[t][0][2];
         const uint64_t eid = uint64_t( ( 6 * c + t ) * S3( Nc + 1 ) + cidx( Nc + 1, ax, ay, az ) );
         P3             xc[4];
         for ( int v = 0; v < 4; ++v )
            xc[v] = microCoord( *m, (int) c, lat - 1, ax + MICRO_CELL[t][v][0], ay + MICRO_CELL[t][v][1], az + MICRO_CELL[t][v][2] );
         double a[10][10];
         p2DivKGradElement( xc, k, eid, a );
         fn( nd, a );
      } );
   }

   void apply( const Function< T >& src, Function< T >& dst, int lat, int flag, bool add ) const override
   {
      src.commDown( lat );
      if ( !add )
         interpolate< T >( dst, 0.0, lat, flag );
      const idx N = idx( 1 ) << lat;
      const idx w = N + 1;
#pragma omp parallel for schedule( dynamic )
      for ( idx c = 0; c < (idx) m->Cl.size(); ++c )
      {
         auto&       dc = dst.c[lat][c];
         const auto& sc = src.c[lat][c];
         for ( idx z = 0; z <= N; ++z )
            for ( idx y = 0; y + z <= N; ++y )
               for ( idx x = 0; x + y + z <= N; ++x )
                  if ( x == 0 || y == 0 || z == 0 || x + y + z == N )
                     dc[cidx( w, x, y, z )] = T( 0 );
         forElements( c, lat, [&]( const idx( *nd )[3], const double( *a )[10] ) {
            idx vi[10];
            for ( int i = 0; i < 10; ++i )
               vi[i] = cidx( w, nd[i][0], nd[i][1], nd[i][2] );
            for ( int i = 0; i < 10; ++i )
            {
               double s = 0.0;
               for ( int j = 0; j < 10; ++j )
                  s += a[i][j] * double( sc[vi[j]] );
               dc[vi[i]] += T( s );
            }
         } );
      }
      commAdditiveCellToLower( dst, lat, All ^ flag, !add );
   }

   void computeInverseDiagonalK( int lat )
   {
      auto&     D = *invDiag;
      const idx w = ( idx( 1 ) << lat ) + 1;
      interpolate< T >( D, 0.0, lat, All );
      for ( idx c = 0; c < (idx) m->Cl.size(); ++c )
      {
         auto& dc = D.c[lat][c];
         std::fill( dc.begin(), dc.end(), T( 0 ) );
         forElements( c, lat, [&]( const idx( *nd )[3], const double( *a )[10] ) {
            for ( int i = 0; i < 10; ++i )
               dc[cidx( w, nd[i][0], nd[i][1], nd[i][2] )] += T( a[i][i] );
         } );
      }
      commAdditiveCellToLower( D, lat, None, true );
      forOwned< T >( *m, lat, All, [&]( int kind, int p, idx i ) {
         T& v = at( D, lat, kind, p, i );
         v    = T( 1.0 ) / v;
      } );
   }
};

// ---------------------------------------------------------------- smoothers and solvers
template < typename T >
struct Solvers
{
   const Mesh*                       m;
   const OperatorBase< T >*          A;
   int                               minL, maxL;
   std::unique_ptr< Function< T > >  tmp, jtmp, ct1, ct2, cgp, cgz, cgap, cgr;
   std::vector< double >             chebCoeffs;
   int                               smoother = 0; // 0 = weighted Jacobi, 1 = Chebyshev
   double                            relax    = 2.0 / 3.0;
   int                               pre = 2, post = 2, inc = 0;
   int                               cgMaxIter = 10000;
   double                            cgTol     = 1e-16;
   const int                         gmgFlag   = Inner | NeumannBoundary;

   Solvers( const Mesh* mesh, const OperatorBase< T >* op, int lo, int hi )
   : m( mesh )
   , A( op )
   , minL( lo )
   , maxL( hi )
   {
      tmp  = std::make_unique< Function< T > >( m, lo, hi );
      jtmp = std::make_unique< Function< T > >( m, lo, hi );
      ct1  = std::make_unique< Function< T > >( m, lo, hi );
      ct2  = std::make_unique< Function< T > >( m, lo, hi );
      cgp  = std::make_unique< Function< T > >( m, lo, hi );
      cgz  = std::make_unique< Function< T > >( m, lo, hi );
      cgap = std::make_unique< Function< T > >( m, lo, hi );
      cgr  = std::make_unique< Function< T > >( m, lo, hi );
   }

   // P1Operator::smooth_jac
   void smoothJac( Function< T >& dst, const Function< T >& rhs, const Function< T >& src, double w, int L, int flag ) const
   {
      A->apply( src, dst, L, flag, false );
      assign< T >( dst, { 1.0, -1.0 }, { &rhs, &dst }, L, flag );
      multElementwise< T >( dst, *A->invDiag, dst, L, flag );
      assign< T >( dst, { 1.0, w }, { &src, &dst }, L, flag );
   }
   void jacobiSolve( Function< T >& x, const Function< T >& b, int L )
   {
      assign< T >( *jtmp, { 1.0 }, { &x }, L, All );
      smoothJac( x, b, *jtmp, relax, L, gmgFlag );
   }
   // ChebyshevSmoother::solve
   void chebyshevSolve( Function< T >& x, const Function< T >& b, int L )
   {
      const int flag = gmgFlag;
      A->apply( x, *ct1, L, flag, false );
      assign< T >( *ct1, { 1.0, -1.0 }, { &b, ct1.get() }, L, flag );
      multElementwise< T >( *ct1, *A->invDiag, *ct1, L, flag );
      assign< T >( x, { 1.0, chebCoeffs[0] }, { &x, ct1.get() }, L, flag );
      for ( size_t k = 1; k < chebCoeffs.size(); ++k )
      {
         A->apply( *ct1, *ct2, L, flag, false );
         multElementwise< T >( *ct1, *A->invDiag, *ct2, L, flag );
         assign< T >( x, { 1.0, chebCoeffs[k] }, { &x, ct1.get() }, L, flag );
      }
   }
   void smooth( Function< T >& x, const Function< T >& b, int L )
   {
      if ( smoother == 0 )
         jacobiSolve( x, b, L );
      else if ( smoother >= 4 )
      {
         // SORSmoother / SymmetricSORSmoother::solve (solvers/SORSmoother.hpp:38-50, SymmetricSORSmoother.hpp:37-60)
         // with smooth_sor in the reference's order
         const auto* c = dynamic_cast< const P1ConstantOperator< T >* >( A );
         if ( !c )
            throw std::invalid_argument( "SOR smoothing needs the constant P1 operator" );
         c->smoothSorOrdered( x, b, relax, L, gmgFlag, false );
         if ( smoother == 5 )
            c->smoothSorOrdered( x, b, relax, L, gmgFlag, true );
      }
      else
         chebyshevSolve( x, b, L );
   }
   // CGSolver::solve with the identity preconditioner
   int cgSolve( Function< T >& x, const Function< T >& b, int L )
   {
      const int flag = Inner | NeumannBoundary | FreeslipBoundary;
      A->apply( x, *cgp, L, flag, false );
      assign< T >( *cgr, { 1.0, -1.0 }, { &b, cgp.get() }, L, flag );
      assign< T >( *cgz, { 1.0 }, { cgr.get() }, L, flag );
      assign< T >( *cgp, { 1.0 }, { cgz.get() }, L, flag );
      double prsold = dot< T >( *cgr, *cgz, L, flag );
      double res0   = std::sqrt( dot< T >( *cgr, *cgr, L, flag ) );
      if ( res0 < cgTol )
         return 0;
      int it = 0;
      for ( ; it < cgMaxIter; ++it )
      {
         A->apply( *cgp, *cgap, L, flag, false );
         double pAp   = dot< T >( *cgp, *cgap, L, flag );
         double alpha = prsold / pAp;
         addScaled< T >( x, alpha, *cgp, L, flag );
         addScaled< T >( *cgr, -alpha, *cgap, L, flag );
         double rsnew = dot< T >( *cgr, *cgr, L, flag );
         if ( std::sqrt( rsnew ) < cgTol )
            break;
         assign< T >( *cgz, { 1.0 }, { cgr.get() }, L, flag );
         double prsnew = dot< T >( *cgr, *cgz, L, flag );
         double beta   = prsnew / prsold;
         assign< T >( *cgp, { 1.0, beta }, { cgz.get(), cgp.get() }, L, flag );
         prsold = prsnew;
      }
      return it + 1;
   }
   // GeometricMultigridSolver::solveRecursively (V-cycle)
   void vcycle( Function< T >& x, Function< T >& b, int L, int invoked )
   {
      if ( L == minL )
      {
         cgSolve( x, b, L );
         return;
      }
      const int preS = pre + inc * ( invoked - L );
      for ( int i = 0; i < preS; ++i )
         smooth( x, b, L );
      A->apply( x, *tmp, L, gmgFlag, false );
      assign< T >( *tmp, { 1.0, -1.0 }, { &b, tmp.get() }, L, gmgFlag );
      A->restrictOp( *tmp, L, gmgFlag );
      assign< T >( b, { 1.0 }, { tmp.get() }, L - 1, gmgFlag );
      interpolate< T >( x, 0.0, L - 1, All );
      vcycle( x, b, L - 1, invoked );
      A->prolongateOp( x, L - 1, gmgFlag, true );
      const int postS = post + inc * ( invoked - L );
      for ( int i = 0; i < postS; ++i )
         smooth( x, b, L );
   }
   // relative residual ||b - Ax|| over Inner DoFs
   double residualNorm( const Function< T >& x, const Function< T >& b, int L )
   {
      A->apply( x, *tmp, L, gmgFlag, false );
      assign< T >( *tmp, { 1.0, -1.0 }, { &b, tmp.get() }, L, gmgFlag );
      return std::sqrt( dot< T >( *tmp, *tmp, L, gmgFlag ) );
   }
};

// ChebyshevSmoother::setupCoefficients (order 1..5, mfem form)
inline int chebCoefficients( int order, double lower, double upper, double* c )
{
   const double theta = 0.5 * ( upper + lower );
   const double delta = 0.5 * ( upper - lower );
   switch ( order - 1 )
   {
   case 0:
      c[0] = 1.0 / theta;
      return 0;
   case 1: {
      const double t0 = 1.0 / ( std::pow( delta, 2 ) - 2 * std::pow( theta, 2 ) );
      c[0]            = -4 * theta * t0;
      c[1]            = 2 * t0;
      return 0;
   }
   case 2: {
      const double t0 = 3 * std::pow( delta, 2 );
      const double t1 = std::pow( theta, 2 );
      const double t2 = 1.0 / ( -4 * std::pow( theta, 3 ) + theta * t0 );
      c[0]            = t2 * ( t0 - 12 * t1 );
      c[1]            = 12 / ( t0 - 4 * t1 );
      c[2]            = -4 * t2;
      return 0;
   }
   case 3: {
      const double t0 = std::pow( delta, 2 );
      const double t1 = std::pow( theta, 2 );
      const double t2 = 8 * t0;
      const double t3 = 1.0 / ( std::pow( delta, 4 ) + 8 * std::pow( theta, 4 ) - t1 * t2 );
      c[0]            = t3 * ( 32 * std::pow( theta, 3 ) - 16 * theta * t0 );
      c[1]            = t3 * ( -48 * t1 + t2 );
      c[2]            = 32 * theta * t3;
      c[3]            = -8 * t3;
      return 0;
   }
   case 4: {
      const double t0 = 5 * std::pow( delta, 4 );
      const double t1 = std::pow( theta, 4 );
      const double t2 = std::pow( theta, 2 );
      const double t3 = std::pow( delta, 2 );
      const double t4 = 60 * t3;
      const double t5 = 20 * t3;
      const double t6 = 1.0 / ( 16 * std::pow( theta, 5 ) - std::pow( theta, 3 ) * t5 + theta * t0 );
      const double t7 = 160 * t2;
      const double t8 = 1.0 / ( t0 + 16 * t1 - t2 * t5 );
      c[0]            = t6 * ( t0 + 80 * t1 - t2 * t4 );
      c[1]            = t8 * ( t4 - t7 );
      c[2]            = t6 * ( -t5 + t7 );
      c[3]            = -80 * t8;
      c[4]            = 16 * t6;
      return 0;
   }
   default:
      return -1;
   }
}

} // namespace orc

// ================================================================ C ABI for ctypes (tests / bench only)
using namespace orc;

struct OrFunc
{
   int                                   prec; // 8 or 4
   std::unique_ptr< Function< double > > d;
   std::unique_ptr< Function< float > >  f;
};

struct OrOp
{
   int                                             prec;
   const Mesh*                                     m;
   std::unique_ptr< OperatorBase< double > >        d;
   std::unique_ptr< OperatorBase< float > >         f;
};

template < typename T >
static Function< T >& fget( OrFunc* h );
template <>
Function< double >& fget< double >( OrFunc* h )
{
   return *h->d;
}
template <>
Function< float >& fget< float >( OrFunc* h )
{
   return *h->f;
}

template < typename T >
static void setCanonical( Function< T >& fn, int L, const double* in )
{
   const Mesh& m = *fn.m;
   const idx   w = ( idx( 1 ) << L ) + 1;
   forEachDof< T >( m, L, [&]( int kind, int p, idx i, idx j, idx k, idx n ) {
      switch ( kind )
      {
      case 0:
         fn.v[L][p][0] = T( in[n] );
         break;
      case 1:
         fn.e[L][p][i] = T( in[n] );
         break;
      case 2:
         fn.f[L][p][fidx( w, i, j )] = T( in[n] );
         break;
      default:
         fn.c[L][p][cidx( w, i, j, k )] = T( in[n] );
      }
   } );
   fn.commDown( L );
}

template < typename T >
static void getCanonical( const Function< T >& fn, int L, double* out )
{
   const Mesh& m = *fn.m;
   const idx   w = ( idx( 1 ) << L ) + 1;
   forEachDof< T >( m, L, [&]( int kind, int p, idx i, idx j, idx k, idx n ) {
      switch ( kind )
      {
      case 0:
         out[n] = double( fn.v[L][p][0] );
         break;
      case 1:
         out[n] = double( fn.e[L][p][i] );
         break;
      case 2:
         out[n] = double( fn.f[L][p][fidx( w, i, j )] );
         break;
      default:
         out[n] = double( fn.c[L][p][cidx( w, i, j, k )] );
      }
   } );
}

// GMG parameters: ip = {minL, maxL, pre, post, inc, smoother(0 jac,1 cheb), chebOrder, cgMaxIter}
//                 dp = {relax, chebLower, chebUpper, cgTol}
// Runs `cycles` V-cycles on x with rhs b at level maxL; res[0] = initial residual norm, res[k] after cycle k.
template < typename T >
static void gmgRun( OrOp* o, OperatorBase< T >* A, Function< T >& x, Function< T >& b, const int* ip, const double* dp,
                    int cycles, double* res )
{
   Solvers< T > S( o->m, A, ip[0], ip[1] );
   S.pre       = ip[2];
   S.post      = ip[3];
   S.inc       = ip[4];
   S.smoother  = ip[5];
   S.cgMaxIter = ip[7];
   S.relax     = dp[0];
   S.cgTol     = dp[3];
   if ( S.smoother == 1 )
   {
      S.chebCoeffs.resize( ip[6] );
      chebCoefficients( ip[6], dp[1], dp[2], S.chebCoeffs.data() );
   }
   const int L = ip[1];
   res[0]      = S.residualNorm( x, b, L );
   for ( int c = 0; c < cycles; ++c )
   {
      S.vcycle( x, b, L, L );
      res[c + 1] = S.residualNorm( x, b, L );
   }
}

// MeshInfo::meshSphericalShell( ntan, nrad, rmin, rmax ) with SHELLMESH_CLASSIC (mesh/MeshGenSphericalShell.cpp):
// elemIdx2Tuple, vertIdx2Tuple, tuple2VertIdx, setupCoordsClassic, computeSphericalShellVertices and the
// six-tetrahedra splittings tNode / sNode of a diamond cell, followed by the usual storage setup.
namespace shell {
struct Gen
{
   idx                     nt, nr;
   std::vector< double >   layers;
   double                  ico[12][3];
   idx                     dn[10][4] = { { 0, 5, 6, 1 },   { 0, 1, 7, 2 },  { 0, 2, 8, 3 },  { 0, 3, 9, 4 },  { 0, 4, 10, 5 },
                                         { 11, 7, 1, 6 }, { 11, 8, 2, 7 }, { 11, 9, 3, 8 }, { 11, 10, 4, 9 }, { 11, 6, 5, 10 } };
   std::vector< P3 >       nc; // [is1][is2][id] on the unit sphere
   P3&                     node( idx a, idx b, idx d ) { return nc[size_t( ( a * nt + b ) * 10 + d )]; }

   void setup()
   {
      const double fifthpi = 0.4 * std::asin( 1.0 );
      const double w       = 2.0 * std::acos( 1.0 / ( 2.0 * std::sin( fifthpi ) ) );
      const double cw = std::cos( w ), sw = std::sin( w );
      ico[0][0] = 0.0, ico[0][1] = 0.0, ico[0][2] = 1.0;
      ico[11][0] = 0.0, ico[11][1] = 0.0, ico[11][2] = -1.0;
      for ( int k = 1; k <= 5; ++k )
      {
         double phi = 2.0 * ( double( k ) - 0.5 ) * fifthpi;
         ico[k][0] = sw * std::cos( phi ), ico[k][1] = sw * std::sin( phi ), ico[k][2] = cw;
         phi = 2.0 * ( double( k ) - 1.0 ) * fifthpi;
         ico[k + 5][0] = sw * std::cos( phi ), ico[k + 5][1] = sw * std::sin( phi ), ico[k + 5][2] = -cw;
      }
      nc.assign( size_t( nt * nt * 10 ), P3{ 0, 0, 0 } );
      const idx n1 = nt - 1;
      for ( idx id = 0; id < 10; ++id )
      {
         const idx L = id < 5 ? 1 : 3, R = id < 5 ? 3 : 1;
         for ( int c = 0; c < 3; ++c )
         {
            node( 0, 0, id )[c]   = ico[dn[id][0]][c];
            node( n1, 0, id )[c]  = ico[dn[id][L]][c];
            node( n1, n1, id )[c] = ico[dn[id][2]][c];
            node( 0, n1, id )[c]  = ico[dn[id][R]][c];
         }
         auto arcMid = [&]( idx a1, idx a2, idx b1, idx b2, idx t1, idx t2 ) {
            P3     v{ node( a1, a2, id )[0] + node( b1, b2, id )[0], node( a1, a2, id )[1] + node( b1, b2, id )[1],
                  node( a1, a2, id )[2] + node( b1, b2, id )[2] };
            double sc = std::sqrt( v[0] * v[0] + v[1] * v[1] + v[2] * v[2] );
            node( t1, t2, id ) = { v[0] / sc, v[1] / sc, v[2] / sc };
         };
         for ( idx m = 1; m < n1; m *= 2 )
         {
            const idx l = n1 / m, l2 = l / 2;
            for ( idx j1 = 0; j1 <= m; ++j1 )
               for ( idx j2 = 0; j2 < m; ++j2 )
               {
                  idx i1 = j1 * l, i2 = j2 * l + l2;
                  arcMid( i1, i2 - l2, i1, i2 + l2, i1, i2 );
               }
            for ( idx j1 = 0; j1 <= m; ++j1 )
               for ( idx j2 = 0; j2 < m; ++j2 )
               {
                  idx i1 = j2 * l + l2, i2 = j1 * l;
                  arcMid( i1 - l2, i2, i1 + l2, i2, i1, i2 );
               }
            for ( idx j1 = 0; j1 < m; ++j1 )
               for ( idx j2 = 0; j2 < m; ++j2 )
               {
                  idx i1 = j1 * l + l2, i2 = j2 * l + l2;
                  arcMid( i1 - l2, i2 + l2, i1 + l2, i2 - l2, i1, i2 );
               }
         }
      }
   }
   void vertIdx2Tuple( idx jv, idx& is1, idx& is2, idx& id, idx& ir ) const
   {
      if ( jv >= 2 * nr )
      {
         idx i = jv - 2 * nr;
         ir    = i / ( 10 * ( nt - 1 ) * ( nt - 1 ) );
         i -= ir * 10 * ( nt - 1 ) * ( nt - 1 );
         id = i / ( ( nt - 1 ) * ( nt - 1 ) );
         i -= id * ( nt - 1 ) * ( nt - 1 );
         is2 = i / ( nt - 1 );
         i -= is2 * ( nt - 1 );
         ++is2;
         is1 = i;
      }
      else if ( jv < nr )
         is1 = 0, is2 = 0, id = 0, ir = jv;
      else
         is1 = 0, is2 = 0, id = 5, ir = jv - nr;
   }
   idx tuple2VertIdx( idx is1, idx is2, idx id, idx ir ) const
   {
      if ( is1 == 0 && is2 == 0 )
         return id < 5 ? ir : nr + ir;
      if ( is2 == 0 )
      {
         is2 = is1;
         is1 = 0;
         id  = id < 5 ? ( id + 4 ) % 5 : ( id + 4 ) % 5 + 5;
      }
      else if ( is1 == nt - 1 )
      {
         is1 = ( nt - 1 ) - is2;
         is2 = nt - 1;
         id  = id < 5 ? ( id + 4 ) % 5 + 5 : id % 5;
      }
      return ir * ( nt - 1 ) * ( nt - 1 ) * 10 + id * ( nt - 1 ) * ( nt - 1 ) + ( is2 - 1 ) * ( nt - 1 ) + is1 + 2 * nr;
   }
};
} // namespace shell

// MeshInfo::meshSymmetricCuboid (mesh/MeshGenSymmetricCuboid.cpp:35-253). Restated with the reference's
// per-sub-cube vertex map: a local vertex shared with the sub-cube below (z - 1: local 0-4 are its 9-13), in
// front (y - 1: 0, 1, 5, 9, 10 are its 3, 4, 8, 12, 13) or to the left (x - 1: 0, 3, 6, 9, 12 are its 1, 4, 7,
// 10, 13) takes that vertex's ID (:110-220); the rest get new IDs in local order. Coordinates are
// offset + local position per sub-cube, the last sub-cube writing a shared vertex winning (:229-239).
static std::unique_ptr< Mesh > makeSymmetricCuboid( const double* lo, const double* hi, int nx, int ny, int nz )
{
   const double wx = ( hi[0] - lo[0] ) / nx, wy = ( hi[1] - lo[1] ) / ny, wz = ( hi[2] - lo[2] ) / nz;
   const P3     local[14] = { { 0, 0, 0 },           { wx, 0, 0 },           { wx / 2, wy / 2, 0 },  { 0, wy, 0 },
                              { wx, wy, 0 },         { wx / 2, 0, wz / 2 },  { 0, wy / 2, wz / 2 },  { wx, wy / 2, wz / 2 },
                              { wx / 2, wy, wz / 2 }, { 0, 0, wz },           { wx, 0, wz },          { wx / 2, wy / 2, wz },
                              { 0, wy, wz },         { wx, wy, wz } };
   const int    cellV[24][4] = { { 0, 1, 2, 5 },   { 0, 3, 2, 6 },   { 1, 2, 4, 7 },    { 2, 4, 3, 8 },    { 2, 0, 5, 6 },   { 2, 1, 5, 7 },
                                 { 2, 3, 6, 8 },   { 2, 4, 7, 8 },   { 0, 9, 5, 6 },    { 1, 10, 5, 7 },   { 3, 12, 6, 8 },  { 4, 13, 7, 8 },
                                 { 2, 5, 6, 8 },   { 2, 5, 7, 8 },   { 11, 5, 6, 8 },   { 11, 5, 7, 8 },   { 11, 9, 5, 6 },  { 11, 10, 5, 7 },
                                 { 11, 12, 6, 8 }, { 11, 13, 7, 8 }, { 11, 9, 10, 5 },  { 11, 9, 12, 6 },  { 11, 10, 13, 7 }, { 11, 12, 13, 8 } };
   const int    fromBelow[14] = { 9, 10, 11, 12, 13, -1, -1, -1, -1, -1, -1, -1, -1, -1 };
   const int    fromFront[14] = { 3, 4, -1, -1, -1, 8, -1, -1, -1, 12, 13, -1, -1, -1 };
   const int    fromLeft[14]  = { 1, -1, -1, 4, -1, -1, 7, -1, -1, 10, -1, -1, 13, -1 };
   std::vector< std::array< idx, 14 > > vmap( size_t( nx ) * ny * nz );
   auto at = [&]( int x, int y, int z ) -> std::array< idx, 14 >& { return vmap[( size_t( z ) * ny + y ) * nx + x]; };
   idx  next = 0;
   std::map< idx, P3 >           nodes;
   std::set< std::vector< idx > > cells;
   for ( int z = 0; z < nz; ++z )
      for ( int y = 0; y < ny; ++y )
         for ( int x = 0; x < nx; ++x )
         {
            auto& m = at( x, y, z );
            for ( int i = 0; i < 14; ++i )
            {
               if ( z > 0 && fromBelow[i] >= 0 )
                  m[i] = at( x, y, z - 1 )[fromBelow[i]];
               else if ( y > 0 && fromFront[i] >= 0 )
                  m[i] = at( x, y - 1, z )[fromFront[i]];
               else if ( x > 0 && fromLeft[i] >= 0 )
                  m[i] = at( x - 1, y, z )[fromLeft[i]];
               else
                  m[i] = next++;
            }
         }
   for ( int z = 0; z < nz; ++z )
      for ( int y = 0; y < ny; ++y )
         for ( int x = 0; x < nx; ++x )
         {
            const P3 off{ lo[0] + wx * x, lo[1] + wy * y, lo[2] + wz * z };
            for ( int i = 0; i < 14; ++i )
               nodes[at( x, y, z )[i]] = P3{ off[0] + local[i][0], off[1] + local[i][1], off[2] + local[i][2] };
            for ( auto& c : cellV )
               cells.insert( { at( x, y, z )[c[0]], at( x, y, z )[c[1]], at( x, y, z )[c[2]], at( x, y, z )[c[3]] } );
         }
   return buildMesh( nodes, cells );
}

static std::unique_ptr< Mesh > makeShell( idx ntan, idx nrad, double rmin, double rmax )
{
   if ( ntan < 2 || ( ( ntan - 1 ) & ( ntan - 2 ) ) != 0 || nrad < 2 || !( rmin > 0 ) || !( rmin < rmax ) )
      return nullptr;
   shell::Gen g;
   g.nt = ntan;
   g.nr = nrad;
   for ( idx k = 0; k < nrad; ++k )
      g.layers.push_back( rmin + ( ( rmax - rmin ) / double( nrad - 1 ) ) * double( k ) );
   g.setup();
   std::map< idx, P3 > nodes;
   const idx           nVerts = ( 10 * ( ntan - 1 ) * ( ntan - 1 ) + 2 ) * nrad;
   for ( idx v = 0; v < nVerts; ++v )
   {
      idx is1, is2, id, ir;
      g.vertIdx2Tuple( v, is1, is2, id, ir );
      const P3& u = g.node( is1, is2, id );
      nodes[v]    = { u[0] * g.layers[ir], u[1] * g.layers[ir], u[2] * g.layers[ir] };
   }
   static const idx tNode[6][4] = { { 0, 1, 3, 7 }, { 4, 7, 5, 0 }, { 0, 5, 1, 7 }, { 1, 2, 3, 6 }, { 5, 7, 6, 1 }, { 3, 6, 7, 1 } };
   static const idx sNode[6][4] = { { 4, 7, 5, 1 }, { 0, 1, 3, 4 }, { 4, 3, 7, 1 }, { 7, 6, 5, 2 }, { 3, 1, 2, 7 }, { 5, 2, 1, 7 } };
   static const idx offset[8][3] = { { 0, 0, 1 }, { 1, 0, 1 }, { 1, 1, 1 }, { 0, 1, 1 },
                                     { 0, 0, 0 }, { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 } };
   std::set< std::vector< idx > > cells;
   const idx                      nElems = ( ntan - 1 ) * ( ntan - 1 ) * ( nrad - 1 ) * 10 * 6;
   for ( idx jelem = 0; jelem < nElems; ++jelem )
   {
      idx x = jelem;
      idx ir = x / ( 60 * ( ntan - 1 ) * ( ntan - 1 ) );
      x -= ir * 60 * ( ntan - 1 ) * ( ntan - 1 );
      idx id = x / ( 6 * ( ntan - 1 ) * ( ntan - 1 ) );
      x -= id * 6 * ( ntan - 1 ) * ( ntan - 1 );
      idx is2 = x / ( 6 * ( ntan - 1 ) );
      x -= is2 * 6 * ( ntan - 1 );
      idx is1 = x / 6;
      idx it  = x - 6 * is1;
      idx vIdx[8];
      for ( int k = 0; k < 8; ++k )
      {
         const idx ks1 = is1 + ( id <= 4 ? offset[k][0] : offset[k][1] );
         const idx ks2 = is2 + ( id <= 4 ? offset[k][1] : offset[k][0] );
         vIdx[k]       = g.tuple2VertIdx( ks1, ks2, id, ir + offset[k][2] );
      }
      const idx* tn = id <= 4 ? tNode[it] : sNode[it];
      cells.insert( { vIdx[tn[0]], vIdx[tn[1]], vIdx[tn[2]], vIdx[tn[3]] } );
   }
   return buildMesh( nodes, cells );
}

// ---------------------------------------------------------------- reference memory with ghost layers
// The per-primitive vertex-DoF memory of one level (p1functionspace/VertexDoFMemory.hpp:44-64) as
// VertexDoFFunction::communicate fills it: the owned arrays after the downward chain, the vertex ghost entries from
// packEdgeForVertex (VertexDoFPackInfo.hpp:204-245), the edge ghost rows from packFaceForEdge / unpackEdgeFromFace
// (:295-408: the face row next to the edge, then the face's cell ghost rows into the edge's cell slots), and the face
// ghost layers from packCellForFace / unpackFaceFromCell (:508-565) with the cell boundary iterator of
// indexing/MacroCellIndexing.cpp:160-300 (start at vertex0, offset towards the fourth vertex, rows along
// vertex0 -> vertex1 wrapping along vertex0 -> vertex2).
namespace halo {
// cell-local unit step from local vertex a towards local vertex b in (x, y, z) lattice coordinates
inline I3 step( int a, int b )
{
   static const I3 pos[4] = { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
   return { pos[b][0] - pos[a][0], pos[b][1] - pos[a][1], pos[b][2] - pos[a][2] };
}
// points of CellBoundaryIterator( w, v0, v1, v2, offset ) in iteration order
inline std::vector< I3 > cellBoundary( idx w, int v0, int v1, int v2, idx off )
{
   const idx N = w - 1;
   int       v3 = 6 - v0 - v1 - v2;
   I3        start{ 0, 0, 0 };
   if ( v0 > 0 )
      start[v0 - 1] = N;
   const I3 toward3 = step( v0, v3 ), d1 = step( v0, v1 ), d2 = step( v0, v2 );
   for ( int k = 0; k < 3; ++k )
      start[k] += off * toward3[k];
   std::vector< I3 > pts;
   const idx         len = w - off;
   for ( idx r = 0; r < len; ++r )
      for ( idx q = 0; q + r < len; ++q )
         pts.push_back( { start[0] + q * d1[0] + r * d2[0], start[1] + q * d1[1] + r * d2[1], start[2] + q * d1[2] + r * d2[2] } );
   return pts;
}
// face-local (x, y) points one row into a face of width `width` from its local edge le (0: v0-v1, 1: v0-v2, 2: v1-v2),
// walked from the given edge vertex end (fromFirst: from the pair's first vertex), `count` points
inline std::vector< std::array< idx, 2 > > faceRow( idx width, int le, bool fromFirst, idx count )
{
   std::vector< std::array< idx, 2 > > pts;
   const idx                           N = width - 1;
   for ( idx i = 0; i < count; ++i )
   {
      const idx j = fromFirst ? i : count - 1 - i;
      if ( le == 0 )
         pts.push_back( { j, 1 } );
      else if ( le == 1 )
         pts.push_back( { 1, j } );
      else
         pts.push_back( { N - 1 - j, j } );
   }
   return pts;
}
} // namespace halo

template < typename T >
static void haloPack( const Function< T >& F, int L, double* out )
{
   F.commDown( L );
   const Mesh& m = *F.m;
   const idx   N = idx( 1 ) << L, w = N + 1;
   const int   nV = int( m.X.size() ), nE = int( m.E.size() ), nF = int( m.F.size() );
   std::vector< std::vector< int > > vEdges( nV ), eFaces( nE );
   for ( int e = 0; e < nE; ++e )
   {
      vEdges[m.E[e][0]].push_back( e );
      vEdges[m.E[e][1]].push_back( e );
   }
   auto edgeOf = [&]( int a, int b ) {
      for ( int e = 0; e < nE; ++e )
         if ( ( m.E[e][0] == a && m.E[e][1] == b ) || ( m.E[e][0] == b && m.E[e][1] == a ) )
            return e;
      return -1;
   };
   for ( int f = 0; f < nF; ++f )
   {
      const auto& t = m.F[f];
      for ( int e : { edgeOf( t[0], t[1] ), edgeOf( t[0], t[2] ), edgeOf( t[1], t[2] ) } )
         eFaces[e].push_back( f );
   }
   auto local = [&]( int c, int v ) {
      for ( int l = 0; l < 4; ++l )
         if ( m.Cl[c][l] == v )
            return l;
      return -1;
   };
   // the ghost triangle of face f from its cell c, as face memory (fidx(N, x, y) order)
   auto faceGhost = [&]( int f, int c ) {
      const auto& t   = m.F[f];
      const auto  pts = halo::cellBoundary( w, local( c, t[0] ), local( c, t[1] ), local( c, t[2] ), 1 );
      std::vector< double > g;
      for ( const auto& q : pts )
         g.push_back( double( F.c[L][c][cidx( w, q[0], q[1], q[2] )] ) );
      return g;
   };
   size_t o = 0;
   for ( int v = 0; v < nV; ++v )
   {
      out[o++] = double( F.v[L][v][0] );
      for ( int e : vEdges[v] )
         out[o++] = double( F.e[L][e][m.E[e][0] == v ? 1 : N - 1] );
   }
   for ( int e = 0; e < nE; ++e )
   {
      const size_t base = o;
      for ( idx i = 0; i <= N; ++i )
         out[o++] = double( F.e[L][e][i] );
      std::vector< int > cells;
      for ( size_t c = 0; c < m.Cl.size(); ++c )
         for ( int k = 0; k < 6; ++k )
            if ( m.cE[c][k] == e )
               cells.push_back( int( c ) );
      const size_t cellBase = base + size_t( w ) + eFaces[e].size() * size_t( w - 1 );
      for ( int f : eFaces[e] )
      {
         const auto& t = m.F[f];
         int         le = 0, first = t[0];
         if ( edgeOf( t[0], t[1] ) == e )
            le = 0, first = t[0];
         else if ( edgeOf( t[0], t[2] ) == e )
            le = 1, first = t[0];
         else
            le = 2, first = t[1];
         const bool fromFirst = m.E[e][0] == first;
         for ( const auto& q : halo::faceRow( w, le, fromFirst, N ) )
            out[o++] = double( F.f[L][f][fidx( w, q[0], q[1] )] );
         // the face's cell ghost rows go to the edge's slots of those cells
         for ( size_t k = 0; k < m.fC[f].size(); ++k )
         {
            const int  c    = m.fC[f][k];
            const auto g    = faceGhost( f, c );
            const int  slot = int( std::find( cells.begin(), cells.end(), c ) - cells.begin() );
            idx        i    = 0;
            for ( const auto& q : halo::faceRow( N, le, fromFirst, N - 1 ) )
               out[cellBase + size_t( slot ) * size_t( w - 2 ) + size_t( i++ )] = g[size_t( fidx( N, q[0], q[1] ) )];
         }
      }
      o = cellBase + cells.size() * size_t( w - 2 );
   }
   for ( int f = 0; f < nF; ++f )
   {
      for ( size_t k = 0; k < F.f[L][f].size(); ++k )
         out[o++] = double( F.f[L][f][k] );
      for ( int c : m.fC[f] )
         for ( double g : faceGhost( f, c ) )
            out[o++] = g;
   }
}

extern "C" {

void* or_mesh_load( const char* path ) { return loadGmsh( path ).release(); }
void* or_mesh_load_tags( const char* path ) { return loadGmsh( path, true ).release(); }
void* or_mesh_shell( int ntan, int nrad, double rmin, double rmax ) { return makeShell( ntan, nrad, rmin, rmax ).release(); }
// MeshInfo::meshCuboid (mesh/MeshGenCuboid.cpp:35-174): grid points numbered x fastest; every grid cube split into
// six tetrahedra given by the reference's corner table (corners counter-clockwise in the bottom layer, then the
// top layer)
void* or_mesh_cuboid( const double* lo, const double* hi, int nx, int ny, int nz )
{
   if ( nx < 1 || ny < 1 || nz < 1 )
      return nullptr;
   const int                      tNode[6][4] = { { 4, 5, 7, 3 }, { 0, 3, 1, 4 }, { 4, 1, 5, 3 }, { 5, 6, 7, 2 }, { 1, 3, 2, 5 }, { 7, 2, 3, 5 } };
   const int                      dx[8] = { 0, 1, 1, 0, 0, 1, 1, 0 }, dy[8] = { 0, 0, 1, 1, 0, 0, 1, 1 }, dz[8] = { 0, 0, 0, 0, 1, 1, 1, 1 };
   const idx                      sx = 1, sy = nx + 1, sz = idx( nx + 1 ) * ( ny + 1 );
   std::map< idx, P3 >            nodes;
   std::set< std::vector< idx > > cells;
   for ( idx v = 0; v < sz * ( nz + 1 ); ++v )
   {
      const int i = int( v % sy ), j = int( ( v / sy ) % ( ny + 1 ) ), k = int( v / sz );
      nodes[v]    = { lo[0] + i * ( ( hi[0] - lo[0] ) / nx ), lo[1] + j * ( ( hi[1] - lo[1] ) / ny ), lo[2] + k * ( ( hi[2] - lo[2] ) / nz ) };
   }
   for ( int k = 0; k < nz; ++k )
      for ( int j = 0; j < ny; ++j )
         for ( int i = 0; i < nx; ++i )
            for ( int t = 0; t < 6; ++t )
            {
               std::vector< idx > cv( 4 );
               for ( int q = 0; q < 4; ++q )
               {
                  const int c = tNode[t][q];
                  cv[q]       = ( i + dx[c] ) * sx + ( j + dy[c] ) * sy + ( k + dz[c] ) * sz;
               }
               cells.insert( cv );
            }
   return buildMesh( nodes, cells ).release();
}

// MeshInfo::meshCubedDomain (mesh/MeshGenCubedDomain.cpp:37-383): unit cubes at integer positions; type 0 splits a
// cube into six tets with the cuboid's corner table, type 1 into 24 around its face centres (local vertices
// 0..13 as in the symmetric cuboid). Vertex IDs per local vertex (the reference's symmCubeID / indexInCube).
void* or_mesh_cubed_domain( const int* cubes, int n, int type )
{
   if ( n < 1 || type < 0 || type > 1 )
      return nullptr;
   std::set< std::array< int, 3 > > cs;
   for ( int q = 0; q < n; ++q )
      cs.insert( { cubes[3 * q], cubes[3 * q + 1], cubes[3 * q + 2] } );
   int xmin = INT_MAX, ymin = INT_MAX, zmin = INT_MAX, xmax = INT_MIN, ymax = INT_MIN, zmax = INT_MIN;
   for ( const auto& c : cs )
   {
      xmin = std::min( xmin, c[0] ), ymin = std::min( ymin, c[1] ), zmin = std::min( zmin, c[2] );
      xmax = std::max( xmax, c[0] + 1 ), ymax = std::max( ymax, c[1] + 1 ), zmax = std::max( zmax, c[2] + 1 );
   }
   const idx mx = xmax - xmin, my = ymax - ymin, mz = zmax - zmin;
   auto      inCube = [&]( int x, int y, int z, idx nx, idx ny ) { return idx( z - zmin ) * nx * ny + idx( y - ymin ) * nx + idx( x - xmin ); };
   const idx nCorner = ( mx + 1 ) * ( my + 1 ) * ( mz + 1 ), nXY = mx * my * ( mz + 1 ), nXZ = mx * mz * ( my + 1 );
   auto      symmId = [&]( int x, int y, int z, int lv ) -> idx {
      switch ( lv )
      {
      case 0: return inCube( x, y, z, mx + 1, my + 1 );
      case 1: return inCube( x + 1, y, z, mx + 1, my + 1 );
      case 3: return inCube( x, y + 1, z, mx + 1, my + 1 );
      case 4: return inCube( x + 1, y + 1, z, mx + 1, my + 1 );
      case 9: return inCube( x, y, z + 1, mx + 1, my + 1 );
      case 10: return inCube( x + 1, y, z + 1, mx + 1, my + 1 );
      case 12: return inCube( x, y + 1, z + 1, mx + 1, my + 1 );
      case 13: return inCube( x + 1, y + 1, z + 1, mx + 1, my + 1 );
      case 2: return nCorner + inCube( x, y, z, mx, my );
      case 11: return nCorner + inCube( x, y, z + 1, mx, my );
      case 5: return nCorner + nXY + inCube( x, y, z, mx, my + 1 );
      case 8: return nCorner + nXY + inCube( x, y + 1, z, mx, my + 1 );
      case 6: return nCorner + nXY + nXZ + inCube( x, y, z, mx + 1, my );
      default: return nCorner + nXY + nXZ + inCube( x + 1, y, z, mx + 1, my ); // 7
      }
   };
   const P3  local[14] = { { 0, 0, 0 },       { 1, 0, 0 }, { 0.5, 0.5, 0 }, { 0, 1, 0 }, { 1, 1, 0 },     { 0.5, 0, 0.5 }, { 0, 0.5, 0.5 },
                           { 1, 0.5, 0.5 },   { 0.5, 1, 0.5 }, { 0, 0, 1 }, { 1, 0, 1 }, { 0.5, 0.5, 1 }, { 0, 1, 1 },     { 1, 1, 1 } };
   const int cellV[24][4] = { { 0, 1, 2, 5 },   { 0, 3, 2, 6 },   { 1, 2, 4, 7 },    { 2, 4, 3, 8 },    { 2, 0, 5, 6 },   { 2, 1, 5, 7 },
                              { 2, 3, 6, 8 },   { 2, 4, 7, 8 },   { 0, 9, 5, 6 },    { 1, 10, 5, 7 },   { 3, 12, 6, 8 },  { 4, 13, 7, 8 },
                              { 2, 5, 6, 8 },   { 2, 5, 7, 8 },   { 11, 5, 6, 8 },   { 11, 5, 7, 8 },   { 11, 9, 5, 6 },  { 11, 10, 5, 7 },
                              { 11, 12, 6, 8 }, { 11, 13, 7, 8 }, { 11, 9, 10, 5 },  { 11, 9, 12, 6 },  { 11, 10, 13, 7 }, { 11, 12, 13, 8 } };
   const int tNode[6][4] = { { 4, 5, 7, 3 }, { 0, 3, 1, 4 }, { 4, 1, 5, 3 }, { 5, 6, 7, 2 }, { 1, 3, 2, 5 }, { 7, 2, 3, 5 } };
   const int dx[8] = { 0, 1, 1, 0, 0, 1, 1, 0 }, dy[8] = { 0, 0, 1, 1, 0, 0, 1, 1 }, dz[8] = { 0, 0, 0, 0, 1, 1, 1, 1 };
   std::map< idx, P3 >            nodes;
   std::set< std::vector< idx > > cells;
   for ( const auto& c : cs )
   {
      if ( type == 0 )
      {
         for ( int a = 0; a <= 1; ++a )
            for ( int b = 0; b <= 1; ++b )
               for ( int d = 0; d <= 1; ++d )
                  nodes.emplace( inCube( c[0] + a, c[1] + b, c[2] + d, mx + 1, my + 1 ), P3{ double( c[0] + a ), double( c[1] + b ), double( c[2] + d ) } );
         for ( int t = 0; t < 6; ++t )
         {
            std::vector< idx > cv;
            for ( int q = 0; q < 4; ++q )
               cv.push_back( inCube( c[0] + dx[tNode[t][q]], c[1] + dy[tNode[t][q]], c[2] + dz[tNode[t][q]], mx + 1, my + 1 ) );
            cells.insert( cv );
         }
      }
      else
      {
         for ( int lv = 0; lv < 14; ++lv )
            nodes.emplace( symmId( c[0], c[1], c[2], lv ),
                           P3{ double( c[0] ) + local[lv][0], double( c[1] ) + local[lv][1], double( c[2] ) + local[lv][2] } );
         for ( int t = 0; t < 24; ++t )
         {
            std::vector< idx > cv;
            for ( int q = 0; q < 4; ++q )
               cv.push_back( symmId( c[0], c[1], c[2], cellV[t][q] ) );
            cells.insert( cv );
         }
      }
   }
   return buildMesh( nodes, cells ).release();
}

void* or_mesh_symmetric_cuboid( const double* lo, const double* hi, int nx, int ny, int nz )
{
   return ( nx < 1 || ny < 1 || nz < 1 ) ? nullptr : makeSymmetricCuboid( lo, hi, nx, ny, nz ).release();
}

// sum of the macro-cell volumes
double or_mesh_volume( void* mh )
{
   const auto* m = static_cast< Mesh* >( mh );
   double      v = 0.0;
   for ( const auto& cv : m->Cl )
   {
      P3 x[4] = { m->X[cv[0]], m->X[cv[1]], m->X[cv[2]], m->X[cv[3]] };
      double G[4][3];
      v += std::fabs( baryGrads( x, G ) ) / 6.0;
   }
   return v;
}
void  or_mesh_free( void* m ) { delete static_cast< Mesh* >( m ); }

// V, E, F, C, boundary V, boundary E, boundary F
void or_mesh_info( void* mh, int64_t* out )
{
   auto* m = static_cast< Mesh* >( mh );
   out[0]  = (int64_t) m->X.size();
   out[1]  = (int64_t) m->E.size();
   out[2]  = (int64_t) m->F.size();
   out[3]  = (int64_t) m->Cl.size();
   out[4]  = std::count( m->vB.begin(), m->vB.end(), 1 );
   out[5]  = std::count( m->eB.begin(), m->eB.end(), 1 );
   out[6]  = std::count( m->fB.begin(), m->fB.end(), 1 );
}

// macro-cell vertices (4 per cell, the cell's local vertex order)
void or_mesh_cells( void* mh, int* out )
{
   auto* m = static_cast< Mesh* >( mh );
   for ( size_t c = 0; c < m->Cl.size(); ++c )
      for ( int k = 0; k < 4; ++k )
         out[4 * c + k] = m->Cl[c][k];
}

int64_t or_ndofs( void* mh, int L ) { return numDofs( *static_cast< Mesh* >( mh ), L ); }

// coordinates (3 per DoF) and boundary marker (1 = Dirichlet) in canonical order
void or_dof_coords( void* mh, int L, double* xyz, uint8_t* bnd )
{
   auto&     m = *static_cast< Mesh* >( mh );
   const idx N = idx( 1 ) << L;
   forEachDof< double >( m, L, [&]( int kind, int p, idx i, idx j, idx k, idx n ) {
      P3  x{};
      int b = 0;
      if ( kind == 0 )
      {
         x = m.X[p];
         b = m.vB[p];
      }
      else if ( kind == 1 )
      {
         const P3 &a = m.X[m.E[p][0]], &c = m.X[m.E[p][1]];
         for ( int r = 0; r < 3; ++r )
            x[r] = a[r] + ( c[r] - a[r] ) * double( i ) / double( N );
         b = m.eB[p];
      }
      else if ( kind == 2 )
      {
         const P3 &a = m.X[m.F[p][0]], &c1 = m.X[m.F[p][1]], &c2 = m.X[m.F[p][2]];
         for ( int r = 0; r < 3; ++r )
            x[r] = a[r] + ( c1[r] - a[r] ) * double( i ) / double( N ) + ( c2[r] - a[r] ) * double( j ) / double( N );
         b = m.fB[p];
      }
      else
      {
         x = microCoord( m, p, L, i, j, k );
         b = 0;
      }
      // blended coordinates: the map of the edge's / face's first neighbour cell (IcosahedralShellMap( Edge / Face )
      // takes neighborCells[0]), the cell's own map; macro-vertices keep their positions
      if ( !m.shell.empty() && kind > 0 )
         x = shellF( m.shell[kind == 1 ? m.eC[p][0] : ( kind == 2 ? m.fC[p][0] : p )], x );
      xyz[3 * n + 0] = x[0];
      xyz[3 * n + 1] = x[1];
      xyz[3 * n + 2] = x[2];
      if ( bnd )
         bnd[n] = (uint8_t) b;
   } );
}

// IcosahedralShellMap::setMap on a spherical-shell mesh: every cell gets its map (DoF coordinates become blended)
int or_mesh_set_shell_map( void* mh )
{
   auto& m = *static_cast< Mesh* >( mh );
   try
   {
      std::vector< Mesh::ShellMap > v;
      for ( int c = 0; c < (int) m.Cl.size(); ++c )
         v.push_back( shellMapOf( m, c ) );
      m.shell = v;
   } catch ( const std::exception& )
   {
      return 1;
   }
   return 0;
}

// P1ElementwiseBlendingLaplaceOperator on a mesh with the shell map
void* or_op_new_blending_laplace( void* mh, int minL, int maxL, int prec )
{
   auto* o = new OrOp;
   o->prec = prec;
   o->m    = static_cast< Mesh* >( mh );
   if ( prec == 4 )
      o->f = std::make_unique< P1ElementwiseBlendingLaplace< float > >( o->m, minL, maxL );
   else
      o->d = std::make_unique< P1ElementwiseBlendingLaplace< double > >( o->m, minL, maxL );
   return o;
}

// the shell map of cell c: ref, ray, thr, fo (3 each), r_ref, r_ray
int or_mesh_shell_map( void* mh, int c, double* out14 )
{
   const auto& m = *static_cast< Mesh* >( mh );
   if ( m.shell.empty() )
      return 1;
   const auto& s = m.shell[c];
   for ( int k = 0; k < 3; ++k )
   {
      out14[k]     = s.ref[k];
      out14[3 + k] = s.ray[k];
      out14[6 + k] = s.thr[k];
      out14[9 + k] = s.fo[k];
   }
   out14[12] = s.rRef;
   out14[13] = s.rRay;
   return 0;
}

void* or_func_new( void* mh, int minL, int maxL, int prec )
{
   auto* h = new OrFunc;
   h->prec = prec;
   if ( prec == 4 )
      h->f = std::make_unique< Function< float > >( static_cast< Mesh* >( mh ), minL, maxL );
   else
      h->d = std::make_unique< Function< double > >( static_cast< Mesh* >( mh ), minL, maxL );
   return h;
}
void or_func_free( void* h ) { delete static_cast< OrFunc* >( h ); }

// edge and face vertex lists in storage order
void or_mesh_prims( void* mh, int* edges, int* faces )
{
   const auto& m = *static_cast< Mesh* >( mh );
   for ( size_t e = 0; e < m.E.size(); ++e )
      edges[2 * e] = m.E[e][0], edges[2 * e + 1] = m.E[e][1];
   for ( size_t f = 0; f < m.F.size(); ++f )
      faces[3 * f] = m.F[f][0], faces[3 * f + 1] = m.F[f][1], faces[3 * f + 2] = m.F[f][2];
}

// the reference's lower-dimensional vertex-DoF memory of level L with ghost layers (vertices, edges, faces)
void or_halo_pack( void* fh, int L, double* out )
{
   auto* h = static_cast< OrFunc* >( fh );
   if ( h->prec == 4 )
      haloPack( *h->f, L, out );
   else
      haloPack( *h->d, L, out );
}

void or_func_set( void* fh, int L, const double* canon )
{
   auto* h = static_cast< OrFunc* >( fh );
   if ( h->prec == 4 )
      setCanonical( *h->f, L, canon );
   else
      setCanonical( *h->d, L, canon );
}
void or_func_get( void* fh, int L, double* canon )
{
   auto* h = static_cast< OrFunc* >( fh );
   if ( h->prec == 4 )
      getCanonical( *h->f, L, canon );
   else
      getCanonical( *h->d, L, canon );
}

void* or_op_new( void* mh, int minL, int maxL, int prec )
{
   auto* o = new OrOp;
   o->prec = prec;
   o->m    = static_cast< Mesh* >( mh );
   if ( prec == 4 )
      o->f = std::make_unique< P1ConstantOperator< float > >( o->m, minL, maxL );
   else
      o->d = std::make_unique< P1ConstantOperator< double > >( o->m, minL, maxL );
   return o;
}
void or_op_free( void* o ) { delete static_cast< OrOp* >( o ); }

// P1ElementwiseOperator<p1_div_k_grad_affine_q3> with the coefficient of KField (modes: nmodes x 5 values
// kappa_x, kappa_y, kappa_z, phi, a)
void* or_op_new_divkgrad( void* mh, int minL, int maxL, int prec, int kind, double k0, double alpha, const double* modes,
                          int nmodes )
{
   KField kf;
   kf.kind  = kind;
   kf.k0    = k0;
   kf.alpha = alpha;
   for ( int i = 0; i < nmodes; ++i )
      kf.modes.push_back( { modes[5 * i], modes[5 * i + 1], modes[5 * i + 2], modes[5 * i + 3], modes[5 * i + 4] } );
   auto* o = new OrOp;
   o->prec = prec;
   o->m    = static_cast< Mesh* >( mh );
   if ( prec == 4 )
      o->f = std::make_unique< P1SurrogateDivKGrad< float > >( o->m, minL, maxL, kf );
   else
      o->d = std::make_unique< P1SurrogateDivKGrad< double > >( o->m, minL, maxL, kf );
   return o;
}

// the same operator with the polynomial coefficient k = sum_i coeffs[i] m_i(x) (monomial basis, ncoef terms)
void* or_op_new_divkgrad_poly( void* mh, int minL, int maxL, int prec, const double* coeffs, int ncoef )
{
   KField kf;
   kf.kind = 2;
   kf.poly.assign( coeffs, coeffs + ncoef );
   auto* o = new OrOp;
   o->prec = prec;
   o->m    = static_cast< Mesh* >( mh );
   if ( prec == 4 )
      o->f = std::make_unique< P1SurrogateDivKGrad< float > >( o->m, minL, maxL, kf );
   else
      o->d = std::make_unique< P1SurrogateDivKGrad< double > >( o->m, minL, maxL, kf );
   return o;
}

// P1SurrogateOperator::interpolateStencils( degree, interpolationLevel ) on a div-k-grad operator of this file
int or_op_interpolate_stencils( void* oh, int degree, int interpolationLevel )
{
   auto* o = static_cast< OrOp* >( oh );
   if ( o->prec == 4 )
   {
      auto* s = dynamic_cast< P1SurrogateDivKGrad< float >* >( o->f.get() );
      if ( !s )
         return 1;
      s->interpolateStencils( degree, interpolationLevel );
   }
   else
   {
      auto* s = dynamic_cast< P1SurrogateDivKGrad< double >* >( o->d.get() );
      if ( !s )
         return 1;
      s->interpolateStencils( degree, interpolationLevel );
   }
   return 0;
}

// the fitted polynomials of cell c at level L (NDIR x monomialCount3(degree), direction-major), FP64 operators
int or_op_surrogate_poly( void* oh, int L, int c, double* out )
{
   auto* s = dynamic_cast< P1SurrogateDivKGrad< double >* >( static_cast< OrOp* >( oh )->d.get() );
   if ( !s || !s->active( L ) )
      return 1;
   std::copy( s->poly[L][c].begin(), s->poly[L][c].end(), out );
   return 0;
}

// the variable stencil (element rows) of lattice point (x, y, z) of cell c at level L, FP64 operators
int or_op_variable_stencil( void* oh, int L, int c, long x, long y, long z, double* out15 )
{
   auto* s = dynamic_cast< P1SurrogateDivKGrad< double >* >( static_cast< OrOp* >( oh )->d.get() );
   if ( !s )
      return 1;
   s->variableStencil( c, L, x, y, z, out15 );
   return 0;
}

// P2ConstantOperator< p2_tet_diffusion > on P2 levels minL..maxL; functions used with it hold lattice levels
// minL+1..maxL+1 and every level argument of the operator calls below is a lattice level
// SORSmoothable::smooth_sor in the reference's order (constant-coefficient operators of or_op_new)
int or_smooth_sor_ordered( void* oh, void* dh, void* rh, double relax, int L, int flag, int backwards )
{
   auto* o = static_cast< OrOp* >( oh );
   if ( o->prec == 4 )
   {
      auto* op = dynamic_cast< P1ConstantOperator< float >* >( o->f.get() );
      if ( !op )
         return 1;
      op->smoothSorOrdered( fget< float >( static_cast< OrFunc* >( dh ) ), fget< float >( static_cast< OrFunc* >( rh ) ), relax, L, flag,
                            backwards != 0 );
   }
   else
   {
      auto* op = dynamic_cast< P1ConstantOperator< double >* >( o->d.get() );
      if ( !op )
         return 1;
      op->smoothSorOrdered( fget< double >( static_cast< OrFunc* >( dh ) ), fget< double >( static_cast< OrFunc* >( rh ) ), relax, L,
                            flag, backwards != 0 );
   }
   return 0;
}

void* or_op_new_p2( void* mh, int minL, int maxL, int prec )
{
   auto* o = new OrOp;
   o->prec = prec;
   o->m    = static_cast< Mesh* >( mh );
   if ( prec == 4 )
      o->f = std::make_unique< P2ConstantLaplace< float > >( o->m, minL, maxL );
   else
      o->d = std::make_unique< P2ConstantLaplace< double > >( o->m, minL, maxL );
   return o;
}

// P2ElementwiseOperator< p2_div_k_grad_affine_q4 > with P2KField (kind, k0, alpha, beta, seed)
void* or_op_new_p2_divkgrad( void* mh, int minL, int maxL, int prec, int kind, double k0, double alpha, double beta,
                             uint64_t seed )
{
   P2KField kf;
   kf.kind  = kind;
   kf.k0    = k0;
   kf.alpha = alpha;
   kf.beta  = beta;
   kf.seed  = seed;
   auto* o  = new OrOp;
   o->prec  = prec;
   o->m     = static_cast< Mesh* >( mh );
   if ( prec == 4 )
      o->f = std::make_unique< P2ElementwiseDivKGrad< float > >( o->m, minL, maxL, kf );
   else
      o->d = std::make_unique< P2ElementwiseDivKGrad< double > >( o->m, minL, maxL, kf );
   return o;
}

// value of the P2 coefficient at x for element id eid
double or_p2k_value( int kind, double k0, double alpha, double beta, uint64_t seed, const double* x, uint64_t eid )
{
   P2KField kf;
   kf.kind  = kind;
   kf.k0    = k0;
   kf.alpha = alpha;
   kf.beta  = beta;
   kf.seed  = seed;
   return kf( { x[0], x[1], x[2] }, eid );
}

// the 11-point rule of p2_div_k_grad_affine_q4: barycentric coordinates lam[4 q + v] and weights w[q]
void or_p2_quadrature( double* lam, double* w )
{
   for ( int q = 0; q < 11; ++q )
   {
      for ( int v = 0; v < 4; ++v )
         lam[4 * q + v] = XG4_LAMBDA[q][v];
      w[q] = XG4_W[q];
   }
}

// P2 div-k-grad element matrix (11-point rule) of the tetrahedron x (12 coordinates), coefficient P2KField kind 1
// (polynomial) -> 100 values, row major
void or_p2_divkgrad_element_poly( const double* x, double* out )
{
   P3 xc[4];
   for ( int v = 0; v < 4; ++v )
      xc[v] = { x[3 * v], x[3 * v + 1], x[3 * v + 2] };
   P2KField kf;
   kf.kind = 1;
   double a[10][10];
   p2DivKGradElement( xc, kf, 0, a );
   for ( int i = 0; i < 10; ++i )
      for ( int j = 0; j < 10; ++j )
         out[10 * i + j] = a[i][j];
}

// P2 element matrix of -Laplace on a tetrahedron (12 coordinates) -> 100 values, row major
void or_p2_element( const double* x, double* out )
{
   P3 xc[4];
   for ( int v = 0; v < 4; ++v )
      xc[v] = { x[3 * v], x[3 * v + 1], x[3 * v + 2] };
   double a[10][10];
   p2LaplaceElement( xc, a );
   for ( int i = 0; i < 10; ++i )
      for ( int j = 0; j < 10; ++j )
         out[10 * i + j] = a[i][j];
}

void or_p2_restrict( void* fh, int srcLat, int flag )
{
   auto* h = static_cast< OrFunc* >( fh );
   if ( h->prec == 4 )
      restrictP2( *h->f, srcLat, flag );
   else
      restrictP2( *h->d, srcLat, flag );
}

void or_p2_prolongate( void* fh, int srcLat, int flag, int add )
{
   auto* h = static_cast< OrFunc* >( fh );
   if ( h->prec == 4 )
      prolongateP2( *h->f, srcLat, flag, add != 0 );
   else
      prolongateP2( *h->d, srcLat, flag, add != 0 );
}

// cell stencil (15 weights, C W E N S NW SE TC TW TS TSE BC BN BE BNW) of cell c at level L
void or_op_cell_stencil( void* oh, int L, int c, double* out )
{
   auto* o = static_cast< OrOp* >( oh );
   for ( int k = 0; k < NDIR; ++k )
      out[k] = 0.0;
   if ( o->prec == 4 )
   {
      if ( auto* cop = dynamic_cast< P1ConstantOperator< float >* >( o->f.get() ) )
         for ( int k = 0; k < NDIR; ++k )
            out[k] = cop->cellSt[L][c][k];
   }
   else if ( auto* cop = dynamic_cast< P1ConstantOperator< double >* >( o->d.get() ) )
      for ( int k = 0; k < NDIR; ++k )
         out[k] = cop->cellSt[L][c][k];
}

void or_apply( void* oh, void* src, void* dst, int L, int flag, int add )
{
   auto* o = static_cast< OrOp* >( oh );
   if ( o->prec == 4 )
      o->f->apply( fget< float >( (OrFunc*) src ), fget< float >( (OrFunc*) dst ), L, flag, add != 0 );
   else
      o->d->apply( fget< double >( (OrFunc*) src ), fget< double >( (OrFunc*) dst ), L, flag, add != 0 );
}

void or_inv_diag( void* oh, int L, double* canon )
{
   auto* o = static_cast< OrOp* >( oh );
   if ( o->prec == 4 )
      getCanonical( *o->f->invDiag, L, canon );
   else
      getCanonical( *o->d->invDiag, L, canon );
}

void or_smooth_jac( void* oh, void* dst, void* rhs, void* src, double relax, int L, int flag )
{
   auto* o = static_cast< OrOp* >( oh );
   if ( o->prec == 4 )
   {
      Solvers< float > s( o->m, o->f.get(), L, L );
      s.smoothJac( fget< float >( (OrFunc*) dst ), fget< float >( (OrFunc*) rhs ), fget< float >( (OrFunc*) src ), relax, L, flag );
   }
   else
   {
      Solvers< double > s( o->m, o->d.get(), L, L );
      s.smoothJac( fget< double >( (OrFunc*) dst ), fget< double >( (OrFunc*) rhs ), fget< double >( (OrFunc*) src ), relax, L, flag );
   }
}

void or_restrict( void* fh, int srcL, int flag )
{
   auto* h = static_cast< OrFunc* >( fh );
   if ( h->prec == 4 )
      restrict3D( *h->f, srcL, flag );
   else
      restrict3D( *h->d, srcL, flag );
}

void or_prolongate( void* fh, int srcL, int flag, int add )
{
   auto* h = static_cast< OrFunc* >( fh );
   if ( h->prec == 4 )
      prolongate3D( *h->f, srcL, flag, add != 0 );
   else
      prolongate3D( *h->d, srcL, flag, add != 0 );
}

double or_dot( void* a, void* b, int L, int flag )
{
   auto* ha = static_cast< OrFunc* >( a );
   if ( ha->prec == 4 )
      return dot( *ha->f, fget< float >( (OrFunc*) b ), L, flag );
   return dot( *ha->d, fget< double >( (OrFunc*) b ), L, flag );
}

int or_cheb_coefficients( int order, double lower, double upper, double* out ) { return chebCoefficients( order, lower, upper, out ); }

// estimateSpectralRadiusWithPowerIteration on D^-1 A (ChebyshevSmoother.hpp:255-265, SpectrumEstimation.hpp:57-80);
// x holds the start vector and is overwritten.
double or_power_iteration( void* oh, void* xh, int iters, int L )
{
   auto*              o = static_cast< OrOp* >( oh );
   auto&              A = *o->d;
   Function< double >& x = fget< double >( (OrFunc*) xh );
   Function< double >  aux( o->m, L, L );
   auto               invDiagApply = [&]( const Function< double >& s, Function< double >& d ) {
      A.apply( s, d, L, All, false );
      multElementwise< double >( d, *A.invDiag, d, L, All );
   };
   double norm = std::sqrt( dot( x, x, L, All ) );
   assign< double >( x, { 1.0 / norm }, { &x }, L, All );
   invDiagApply( x, aux );
   double radius = 0.0;
   for ( int it = 1; it <= iters; ++it )
   {
      norm = std::sqrt( dot( aux, aux, L, All ) );
      assign< double >( x, { 1.0 / norm }, { &aux }, L, All );
      invDiagApply( x, aux );
      radius = dot( x, aux, L, All );
   }
   return radius;
}

void or_gmg( void* oh, void* xh, void* bh, const int* ip, const double* dp, int cycles, double* res )
{
   auto* o = static_cast< OrOp* >( oh );
   if ( o->prec == 4 )
      gmgRun< float >( o, o->f.get(), fget< float >( (OrFunc*) xh ), fget< float >( (OrFunc*) bh ), ip, dp, cycles, res );
   else
      gmgRun< double >( o, o->d.get(), fget< double >( (OrFunc*) xh ), fget< double >( (OrFunc*) bh ), ip, dp, cycles, res );
}

// Mixed-precision iterative refinement (mixedPrecisionIterativeRefinement.cpp:162-172): FP64 residual, FP32 V-cycles
// on the correction, FP64 update. opd/ops are FP64/FP32 operators on the same mesh and levels.
void or_ir( void* opd, void* ops, void* xh, void* bh, const int* ip, const double* dp, int outer, int inner, double* res )
{
   auto* od = static_cast< OrOp* >( opd );
   auto* os = static_cast< OrOp* >( ops );
   auto& x  = fget< double >( (OrFunc*) xh );
   auto& b  = fget< double >( (OrFunc*) bh );
   const int L = ip[1];
   Solvers< double > Sd( od->m, od->d.get(), ip[0], ip[1] );
   Solvers< float >  Ss( os->m, os->f.get(), ip[0], ip[1] );
   Ss.pre = ip[2];
   Ss.post = ip[3];
   Ss.inc = ip[4];
   Ss.smoother = ip[5];
   Ss.cgMaxIter = ip[7];
   Ss.relax = dp[0];
   Ss.cgTol = dp[3];
   if ( Ss.smoother == 1 )
   {
      Ss.chebCoeffs.resize( ip[6] );
      chebCoefficients( ip[6], dp[1], dp[2], Ss.chebCoeffs.data() );
   }
   Function< double > r( od->m, ip[0], ip[1] ), t( od->m, ip[0], ip[1] );
   Function< float >  rs( od->m, ip[0], ip[1] ), cs( od->m, ip[0], ip[1] );
   const int flag = Inner | NeumannBoundary;
   res[0] = Sd.residualNorm( x, b, L );
   for ( int it = 0; it < outer; ++it )
   {
      od->d->apply( x, t, L, flag, false );
      assign< double >( r, { 1.0, -1.0 }, { &b, &t }, L, flag );
      copyFrom< float >( rs, r, L );
      interpolate< float >( cs, 0.0, L, All );
      for ( int k = 0; k < inner; ++k )
         Ss.vcycle( cs, rs, L, L );
      copyTo< float >( t, cs, L );
      addScaled< double >( x, 1.0, t, L, flag );
      res[it + 1] = Sd.residualNorm( x, b, L );
   }
}

} // extern "C"
