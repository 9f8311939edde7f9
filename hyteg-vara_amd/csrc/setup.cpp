// Host-side setup: topology tables, per-level work lists and stencil assembly.
//
// Stencils. The P1 Laplace stiffness of a uniformly refined affine macro-cell is a sum over
// the micro-tetrahedra around a micro-vertex (src/hyteg/p1functionspace/P1Elements.hpp:89-143,
// 463-527). All micro-tetrahedra of one of the six types (CellDoFIndexing.hpp:155-198) are
// translates of each other, so one 4x4 element matrix per type and cell suffices. The element
// matrix is K_ij = vol * grad(l_i) . grad(l_j) with grad(l_i) = n_i / (n_i . (P_i - P_a)),
// n_i a normal of the face opposite P_i and P_a one of its vertices. The stencil of a point
// in a cell is the sum, over the (type, corner) pairs whose tetrahedron lies inside the cell,
// of the corresponding element-matrix row; at a macro-face, -edge or -vertex only the
// micro-tetrahedra inside the cell contribute (one-sided stencil), and the full row is the sum
// over the incident cells (P1Operator.hpp:1792-2161).
#include "hyg_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>

namespace hyg {

namespace {

using V3 = std::array< double, 3 >;
V3     sub( const V3& a, const V3& b ) { return { a[0] - b[0], a[1] - b[1], a[2] - b[2] }; }
V3     cross( const V3& a, const V3& b ) { return { a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0] }; }
double dot3( const V3& a, const V3& b ) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// lattice offsets of the vertices of the six micro-tetrahedron types
const int kMicroTet[6][4][3] = { { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } },   // white up
                                 { { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 }, { 1, 0, 1 } },   // blue up
                                 { { 1, 0, 0 }, { 0, 1, 0 }, { 1, 0, 1 }, { 0, 0, 1 } },   // green up
                                 { { 1, 1, 0 }, { 1, 1, 1 }, { 0, 1, 1 }, { 1, 0, 1 } },   // white down
                                 { { 1, 0, 1 }, { 0, 1, 1 }, { 0, 0, 1 }, { 0, 1, 0 } },   // blue down
                                 { { 0, 1, 0 }, { 1, 1, 0 }, { 1, 0, 1 }, { 0, 1, 1 } } }; // green down

int dirIndex( int dx, int dy, int dz )
{
   for ( int d = 0; d < NDIR; ++d )
      if ( dirX( d ) == dx && dirY( d ) == dy && dirZ( d ) == dz )
         return d;
   throw std::logic_error( "micro-tetrahedron edge is not a stencil direction" );
}

} // namespace

// element matrices of the six micro-tetrahedron types of a cell with vertices cv (local order) at level L
void elementMatrices( const Mesh& m, const std::array< int, 4 >& cv, int L, double K[6][4][4] )
{
   const double N  = double( 1 << L );
   const V3&    P0 = m.X[cv[0]];
   const V3     a1 = sub( m.X[cv[1]], P0 ), a2 = sub( m.X[cv[2]], P0 ), a3 = sub( m.X[cv[3]], P0 );
   for ( int t = 0; t < 6; ++t )
   {
      V3 P[4];
      for ( int k = 0; k < 4; ++k )
         for ( int r = 0; r < 3; ++r )
            P[k][r] = ( kMicroTet[t][k][0] * a1[r] + kMicroTet[t][k][1] * a2[r] + kMicroTet[t][k][2] * a3[r] ) / N;
      const double vol = std::fabs( dot3( sub( P[1], P[0] ), cross( sub( P[2], P[0] ), sub( P[3], P[0] ) ) ) ) / 6.0;
      V3           g[4];
      for ( int i = 0; i < 4; ++i )
      {
         const int a = ( i + 1 ) % 4, b = ( i + 2 ) % 4, d = ( i + 3 ) % 4;
         const V3  n = cross( sub( P[b], P[a] ), sub( P[d], P[a] ) );
         const double s = dot3( n, sub( P[i], P[a] ) );
         for ( int r = 0; r < 3; ++r )
            g[i][r] = n[r] / s;
      }
      for ( int i = 0; i < 4; ++i )
         for ( int j = 0; j < 4; ++j )
            K[t][i][j] = vol * dot3( g[i], g[j] );
   }
}

// one-sided stencil of cell c at lattice point (qx, qy, qz)
void stencilAt( const double K[6][4][4], int N, int qx, int qy, int qz, double* w )
{
   for ( int d = 0; d < NDIR; ++d )
      w[d] = 0.0;
   for ( int t = 0; t < 6; ++t )
      for ( int k = 0; k < 4; ++k )
      {
         // tetrahedron of type t whose corner k sits at q
         bool inside = true;
         for ( int j = 0; j < 4; ++j )
         {
            const int x = qx - kMicroTet[t][k][0] + kMicroTet[t][j][0];
            const int y = qy - kMicroTet[t][k][1] + kMicroTet[t][j][1];
            const int z = qz - kMicroTet[t][k][2] + kMicroTet[t][j][2];
            inside      = inside && inCell( N, x, y, z );
         }
         if ( !inside )
            continue;
         for ( int j = 0; j < 4; ++j )
            w[dirIndex( kMicroTet[t][j][0] - kMicroTet[t][k][0], kMicroTet[t][j][1] - kMicroTet[t][k][1],
                        kMicroTet[t][j][2] - kMicroTet[t][k][2] )] += K[t][k][j];
      }
}

namespace {

void entryPosHost( const LowerEntry& e, int N, int i, int j, int& qx, int& qy, int& qz )
{
   auto u = []( int l, int r ) { return l == r + 1 ? 1 : 0; };
   qx     = N * u( e.l0, 0 ) + i * ( u( e.l1, 0 ) - u( e.l0, 0 ) ) + j * ( u( e.l2, 0 ) - u( e.l0, 0 ) );
   qy     = N * u( e.l0, 1 ) + i * ( u( e.l1, 1 ) - u( e.l0, 1 ) ) + j * ( u( e.l2, 1 ) - u( e.l0, 1 ) );
   qz     = N * u( e.l0, 2 ) + i * ( u( e.l1, 2 ) - u( e.l0, 2 ) ) + j * ( u( e.l2, 2 ) - u( e.l0, 2 ) );
}

template < typename T >
T* upload( const std::vector< T >& h )
{
   if ( h.empty() )
      return nullptr;
   void* d = nullptr;
   if ( hipMalloc( &d, h.size() * sizeof( T ) ) != hipSuccess )
      throw std::runtime_error( "hipMalloc failed for setup table" );
   if ( hipMemcpy( d, h.data(), h.size() * sizeof( T ), hipMemcpyHostToDevice ) != hipSuccess )
      throw std::runtime_error( "hipMemcpy failed for setup table" );
   return static_cast< T* >( d );
}

} // namespace

void primEntriesPos( const LowerEntry& e, int N, int i, int j, int& qx, int& qy, int& qz ) { entryPosHost( e, N, i, j, qx, qy, qz ); }

// Topology tables: LowerPrim order = vertices, edges, faces; CellInfo per cell.
void MeshCtx::buildTopology()
{
   const Mesh& m = *mesh;
   auto        localOf = [&]( int c, int v ) {
      for ( int l = 0; l < 4; ++l )
         if ( m.Cv[c][l] == v )
            return l;
      throw std::logic_error( "vertex not in cell" );
   };
   auto addPrim = [&]( int kind, int bnd, const std::vector< int >& cells, const int* pv, int nv ) {
      LowerPrim p;
      p.kind   = kind;
      p.bc     = bcOfFlag( bnd ); // create0123BC: mesh flag -> DoF type
      p.nCells = (int) cells.size();
      p.entry0 = (int) entriesHost.size();
      primsHost.push_back( p );
      for ( int c : cells )
      {
         LowerEntry e;
         e.cell = c;
         e.l0   = localOf( c, pv[0] );
         e.l1   = nv > 1 ? localOf( c, pv[1] ) : e.l0;
         e.l2   = nv > 2 ? localOf( c, pv[2] ) : e.l0;
         entriesHost.push_back( e );
      }
   };
   for ( int v = 0; v < m.nV(); ++v )
      addPrim( 0, m.vB[v], m.vC[v], &v, 1 );
   for ( int e = 0; e < m.nE(); ++e )
      addPrim( 1, m.eB[e], m.eC[e], m.E[e].data(), 2 );
   for ( int f = 0; f < m.nF(); ++f )
      addPrim( 2, m.fB[f], m.fC[f], m.F[f].data(), 3 );

   // Multi-colour SOR. A lattice point with barycentric lattice coordinates b_i (sum N) over the vertices v_i of a
   // cell containing it gets the colour (sum_i c(v_i) b_i) mod nColours, a value that depends only on the point.
   // The 15-point stencil couples points that differ by e_i - e_j (i != j) or, along the cell's inner diagonal
   // (x, y, z) +- (1, -1, 1), by e_1 - e_2 + e_3 - e_0 in the cell's local vertex order, so coupled points never
   // share a colour if the vertex colours differ along every macro-edge and c_1 + c_3 != c_0 + c_2 (mod
   // nColours) in every cell. The vertex colours are chosen greedily in index order (smallest admissible colour)
   // for the smallest nColours >= 4 for which that succeeds.
   {
      std::vector< std::vector< int > > nb( m.nV() );
      for ( const auto& e : m.E )
      {
         nb[e[0]].push_back( e[1] );
         nb[e[1]].push_back( e[0] );
      }
      std::vector< int > col;
      for ( nColours = 4; nColours <= 64; ++nColours )
      {
         col.assign( m.nV(), -1 );
         bool ok = true;
         for ( int v = 0; v < m.nV() && ok; ++v )
         {
            ok = false;
            for ( int k = 0; k < nColours && !ok; ++k )
            {
               bool good = true;
               for ( int u : nb[v] )
                  good = good && col[u] != k;
               col[v] = k;
               for ( int c : m.vC[v] )
               {
                  const auto& cv = m.Cv[c];
                  if ( good && col[cv[0]] >= 0 && col[cv[1]] >= 0 && col[cv[2]] >= 0 && col[cv[3]] >= 0 )
                     good = ( ( col[cv[1]] + col[cv[3]] - col[cv[0]] - col[cv[2]] ) % nColours + nColours ) % nColours != 0;
               }
               ok = good;
               if ( !ok )
                  col[v] = -1;
            }
         }
         if ( ok )
            break;
      }
      if ( nColours > 64 )
         throw std::runtime_error( "no SOR colouring with at most 64 colours" );
      vertexColours = col;
   }
   cellInfoHost.resize( m.nC() );
   for ( int c = 0; c < m.nC(); ++c )
   {
      CellInfo& ci = cellInfoHost[c];
      for ( int k = 0; k < 4; ++k )
         ci.colour[k] = vertexColours[m.Cv[c][k]];
      ci.bcword    = 1; // slot 0: cells are always inner
      ci.ownmask   = 1;
      ci.prim[0]   = -1;
      ci.invNC[0]  = 1.0;
      ci.invNC[15] = 0.0;
      for ( int k = 0; k < 15; ++k )
         ci.entry[k] = -1;
      auto setSlot = [&]( int slot, int prim, int bnd, const std::vector< int >& cells ) {
         ci.bcword |= uint64_t( bcOfFlag( bnd ) ) << ( 4 * slot );
         if ( cells.front() == c )
            ci.ownmask |= 1u << slot;
         ci.prim[slot]  = prim;
         ci.invNC[slot] = 1.0 / double( cells.size() );
      };
      for ( int k = 0; k < 4; ++k )
      {
         const int f = m.cF[c][k];
         setSlot( 1 + k, m.nV() + m.nE() + f, m.fB[f], m.fC[f] );
      }
      for ( int k = 0; k < 6; ++k )
      {
         const int e = m.cE[c][k];
         setSlot( 5 + k, m.nV() + e, m.eB[e], m.eC[e] );
      }
      for ( int k = 0; k < 4; ++k )
      {
         const int v = m.Cv[c][k];
         setSlot( 11 + k, v, m.vB[v], m.vC[v] );
      }
   }
   // entry of every (primitive, cell) pair in the cell's slot table
   for ( size_t p = 0; p < primsHost.size(); ++p )
   {
      const LowerPrim& pr = primsHost[p];
      for ( int k = 0; k < pr.nCells; ++k )
      {
         const LowerEntry& e    = entriesHost[pr.entry0 + k];
         CellInfo&         ci   = cellInfoHost[e.cell];
         int               slot = -1;
         if ( pr.kind == 2 )
            slot = 1 + entrySlot( e );
         else if ( pr.kind == 0 )
            slot = 11 + e.l0;
         else
         {
            // local edge of the vertex pair (l0, l1): e0..e5 = (0,1) (0,2) (1,2) (0,3) (1,3) (2,3)
            const int a = std::min( e.l0, e.l1 ), b = std::max( e.l0, e.l1 );
            const int le = a == 0 ? ( b == 1 ? 0 : ( b == 2 ? 1 : 3 ) ) : ( a == 1 ? ( b == 2 ? 2 : 4 ) : 5 );
            slot         = 5 + le;
         }
         ci.entry[slot] = pr.entry0 + k;
      }
   }
   // face slots of every cell (layout order), with the entry of the (face, cell) pair
   cellFacesHost.assign( size_t( m.nC() ) * 4, CellFaceSlot{ -1, -1, -1, 0 } );
   for ( int f = 0; f < m.nF(); ++f )
   {
      const int       p  = m.nV() + m.nE() + f;
      const LowerPrim& pr = primsHost[p];
      for ( int k = 0; k < pr.nCells; ++k )
      {
         const LowerEntry& e = entriesHost[pr.entry0 + k];
         CellFaceSlot&     s = cellFacesHost[4 * e.cell + entrySlot( e )];
         s.entry             = pr.entry0 + k;
         s.prim              = p;
         s.block             = 2 * f + k;
         s.bc                = pr.bc;
      }
   }
   prims     = upload( primsHost );
   entries   = upload( entriesHost );
   cellInfo  = upload( cellInfoHost );
   cellFaces = upload( cellFacesHost );
}

LevelGeom& MeshCtx::geom( int L )
{
   if ( (int) levels.size() <= L )
      levels.resize( L + 1 );
   LevelGeom& g = levels[L];
   if ( g.L == L )
      return g;
   const int N = 1 << L, w = N + 1;
   g.stride    = cellStride( L );
   g.tilesHost.clear();
   g.interiorTile0  = -1;
   g.nInteriorTiles = 0;
   for ( int z = 0; z <= N; ++z )
   {
      const int W = w - z, np = W * ( W + 1 ) / 2;
      for ( int s0 = 0; s0 < np; s0 += 1024 ) // 256 threads x 4 points
      {
         if ( z >= 1 && z <= N - 3 )
         {
            if ( g.interiorTile0 < 0 )
               g.interiorTile0 = (int) g.tilesHost.size();
            ++g.nInteriorTiles;
         }
         g.tilesHost.push_back( { z, s0 } );
      }
   }
   if ( g.interiorTile0 < 0 )
      g.interiorTile0 = 0;
   g.nTiles = (int) g.tilesHost.size();
   // z-chunk major, then segment column, then row group: the four waves of a block march adjacent rows.
   // A segment of rows y.. starting at column x0 covers the planes where row y still has a point at
   // x >= x0 (z <= N - y - x0): the last of them holds no interior output, only the row end that lies
   // on the diagonal face.
   // HYG_ZCHUNK (planes per segment) and HYG_SEG_ORDER (1: columns vary fastest) override the
   // defaults for experiments.
   const char* ec    = std::getenv( "HYG_ZCHUNK" );
   const char* eo    = std::getenv( "HYG_SEG_ORDER" );
   const int   chunk = ec ? std::max( 4, std::min( 1024, std::atoi( ec ) ) ) : kZSegChunk;
   const bool  xFast = eo && std::atoi( eo ) == 1;
   for ( int k = 0; k < kNumCellCfgs; ++k )
   {
      const int R = kCellRows[k], WO = 64 * kCellPoints[k] - 2;
      auto&     v = g.zSegsHost[k];
      v.clear();
      auto add = [&]( int y, int x0, int z0 ) {
         const int zEnd = N - y - x0 + 1;
         if ( z0 < zEnd )
            v.push_back( { y, x0, z0, std::min( z0 + chunk, zEnd ) } );
      };
      for ( int z0 = 1; z0 <= N - 2; z0 += chunk )
         if ( xFast )
         {
            for ( int y = 1; y <= N - 2; y += R )
               for ( int x0 = 1; x0 <= N - 2; x0 += WO )
                  add( y, x0, z0 );
         }
         else
         {
            for ( int x0 = 1; x0 <= N - 2; x0 += WO )
               for ( int y = 1; y <= N - 2; y += R )
                  add( y, x0, z0 );
         }
      g.nZSegs[k] = (int) v.size();
      g.zSegs[k]  = upload( v );
   }
   // P2 row pairs (y0, y0+1), y0 even: segment k holds the x values 124k .. 124k+123 (k = 0 from x = 1); the
   // pair has interior outputs while z <= N - 1 - x_first - max(y0, 1). Planes are chunked from z = 0 so every
   // segment starts on an even plane (plane 0 itself has no interior output).
   // Coarse lattices take shorter segments: more waves, fewer serial plane steps each.
   for ( int ci = 0; ci < 4; ++ci )
   {
      const int p2Chunk = kP2Chunks[ci];
      auto&     v       = g.p2ZSegsHost[ci];
      v.clear();
      for ( int z0 = 0; z0 <= N - 3; z0 += p2Chunk )
         for ( int k = 0; kP2ZSeg * k <= N - 3; ++k )
            for ( int y0 = 0; y0 <= N - 3; y0 += 2 )
            {
               const int xs = k == 0 ? 1 : kP2ZSeg * k, zEnd = N - xs - std::max( y0, 1 );
               if ( zEnd > std::max( z0, 1 ) )
                  v.push_back( { y0, k, z0, std::min( z0 + p2Chunk, zEnd ) } );
            }
      g.nP2ZSegs[ci] = (int) v.size();
      g.p2ZSegs[ci]  = upload( v );
   }
   // z = 0 and y = 0 faces of every cell: rows of 62 points
   g.flatItemsHost.clear();
   for ( int c = 0; c < mesh->nC(); ++c )
      for ( int slot = 0; slot < 2; ++slot )
         for ( int line = 1; line <= N - 2; ++line )
            for ( int x0 = 1; x0 <= N - 1 - line; x0 += 62 )
               g.flatItemsHost.push_back( int4{ c, slot, line, x0 } );
   g.nFlatItems = (int) g.flatItemsHost.size();
   g.flatItems  = upload( g.flatItemsHost );
   g.chunksHost.clear();
   for ( int p = 0; p < (int) primsHost.size(); ++p )
   {
      const int kind = primsHost[p].kind;
      if ( kind == 0 )
         g.chunksHost.push_back( { p, 0, 0, 1 } );
      else if ( kind == 1 )
      {
         for ( int i0 = 1; i0 <= N - 1; i0 += 64 )
            g.chunksHost.push_back( { p, 0, i0, std::min( 64, N - i0 ) } );
      }
      else
      {
         for ( int j = 1; j <= N - 2; ++j )
            for ( int i0 = 1; i0 <= N - 1 - j; i0 += 64 )
               g.chunksHost.push_back( { p, j, i0, std::min( 64, N - j - i0 ) } );
      }
   }
   g.nChunks        = (int) g.chunksHost.size();
   g.firstFaceChunk = g.nChunks;
   for ( int k = 0; k < g.nChunks; ++k )
      if ( primsHost[g.chunksHost[k].prim].kind == 2 )
      {
         g.firstFaceChunk = k;
         break;
      }
   g.tiles   = upload( g.tilesHost );
   g.chunks  = upload( g.chunksHost );
   g.L       = L;
   return g;
}

MeshCtx::~MeshCtx()
{
   for ( auto& g : levels )
   {
      (void) hipFree( g.tiles );
      for ( auto* zs : g.zSegs )
         (void) hipFree( zs );
      (void) hipFree( g.flatItems );
      for ( auto* zs : g.p2ZSegs )
         (void) hipFree( zs );
      (void) hipFree( g.chunks );
   }
   (void) hipFree( prims );
   (void) hipFree( shell );
   (void) hipFree( p2tab );
   (void) hipFree( cellFaces );
   (void) hipFree( entries );
   (void) hipFree( cellInfo );
}

// Assemble the per-level stencil tables of the constant P1 Laplace operator.
void assembleP1Const( hyg_op_s& op )
{
   const MeshCtx& mc = op.mesh->ctx;
   const Mesh&    m  = *mc.mesh;
   op.levels.resize( op.maxL + 1 );
   for ( int L = op.minL; L <= op.maxL; ++L )
   {
      hyg_op_s::Level& ol = op.levels[L];
      const int        N  = 1 << L;
      ol.cellStHost.assign( size_t( m.nC() ) * 16, 0.0 );
      ol.entryWHost.assign( mc.entriesHost.size() * 16, 0.0 );
      ol.primInvHost.assign( mc.primsHost.size(), 0.0 );
      std::vector< std::array< std::array< std::array< double, 4 >, 4 >, 6 > > Ks( m.nC() );
      for ( int c = 0; c < m.nC(); ++c )
      {
         double K[6][4][4];
         elementMatrices( m, m.Cv[c], L, K );
         for ( int t = 0; t < 6; ++t )
            for ( int i = 0; i < 4; ++i )
               for ( int j = 0; j < 4; ++j )
                  Ks[c][t][i][j] = K[t][i][j];
         if ( L >= 2 )
         {
            double* w = &ol.cellStHost[16 * c];
            stencilAt( K, N, 1, 1, 1, w );
            w[15] = 1.0 / w[0];
         }
      }
      for ( size_t p = 0; p < mc.primsHost.size(); ++p )
      {
         const LowerPrim& pr = mc.primsHost[p];
         // representative interior point of the primitive at this level
         int i = 0, j = 0;
         if ( pr.kind == 1 )
         {
            if ( L < 1 )
               continue;
            i = 1;
         }
         else if ( pr.kind == 2 )
         {
            if ( L < 2 )
               continue;
            i = j = 1;
         }
         double centre = 0.0;
         for ( int k = 0; k < pr.nCells; ++k )
         {
            const LowerEntry& e = mc.entriesHost[pr.entry0 + k];
            double            K[6][4][4];
            for ( int t = 0; t < 6; ++t )
               for ( int a = 0; a < 4; ++a )
                  for ( int b = 0; b < 4; ++b )
                     K[t][a][b] = Ks[e.cell][t][a][b];
            int qx, qy, qz;
            entryPosHost( e, N, i, j, qx, qy, qz );
            double* w = &ol.entryWHost[16 * size_t( pr.entry0 + k )];
            stencilAt( K, N, qx, qy, qz, w );
            centre += w[0];
         }
         ol.primInvHost[p] = 1.0 / centre;
      }
      std::vector< float > f1( ol.cellStHost.begin(), ol.cellStHost.end() );
      std::vector< float > f2( ol.entryWHost.begin(), ol.entryWHost.end() );
      std::vector< float > f3( ol.primInvHost.begin(), ol.primInvHost.end() );
      ol.cellSt   = upload( ol.cellStHost );
      ol.entryW   = upload( ol.entryWHost );
      ol.primInv  = upload( ol.primInvHost );
      ol.cellStF  = upload( f1 );
      ol.entryWF  = upload( f2 );
      ol.primInvF = upload( f3 );
      // face partial-sum blocks (allocated here so that no stencil call allocates)
      const size_t np = size_t( m.nF() ) * 2 * size_t( N + 1 ) * size_t( N + 1 );
      if ( np > 0 )
      {
         void* d64 = nullptr;
         void* d32 = nullptr;
         if ( hipMalloc( &d64, np * sizeof( double ) ) != hipSuccess || hipMalloc( &d32, np * sizeof( float ) ) != hipSuccess )
         {
            (void) hipFree( d64 );
            throw std::runtime_error( "hipMalloc failed for face partial sums" );
         }
         ol.facePart  = static_cast< double* >( d64 );
         ol.facePartF = static_cast< float* >( d32 );
      }
   }
}

// Variable-coefficient elementwise operator: per level the geometry of every cell (affine map and the
// element matrices of the six micro-cell types for k = 1), the buffers of the per-element coefficient
// integrals and of the one-sided face/edge/vertex sums, and the inverse diagonal, computed on the device.
void assembleP1DivKGrad( hyg_op_s& op )
{
   const MeshCtx& mc = op.mesh->ctx;
   const Mesh&    m  = *mc.mesh;
   op.levels.resize( op.maxL + 1 );
   auto alloc = []( size_t bytes ) {
      void* d = nullptr;
      if ( bytes > 0 && hipMalloc( &d, bytes ) != hipSuccess )
         throw std::runtime_error( "hipMalloc failed for the elementwise operator" );
      return d;
   };
   if ( !op.modesHost.empty() )
      op.modes = upload( op.modesHost );
   op.nModes = int( op.modesHost.size() / 5 );
   if ( op.kPolyDeg >= 0 )
      op.kPoly = upload( op.kPolyHost );
   for ( int L = op.minL; L <= op.maxL; ++L )
   {
      hyg_op_s::Level& ol = op.levels[L];
      const int        N  = 1 << L;
      std::vector< ElemCell > ec( m.nC() );
      for ( int c = 0; c < m.nC(); ++c )
      {
         const auto& cv = m.Cv[c];
         for ( int r = 0; r < 3; ++r )
         {
            ec[c].v0[r] = m.X[cv[0]][r];
            for ( int k = 0; k < 3; ++k )
               ec[c].a[r][k] = ( m.X[cv[k + 1]][r] - m.X[cv[0]][r] ) / double( N );
         }
         double K[6][4][4];
         elementMatrices( m, cv, L, K );
         for ( int t = 0; t < 6; ++t )
            for ( int i = 0; i < 4; ++i )
               for ( int j = 0; j < 4; ++j )
                  ec[c].mt[t][4 * i + j] = K[t][i][j];
      }
      ol.elem                = upload( ec );
      const size_t nKb       = size_t( m.nC() ) * 6 * size_t( S3( N ) );
      const size_t nLow      = mc.entriesHost.size() * size_t( N + 1 );
      const size_t nFace     = size_t( m.nF() ) * 2 * size_t( N + 1 ) * size_t( N + 1 );
      const size_t nFunc     = size_t( m.nC() ) * size_t( cellStride( L ) ) + size_t( kTailSlack );
      ol.kb        = static_cast< double* >( alloc( nKb * sizeof( double ) ) );
      ol.kbF       = static_cast< float* >( alloc( nKb * sizeof( float ) ) );
      ol.lowPart   = static_cast< double* >( alloc( nLow * sizeof( double ) ) );
      ol.lowPartF  = static_cast< float* >( alloc( nLow * sizeof( float ) ) );
      ol.facePart  = static_cast< double* >( alloc( nFace * sizeof( double ) ) );
      ol.facePartF = static_cast< float* >( alloc( nFace * sizeof( float ) ) );
      ol.dinv      = static_cast< double* >( alloc( nFunc * sizeof( double ) ) );
      ol.dinvF     = static_cast< float* >( alloc( nFunc * sizeof( float ) ) );
      if ( op.nModes > 0 )
      {
         const size_t nm = size_t( m.nC() ) * size_t( op.nModes );
         ol.kbPhase      = static_cast< double* >( alloc( nm * 5 * sizeof( double ) ) );
         ol.kbTab        = static_cast< double* >( alloc( nm * 60 * sizeof( double ) ) );
         ol.kbTabF       = static_cast< float* >( alloc( nm * 60 * sizeof( float ) ) );
      }
   }
}

// ------------------------------------------------------------------ P2 (quadratic) path
namespace {
// local P2 nodes: vertices 0..3, then the edges (2,3) (1,3) (1,2) (0,3) (0,2) (0,1)
const int kP2Edge[6][2] = { { 2, 3 }, { 1, 3 }, { 1, 2 }, { 0, 3 }, { 0, 2 }, { 0, 1 } };

// basis and quarter-point barycentric coordinates: the compile-time versions the transfer kernels use
void p2Basis( const double lam[4], double phi[10] )
{
   p2BasisAt( lam, phi );
}
bool quarterBary( int t, const int o[3], const int r[3], int l4[4] )
{
   return p2QuarterBary( t, o, r, l4 );
}
} // namespace

void buildP2Tables( P2Tables& tb )
{
   tb = P2Tables{};
   for ( int t = 0; t < 6; ++t )
   {
      for ( int v = 0; v < 4; ++v )
         for ( int c = 0; c < 3; ++c )
            tb.nd[t][v][c] = 2 * kMicroTet[t][v][c];
      for ( int e = 0; e < 6; ++e )
         for ( int c = 0; c < 3; ++c )
            tb.nd[t][4 + e][c] = kMicroTet[t][kP2Edge[e][0]][c] + kMicroTet[t][kP2Edge[e][1]][c];
      for ( int l = 0; l < 10; ++l )
      {
         const int cls = ( tb.nd[t][l][0] & 1 ) | ( ( tb.nd[t][l][1] & 1 ) << 1 ) | ( ( tb.nd[t][l][2] & 1 ) << 2 );
         tb.ent[cls][tb.nEnt[cls]++] = 16 * t + l;
      }
   }
   // prolongation: micro-cells containing each residue of the fine lattice modulo 4
   for ( int rz = 0; rz < 4; ++rz )
      for ( int ry = 0; ry < 4; ++ry )
         for ( int rx = 0; rx < 4; ++rx )
         {
            const int r[3] = { rx, ry, rz }, ri = rx + 4 * ry + 16 * rz;
            if ( ( ( rx | ry | rz ) & 1 ) == 0 )
               continue; // points of the coarse lattice take the coarse value itself
            for ( int t = 0; t < 6; ++t )
               for ( int o = 0; o < 8; ++o )
               {
                  const int oo[3] = { o & 1, ( o >> 1 ) & 1, ( o >> 2 ) & 1 };
                  int       l4[4];
                  if ( !quarterBary( t, oo, r, l4 ) )
                     continue;
                  if ( tb.nCand[ri] == 6 )
                     throw std::logic_error( "more than six micro-cells contain a lattice point" );
                  P2Tables::Cand& cd = tb.cand[ri][tb.nCand[ri]++];
                  cd.t               = t;
                  cd.o               = o;
                  const double lam[4] = { l4[0] / 4.0, l4[1] / 4.0, l4[2] / 4.0, l4[3] / 4.0 };
                  p2Basis( lam, cd.phi );
               }
         }
   // restriction: support of the basis function of a coarse point of each parity class (point n = 8 + cls,
   // away from any boundary), on the fine lattice: micro-cell (t, a) with n as node l, its 35 quarter points
   for ( int cls = 0; cls < 8; ++cls )
   {
      const int n[3] = { 8 + ( cls & 1 ), 8 + ( ( cls >> 1 ) & 1 ), 8 + ( ( cls >> 2 ) & 1 ) };
      for ( int k = 0; k < tb.nEnt[cls]; ++k )
      {
         const int t = tb.ent[cls][k] >> 4, l = tb.ent[cls][k] & 15;
         int       a[3];
         for ( int c = 0; c < 3; ++c )
            a[c] = ( n[c] - tb.nd[t][l][c] ) / 2;
         for ( int i = 0; i <= 4; ++i )
            for ( int j = 0; i + j <= 4; ++j )
               for ( int q = 0; i + j + q <= 4; ++q )
               {
                  const int    l4[4]  = { 4 - i - j - q, i, j, q };
                  const double lam[4] = { l4[0] / 4.0, l4[1] / 4.0, l4[2] / 4.0, l4[3] / 4.0 };
                  double       phi[10];
                  p2Basis( lam, phi );
                  if ( phi[l] == 0.0 )
                     continue;
                  int d[3];
                  for ( int c = 0; c < 3; ++c )
                  {
                     int p = 4 * a[c];
                     for ( int v = 0; v < 4; ++v )
                        p += l4[v] * kMicroTet[t][v][c];
                     d[c] = p - 2 * n[c];
                  }
                  int m = 0;
                  while ( m < tb.nSup[cls] && !( tb.sup[cls][m].d[0] == d[0] && tb.sup[cls][m].d[1] == d[1] && tb.sup[cls][m].d[2] == d[2] ) )
                     ++m;
                  if ( m < tb.nSup[cls] )
                  {
                     if ( tb.sup[cls][m].w != phi[l] )
                        throw std::logic_error( "inconsistent P2 interpolation weight" );
                     continue;
                  }
                  if ( tb.nSup[cls] == 128 )
                     throw std::logic_error( "P2 restriction support exceeds its table" );
                  P2Tables::Sup& sp = tb.sup[cls][tb.nSup[cls]++];
                  sp.d[0] = d[0];
                  sp.d[1] = d[1];
                  sp.d[2] = d[2];
                  sp.w    = phi[l];
               }
      }
   }
}

const P2Tables* MeshCtx::p2Tables()
{
   if ( !p2tab )
   {
      std::vector< P2Tables > tb( 1 );
      buildP2Tables( tb[0] );
      p2tab = upload( tb );
   }
   return p2tab;
}

// P2ConstantOperator< p2_tet_diffusion >: per lattice level the element matrices of the six micro-cell types
// of every cell (micro-cells of level lattice-1), with the exact integrals of the products of the P2 basis
// gradients (grad phi_i = sum_k lambda_k a_ik + b_i, int lambda_k lambda_l = vol (1 + delta_kl) / 20,
// int lambda_k = vol / 4), the partial-sum buffers and the inverse diagonal (computed on the device).
void assembleP2Const( hyg_op_s& op )
{
   const MeshCtx& mc = op.mesh->ctx;
   const Mesh&    m  = *mc.mesh;
   op.levels.resize( op.maxL + 1 );
   auto alloc = []( size_t bytes ) {
      void* d = nullptr;
      if ( bytes > 0 && hipMalloc( &d, bytes ) != hipSuccess )
         throw std::runtime_error( "hipMalloc failed for the P2 operator" );
      return d;
   };
   for ( int L = op.minL; L <= op.maxL; ++L )
   {
      hyg_op_s::Level& ol = op.levels[L];
      const int        N  = 1 << L;
      const double     Ne = double( N / 2 ); // micro-cells of the P2 elements: level L-1
      std::vector< double > el( size_t( m.nC() ) * 600 );
      for ( int c = 0; c < m.nC(); ++c )
      {
         const auto& cv = m.Cv[c];
         const V3&   P0 = m.X[cv[0]];
         const V3    a1 = sub( m.X[cv[1]], P0 ), a2 = sub( m.X[cv[2]], P0 ), a3 = sub( m.X[cv[3]], P0 );
         for ( int t = 0; t < 6; ++t )
         {
            V3 P[4];
            for ( int k = 0; k < 4; ++k )
               for ( int r = 0; r < 3; ++r )
                  P[k][r] = ( kMicroTet[t][k][0] * a1[r] + kMicroTet[t][k][1] * a2[r] + kMicroTet[t][k][2] * a3[r] ) / Ne;
            const double vol = std::fabs( dot3( sub( P[1], P[0] ), cross( sub( P[2], P[0] ), sub( P[3], P[0] ) ) ) ) / 6.0;
            V3           g[4];
            for ( int i = 0; i < 4; ++i )
            {
               const int    a = ( i + 1 ) % 4, b = ( i + 2 ) % 4, d = ( i + 3 ) % 4;
               const V3     nrm = cross( sub( P[b], P[a] ), sub( P[d], P[a] ) );
               const double sc  = dot3( nrm, sub( P[i], P[a] ) );
               for ( int r = 0; r < 3; ++r )
                  g[i][r] = nrm[r] / sc;
            }
            // grad phi_i = sum_k lambda_k A[i][k] + B[i]
            V3 A[10][4], B[10];
            for ( int i = 0; i < 10; ++i )
            {
               B[i] = { 0, 0, 0 };
               for ( int k = 0; k < 4; ++k )
                  A[i][k] = { 0, 0, 0 };
            }
            for ( int v = 0; v < 4; ++v )
               for ( int r = 0; r < 3; ++r )
               {
                  A[v][v][r] = 4.0 * g[v][r];
                  B[v][r]    = -g[v][r];
               }
            for ( int e = 0; e < 6; ++e )
            {
               const int pa = kP2Edge[e][0], pb = kP2Edge[e][1];
               for ( int r = 0; r < 3; ++r )
               {
                  A[4 + e][pa][r] = 4.0 * g[pb][r];
                  A[4 + e][pb][r] = 4.0 * g[pa][r];
               }
            }
            for ( int i = 0; i < 10; ++i )
               for ( int j = 0; j < 10; ++j )
               {
                  double v = dot3( B[i], B[j] );
                  for ( int k = 0; k < 4; ++k )
                  {
                     v += 0.25 * ( dot3( A[i][k], B[j] ) + dot3( B[i], A[j][k] ) );
                     for ( int l = 0; l < 4; ++l )
                        v += dot3( A[i][k], A[j][l] ) * ( k == l ? 2.0 : 1.0 ) / 20.0;
                  }
                  el[size_t( c ) * 600 + 100 * t + 10 * i + j] = vol * v;
               }
         }
      }
      std::vector< float > elF( el.begin(), el.end() );
      ol.p2el              = upload( el );
      ol.p2elF             = upload( elF );
      // merged stencils of the cell-interior points (f = -1) and the one-sided stencils of the cell-face points
      // (f = 0..3): sum over the (type, node) pairs of a class of the element rows, keyed by the offset to the
      // other node; the face stencils take only the micro-cells on the cell's side of the face
      std::vector< double > st( size_t( m.nC() ) * 8 * kP2StencilMax, 0.0 ), fst( size_t( m.nC() ) * 32 * kP2StencilMax, 0.0 );
      for ( int f = -1; f < 4; ++f )
         for ( int cls = 0; cls < 8; ++cls )
         {
            const P2StencilOffsets off = p2StencilOffsets( cls, f );
            for ( int c = 0; c < m.nC(); ++c )
            {
               double* dst = f < 0 ? &st[( size_t( c ) * 8 + cls ) * kP2StencilMax] : &fst[( ( size_t( c ) * 4 + f ) * 8 + cls ) * kP2StencilMax];
               for ( int t = 0; t < 6; ++t )
                  for ( int l = 0; l < 10; ++l )
                  {
                     if ( p2NodeClass( t, l ) != cls || !p2OnCellSide( f, t, l ) )
                        continue;
                     for ( int j = 0; j < 10; ++j )
                     {
                        int e = 0;
                        while ( !( off.d[e][0] == p2NodeOffset( t, j, 0 ) - p2NodeOffset( t, l, 0 ) &&
                                   off.d[e][1] == p2NodeOffset( t, j, 1 ) - p2NodeOffset( t, l, 1 ) &&
                                   off.d[e][2] == p2NodeOffset( t, j, 2 ) - p2NodeOffset( t, l, 2 ) ) )
                           ++e;
                        dst[e] += el[size_t( c ) * 600 + 100 * t + 10 * l + j];
                     }
                  }
            }
         }
      std::vector< float > stF( st.begin(), st.end() ), fstF( fst.begin(), fst.end() );
      ol.p2st              = upload( st );
      ol.p2stF             = upload( stF );
      ol.p2fst             = upload( fst );
      ol.p2fstF            = upload( fstF );
      const size_t nLow    = mc.entriesHost.size() * size_t( N + 1 );
      const size_t nFace   = size_t( m.nF() ) * 2 * size_t( N + 1 ) * size_t( N + 1 );
      const size_t nFunc   = size_t( m.nC() ) * size_t( cellStride( L ) ) + size_t( kTailSlack );
      ol.lowPart   = static_cast< double* >( alloc( nLow * sizeof( double ) ) );
      ol.lowPartF  = static_cast< float* >( alloc( nLow * sizeof( float ) ) );
      ol.facePart  = static_cast< double* >( alloc( nFace * sizeof( double ) ) );
      ol.facePartF = static_cast< float* >( alloc( nFace * sizeof( float ) ) );
      ol.dinv      = static_cast< double* >( alloc( nFunc * sizeof( double ) ) );
      ol.dinvF     = static_cast< float* >( alloc( nFunc * sizeof( float ) ) );
   }
}

// P2ElementwiseOperator< p2_div_k_grad_affine_q4 >: per lattice level and per cell and type the 11 quadrature-point
// matrices w_q |det J| grad phi(l_q) grad phi(l_q)^T of the micro-cells of level lattice-1, the cell geometry
// for evaluating the coefficient, and the buffers (coefficient values, partial sums, inverse diagonal) filled
// on the device.
void assembleP2DivKGrad( hyg_op_s& op )
{
   const MeshCtx& mc = op.mesh->ctx;
   const Mesh&    m  = *mc.mesh;
   op.levels.resize( op.maxL + 1 );
   auto alloc = []( size_t bytes ) {
      void* d = nullptr;
      if ( bytes > 0 && hipMalloc( &d, bytes ) != hipSuccess )
         throw std::runtime_error( "hipMalloc failed for the P2 variable-coefficient operator" );
      return d;
   };
   for ( int L = op.minL; L <= op.maxL; ++L )
   {
      hyg_op_s::Level& ol = op.levels[L];
      const int        N  = 1 << L;
      const int        Ne = N / 2;
      std::vector< double >   el( size_t( m.nC() ) * 6600 );
      std::vector< ElemCell > ec( m.nC() );
      for ( int c = 0; c < m.nC(); ++c )
      {
         const auto& cv = m.Cv[c];
         const V3&   P0 = m.X[cv[0]];
         const V3    a1 = sub( m.X[cv[1]], P0 ), a2 = sub( m.X[cv[2]], P0 ), a3 = sub( m.X[cv[3]], P0 );
         for ( int r = 0; r < 3; ++r )
         {
            ec[c].v0[r]   = P0[r];
            ec[c].a[r][0] = a1[r] / double( Ne );
            ec[c].a[r][1] = a2[r] / double( Ne );
            ec[c].a[r][2] = a3[r] / double( Ne );
         }
         for ( int t = 0; t < 6; ++t )
         {
            V3 P[4];
            for ( int k = 0; k < 4; ++k )
               for ( int r = 0; r < 3; ++r )
                  P[k][r] = ( kMicroTet[t][k][0] * a1[r] + kMicroTet[t][k][1] * a2[r] + kMicroTet[t][k][2] * a3[r] ) / double( Ne );
            const double adet = std::fabs( dot3( sub( P[1], P[0] ), cross( sub( P[2], P[0] ), sub( P[3], P[0] ) ) ) );
            V3           g[4];
            for ( int i = 0; i < 4; ++i )
            {
               const int    a = ( i + 1 ) % 4, b = ( i + 2 ) % 4, d = ( i + 3 ) % 4;
               const V3     nrm = cross( sub( P[b], P[a] ), sub( P[d], P[a] ) );
               const double sc  = dot3( nrm, sub( P[i], P[a] ) );
               for ( int r = 0; r < 3; ++r )
                  g[i][r] = nrm[r] / sc;
            }
            for ( int q = 0; q < 11; ++q )
            {
               const double* lam = kXG4Lambda[q];
               V3            d[10];
               for ( int v = 0; v < 4; ++v )
                  for ( int r = 0; r < 3; ++r )
                     d[v][r] = ( 4.0 * lam[v] - 1.0 ) * g[v][r];
               for ( int e = 0; e < 6; ++e )
               {
                  const int pa = kP2Edge[e][0], pb = kP2Edge[e][1];
                  for ( int r = 0; r < 3; ++r )
                     d[4 + e][r] = 4.0 * ( lam[pa] * g[pb][r] + lam[pb] * g[pa][r] );
               }
               for ( int i = 0; i < 10; ++i )
                  for ( int j = 0; j < 10; ++j )
                     el[size_t( c ) * 6600 + 1100 * t + 100 * q + 10 * i + j] = kXG4Weight[q] * adet * dot3( d[i], d[j] );
            }
         }
      }
      std::vector< float > elF( el.begin(), el.end() );
      ol.p2el              = upload( el );
      ol.p2elF             = upload( elF );
      ol.elem              = upload( ec );
      const size_t nKq     = size_t( m.nC() ) * 6 * size_t( S3( Ne + 1 ) ) * 11;
      const size_t nLow    = mc.entriesHost.size() * size_t( N + 1 );
      const size_t nFace   = size_t( m.nF() ) * 2 * size_t( N + 1 ) * size_t( N + 1 );
      const size_t nFunc   = size_t( m.nC() ) * size_t( cellStride( L ) ) + size_t( kTailSlack );
      ol.p2kq      = static_cast< double* >( alloc( nKq * sizeof( double ) ) );
      ol.p2kqF     = static_cast< float* >( alloc( nKq * sizeof( float ) ) );
      // partial sums for up to three components (a vector function's components applied in one pass)
      ol.lowPart   = static_cast< double* >( alloc( 3 * nLow * sizeof( double ) ) );
      ol.lowPartF  = static_cast< float* >( alloc( 3 * nLow * sizeof( float ) ) );
      ol.facePart  = static_cast< double* >( alloc( 3 * nFace * sizeof( double ) ) );
      ol.facePartF = static_cast< float* >( alloc( 3 * nFace * sizeof( float ) ) );
      ol.dinv      = static_cast< double* >( alloc( nFunc * sizeof( double ) ) );
      ol.dinvF     = static_cast< float* >( alloc( nFunc * sizeof( float ) ) );
   }
}

} // namespace hyg

hyg_op_s::~hyg_op_s()
{
   for ( auto& l : levels )
   {
      (void) hipFree( l.cellSt );
      (void) hipFree( l.cellStF );
      (void) hipFree( l.entryW );
      (void) hipFree( l.entryWF );
      (void) hipFree( l.primInv );
      (void) hipFree( l.primInvF );
      (void) hipFree( l.facePart );
      (void) hipFree( l.facePartF );
      (void) hipFree( l.elem );
      (void) hipFree( l.kb );
      (void) hipFree( l.kbF );
      (void) hipFree( l.kbPhase );
      (void) hipFree( l.kbTab );
      (void) hipFree( l.kbTabF );
      (void) hipFree( l.lowPart );
      (void) hipFree( l.lowPartF );
      (void) hipFree( l.dinv );
      (void) hipFree( l.dinvF );
      (void) hipFree( l.p2el );
      (void) hipFree( l.p2elF );
      (void) hipFree( l.p2kq );
      (void) hipFree( l.p2kqF );
      (void) hipFree( l.p2st );
      (void) hipFree( l.p2stF );
      (void) hipFree( l.p2fst );
      (void) hipFree( l.p2fstF );
      (void) hipFree( l.p2s );
      (void) hipFree( l.p2sF );
      (void) hipFree( l.surPoly );
      (void) hipFree( l.bw );
      (void) hipFree( l.bwF );
      (void) hipFree( l.surBnd );
   }
   (void) hipFree( modes );
   (void) hipFree( kPoly );
}
