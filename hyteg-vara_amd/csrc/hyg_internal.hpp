// Internal declarations of the MI355X P1 multigrid library (host + device).
//
// Data layout in HBM. A function at level L stores, for every macro-cell, one array of
// S(w) = w(w+1)(w+2)/6 values (w = 2^L + 1) in the reference's tetrahedral linearisation
// index(x,y,z) = S(w) - S(w-z) + y(w-z+1) - y(y+1)/2 + x
// (src/hyteg/indexing/MacroCellIndexing.hpp:40-53), padded to a multiple of 64 values so
// every cell array starts 512-byte aligned. The cell arrays of all macro-cells of a level are
// contiguous. Values on macro-faces, -edges and -vertices are held by every incident cell
// (a replica per cell) and every kernel that writes such a value writes all replicas, so a
// function is always consistent and no halo refresh precedes an operator.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace hyg {

using i64 = int64_t;

// ------------------------------------------------------------------ indexing
__host__ __device__ inline i64 S3( i64 w ) { return w * ( w + 1 ) * ( w + 2 ) / 6; }
__host__ __device__ inline i64 T2( i64 w ) { return w * ( w + 1 ) / 2; }
// 32-bit variant: w (w + 1) (w + 2) < 2^31 up to w = 1025 (level 10)
__host__ __device__ inline int S3i( int w ) { return ( w * ( w + 1 ) * ( w + 2 ) ) / 6; }
__host__ __device__ inline int T2i( int w ) { return ( w * ( w + 1 ) ) >> 1; }
__host__ __device__ inline int cidx( int w, int x, int y, int z )
{
   return S3i( w ) - S3i( w - z ) + y * ( w - z + 1 ) - ( y * ( y + 1 ) ) / 2 + x;
}
// start of row y in a plane whose row 0 has W points
__host__ __device__ inline int rowStart( int W, int y ) { return y * W - ( y * ( y - 1 ) ) / 2; }

inline i64 cellStride( int L )
{
   const i64 w = ( i64( 1 ) << L ) + 1;
   return ( S3( w ) + 63 ) / 64 * 64;
}
// values allocated after the last cell array of a level: row-window kernels read up to 128 values
// past a cell's last point (and discard them)
constexpr i64 kTailSlack = 256;

// stencil directions in the reference order C W E N S NW SE TC TW TS TSE BC BN BE BNW
// (src/hyteg/p1functionspace/VertexDoFIndexing.cpp:484-520)
constexpr int NDIR = 15;
__host__ __device__ constexpr int dirX( int d )
{
   constexpr int t[NDIR] = { 0, -1, 1, 0, 0, -1, 1, 0, -1, 0, 1, 0, 0, 1, -1 };
   return t[d];
}
__host__ __device__ constexpr int dirY( int d )
{
   constexpr int t[NDIR] = { 0, 0, 0, 1, -1, 1, -1, 0, 0, -1, -1, 0, 1, 0, 1 };
   return t[d];
}
__host__ __device__ constexpr int dirZ( int d )
{
   constexpr int t[NDIR] = { 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1 };
   return t[d];
}

// Local primitive slot of a cell point: 0 interior, 1..4 faces 0..3, 5..10 edges 0..5, 11..14 vertices 0..3.
// Cell-local faces: f0=(0,1,2) f1=(0,1,3) f2=(0,2,3) f3=(1,2,3); edges e0=(0,1) e1=(0,2) e2=(1,2) e3=(0,3) e4=(1,3) e5=(2,3).
__host__ __device__ inline int classifySlot( int N, int x, int y, int z )
{
   const int b0 = N - x - y - z;
   const int zm = ( b0 == 0 ) | ( ( x == 0 ) << 1 ) | ( ( y == 0 ) << 2 ) | ( ( z == 0 ) << 3 );
   // lookup by zero-mask of the barycentric coordinates (b0, x, y, z)
   // one zero -> face opposite; two zeros -> edge of the two non-zero ones; three zeros -> vertex
   constexpr signed char lut[16] = {
       0,      // none zero: interior
       1 + 3,  // b0=0  -> face 3 (1,2,3)
       1 + 2,  // x=0   -> face 2 (0,2,3)
       5 + 5,  // b0,x  -> edge (2,3) = e5
       1 + 1,  // y=0   -> face 1 (0,1,3)
       5 + 4,  // b0,y  -> edge (1,3) = e4
       5 + 3,  // x,y   -> edge (0,3) = e3
       11 + 3, // b0,x,y -> vertex 3
       1 + 0,  // z=0   -> face 0 (0,1,2)
       5 + 2,  // b0,z  -> edge (1,2) = e2
       5 + 1,  // x,z   -> edge (0,2) = e1
       11 + 2, // b0,x,z -> vertex 2
       5 + 0,  // y,z   -> edge (0,1) = e0
       11 + 1, // b0,y,z -> vertex 1
       11 + 0, // x,y,z -> vertex 0
       -1 };
   return lut[zm];
}

__host__ __device__ inline bool inCell( int N, int x, int y, int z ) { return x >= 0 && y >= 0 && z >= 0 && x + y + z <= N; }

// ------------------------------------------------------------------ device tables
// A tile of 1024 consecutive points of one z-plane of a cell array (256 threads x 4 points).
struct PlaneTile
{
   int z;
   int s0;
};

// Work of one z-marching wavefront: rows y .. y+R-1 and points x0-1 .. x0+64P-2 (outputs
// x0 .. x0+64P-3), R rows and P points per lane of the table's configuration, planes [z0, z1).
struct ZSeg
{
   int y;
   int x0;
   int z0;
   int z1;
};
constexpr int kZSegChunk = 24; // planes per z-marching segment (short segments balance the tail of the launch)
// P2 constant operator, z-marching row pairs: a wavefront computes rows y0 (even) and y0+1 of a segment of
// kP2ZSeg x values (x0 = kP2ZSeg k - 2 + 2 lane, lanes 0 and 63 halo) over planes [z0, z1), z0 even; the
// table stores (y0, k, z0, z1) in a ZSeg.
constexpr int kP2ZSeg   = 124;
constexpr int kP2ZChunk = 32; // planes per segment where N >= 256 (N = 128: 16, smaller lattices: 8)
// configurations of the P2 row kernel (tuned by hyg_op_tune): planes per z-segment kP2Chunks[cfg / 2] and planes in
// flight cfg % 2 + 1
constexpr int kP2Chunks[4] = { 8, 16, 32, 64 };
constexpr int kNumP2Cfgs   = 8;
inline int    p2DefaultConfig( int N ) { return 2 * ( N >= 256 ? 2 : ( N >= 128 ? 1 : 0 ) ); }
// cell-interior kernel configurations: rows per wavefront, points per lane, planes in flight
constexpr int kCellRows[]   = { 1, 2, 4, 1, 2, 4, 8 };
constexpr int kCellPoints[] = { 1, 1, 1, 1, 1, 1, 1 };
constexpr int kCellDepth[]  = { 2, 2, 2, 4, 4, 4, 2 };
constexpr int kNumCellCfgs  = 7;

// Lower-dimensional primitive (macro-vertex, -edge or -face); order: vertices, edges, faces.
struct LowerPrim
{
   int kind;   // 0 vertex, 1 edge, 2 face
   int bc;     // DoFType of the primitive
   int nCells; // incident macro-cells
   int entry0; // first LowerEntry
};

// One (lower-dimensional primitive, incident cell) pair. The primitive's vertices are the cell-local
// vertices l0, l1, l2 (l1/l2 unused for vertices, l2 for edges); its point (i, j) has cell coordinates
// N*e(l0) + i*(e(l1) - e(l0)) + j*(e(l2) - e(l0)), e(0) = 0, e(1..3) = unit vectors.
struct LowerEntry
{
   int cell;
   int l0, l1, l2;
};

// A run of up to 64 points of one lower-dimensional primitive: face row j (i from i0), edge (i from i0), vertex.
struct LowerChunk
{
   int prim;
   int j;
   int i0;
   int n;
};

// One face of one macro-cell, for the stencil kernels that compute one-sided face sums. Local face slots
// in the cell's layout: 0 z = 0, 1 y = 0, 2 x = 0, 3 x+y+z = N.
struct CellFaceSlot
{
   int entry; // LowerEntry of this (face, cell) pair: its one-sided weights are entryW[16 * entry ..]
   int prim;  // LowerPrim index of the face
   int block; // 2 * face index + side (position of the cell in the face's cell list): partial-sum block
   int bc;    // DoFType of the face
};
// Partial sums of a face side at level L are stored in a block of (N+1)^2 values, indexed by the cell
// coordinates of the point: slot 0 x + y (N+1), slot 1 x + z (N+1), slots 2 and 3 z + y (N+1).
__host__ __device__ inline int facePartIndex( int slot, int N, int x, int y, int z )
{
   return slot == 0 ? x + y * ( N + 1 ) : ( slot == 1 ? x + z * ( N + 1 ) : z + y * ( N + 1 ) );
}
// local face slot of a face entry (the slot is the face opposite the local vertex not among l0, l1, l2)
__host__ __device__ inline int entrySlot( const LowerEntry& e ) { return e.l0 + e.l1 + e.l2 - 3; }

// Per-cell data used by pointwise kernels.
struct CellInfo
{
   uint64_t bcword;    // DoFType (4 bits) of local slots 0..14
   uint32_t ownmask;   // bit k: this cell is the designated owner of local slot k (bit 0 always)
   int      prim[15];  // LowerPrim index of local slot k >= 1 (slot 0: -1)
   int      entry[15]; // LowerEntry of the (primitive, this cell) pair of local slot k >= 1 (slot 0: -1)
   double   invNC[16]; // 1 / #cells of the local slot's primitive (slot 0: 1)
   int      colour[4]; // multi-colour SOR: colour of each local vertex in the macro-vertex colouring
};

// Geometry of one macro-cell for the variable-coefficient elementwise operator at one level.
struct ElemCell
{
   double v0[3];       // position of local vertex 0
   double a[3][3];     // a[r][k]: component r of (v_{k+1} - v0) / N, the lattice step along local axis k
   double mt[6][16];   // element matrix of each micro-cell type for k = 1 (vertex order of the type)
};

// P2 (quadratic Lagrange) on the refined vertex lattice: a P2 function of level L keeps its vertex DoFs at
// the even points and its edge DoFs at the odd points of the vertex lattice of level L+1 ("lattice level").
// Tables shared by the P2 kernels (built on the host once per operator). Local node order of a micro-cell:
// vertices 0..3, edges (2,3) (1,3) (1,2) (0,3) (0,2) (0,1).
struct P2Tables
{
   // lattice offset of node l of a micro-cell of type t from 2 * anchor (vertices 2 e_v, edges e_a + e_b)
   int nd[6][10][3];
   // per parity class px + 2 py + 4 pz of a lattice point: the (type, node) pairs (16 t + l) it can be
   int nEnt[8];
   int ent[8][24];
   // prolongation: per residue (px & 3) + 4 (py & 3) + 16 (pz & 3) of a point of the fine lattice, the
   // coarse micro-cells containing it (anchor = (p >> 2) - offset bits o) with the basis values there
   struct Cand
   {
      int    t, o;
      double phi[10];
   };
   int  nCand[64];
   Cand cand[64][6];
   // restriction: per parity class of a coarse lattice point n, the fine lattice points 2 n + d in the
   // support of its basis function with the weights phi_n
   struct Sup
   {
      int    d[3];
      double w;
   };
   int nSup[8];
   Sup sup[8][128];
};
void buildP2Tables( P2Tables& t );

// micro-cell types WHITE_UP, BLUE_UP, GREEN_UP, WHITE_DOWN, BLUE_DOWN, GREEN_DOWN: vertex offsets from the anchor
// (volumedofspace/CellDoFIndexing.hpp:155-198)
constexpr int kMicroCell[6][4][3] = { { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } },
                                      { { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 }, { 1, 0, 1 } },
                                      { { 1, 0, 0 }, { 0, 1, 0 }, { 1, 0, 1 }, { 0, 0, 1 } },
                                      { { 1, 1, 0 }, { 1, 1, 1 }, { 0, 1, 1 }, { 1, 0, 1 } },
                                      { { 1, 0, 1 }, { 0, 1, 1 }, { 0, 0, 1 }, { 0, 1, 0 } },
                                      { { 0, 1, 0 }, { 1, 1, 0 }, { 1, 0, 1 }, { 0, 1, 1 } } };
constexpr int kP2EdgeNodes[6][2] = { { 2, 3 }, { 1, 3 }, { 1, 2 }, { 0, 3 }, { 0, 2 }, { 0, 1 } };

// lattice offset of P2 node l of a micro-cell of type t from 2 * anchor
constexpr int p2NodeOffset( int t, int l, int c )
{
   return l < 4 ? 2 * kMicroCell[t][l][c] : kMicroCell[t][kP2EdgeNodes[l - 4][0]][c] + kMicroCell[t][kP2EdgeNodes[l - 4][1]][c];
}
constexpr int p2NodeClass( int t, int l )
{
   return ( p2NodeOffset( t, l, 0 ) & 1 ) | ( ( p2NodeOffset( t, l, 1 ) & 1 ) << 1 ) | ( ( p2NodeOffset( t, l, 2 ) & 1 ) << 2 );
}

// largest coordinate sum of the vertices of a micro-cell type (in level-L units)
constexpr int p2CellMaxSum( int t )
{
   int m = 0;
   for ( int v = 0; v < 4; ++v )
   {
      const int sum = kMicroCell[t][v][0] + kMicroCell[t][v][1] + kMicroCell[t][v][2];
      m             = sum > m ? sum : m;
   }
   return m;
}
// Does the micro-cell of type t that has a point of cell face f as its node l lie on the cell's side of the
// face? Faces 0..3 are z = 0, y = 0, x = 0 and x + y + z = N; f < 0 means no face (every micro-cell).
constexpr bool p2OnCellSide( int f, int t, int l )
{
   return f < 0 || ( f < 3 ? p2NodeOffset( t, l, 2 - f ) == 0
                           : p2NodeOffset( t, l, 0 ) + p2NodeOffset( t, l, 1 ) + p2NodeOffset( t, l, 2 ) >= 2 * p2CellMaxSum( t ) );
}

// Merged P2 stencil of a cell-interior point of parity class cls: the distinct offsets (within [-2, 2]^3) to
// the nodes of all micro-cells that have the point as a node, in order of first appearance over (t, l, j).
// Sizes 65 (vertex DoFs), 27 or 19 (edge DoFs). With a face f >= 0, the one-sided stencil of a point inside
// that cell face (not on a cell edge): only the micro-cells on the cell's side of the face.
struct P2StencilOffsets
{
   int n = 0;
   int d[65][3] = {};
};
constexpr P2StencilOffsets p2StencilOffsets( int cls, int f = -1 )
{
   P2StencilOffsets s;
   for ( int t = 0; t < 6; ++t )
      for ( int l = 0; l < 10; ++l )
      {
         if ( p2NodeClass( t, l ) != cls || !p2OnCellSide( f, t, l ) )
            continue;
         for ( int j = 0; j < 10; ++j )
         {
            const int dx = p2NodeOffset( t, j, 0 ) - p2NodeOffset( t, l, 0 ), dy = p2NodeOffset( t, j, 1 ) - p2NodeOffset( t, l, 1 ),
                      dz = p2NodeOffset( t, j, 2 ) - p2NodeOffset( t, l, 2 );
            bool seen = false;
            for ( int e = 0; e < s.n; ++e )
               seen = seen || ( s.d[e][0] == dx && s.d[e][1] == dy && s.d[e][2] == dz );
            if ( !seen )
            {
               s.d[s.n][0] = dx;
               s.d[s.n][1] = dy;
               s.d[s.n][2] = dz;
               ++s.n;
            }
         }
      }
   return s;
}
constexpr int kP2StencilMax = 65;

// ------------------------------------------------------------------ P2 transfer stencils (compile-time tables)
// P2 basis (vertices 0..3, then the edges of kP2EdgeNodes) at barycentric coordinates lam
constexpr void p2BasisAt( const double lam[4], double phi[10] )
{
   for ( int v = 0; v < 4; ++v )
      phi[v] = lam[v] * ( 2.0 * lam[v] - 1.0 );
   for ( int e = 0; e < 6; ++e )
      phi[4 + e] = 4.0 * lam[kP2EdgeNodes[e][0]] * lam[kP2EdgeNodes[e][1]];
}
// 4 x barycentric coordinates of point r (units of a quarter micro-edge) in the micro-cell of type t whose
// anchor is at -o; false if outside. The vertex offsets of a type span a unimodular lattice basis.
constexpr bool p2QuarterBary( int t, const int o[3], const int r[3], int l4[4] )
{
   int v[4][3] = {};
   for ( int k = 0; k < 4; ++k )
      for ( int c = 0; c < 3; ++c )
         v[k][c] = kMicroCell[t][k][c] - o[c];
   long J[3][3] = {};
   for ( int a = 0; a < 3; ++a )
      for ( int c = 0; c < 3; ++c )
         J[a][c] = v[c + 1][a] - v[0][a];
   const long det = J[0][0] * ( J[1][1] * J[2][2] - J[1][2] * J[2][1] ) - J[0][1] * ( J[1][0] * J[2][2] - J[1][2] * J[2][0] ) +
                    J[0][2] * ( J[1][0] * J[2][1] - J[1][1] * J[2][0] );
   const long A[3][3] = { { J[1][1] * J[2][2] - J[1][2] * J[2][1], J[0][2] * J[2][1] - J[0][1] * J[2][2], J[0][1] * J[1][2] - J[0][2] * J[1][1] },
                          { J[1][2] * J[2][0] - J[1][0] * J[2][2], J[0][0] * J[2][2] - J[0][2] * J[2][0], J[0][2] * J[1][0] - J[0][0] * J[1][2] },
                          { J[1][0] * J[2][1] - J[1][1] * J[2][0], J[0][1] * J[2][0] - J[0][0] * J[2][1], J[0][0] * J[1][1] - J[0][1] * J[1][0] } };
   const long d[3]    = { r[0] - 4 * v[0][0], r[1] - 4 * v[0][1], r[2] - 4 * v[0][2] };
   l4[0]              = 4;
   for ( int k = 0; k < 3; ++k )
   {
      const long num = A[k][0] * d[0] + A[k][1] * d[1] + A[k][2] * d[2];
      l4[k + 1]      = int( num / det );
      l4[0] -= l4[k + 1];
   }
   for ( int k = 0; k < 4; ++k )
      if ( l4[k] < 0 )
         return false;
   return true;
}

// Quadratic interpolation at a fine lattice point P = 4 Q + r (fine lattice = 2 x coarse lattice): the coarse
// lattice points 2 Q + d and weights of the first micro-cell (type-major, then anchor offset o) of the coarse
// P2 level that contains it; points of the coarse lattice (r even) take the coarse value itself. Every
// micro-cell containing a cell-interior point lies in the cell, so any of them gives the interpolant.
struct P2ProlongStencil
{
   int    n = 0;
   int    d[10][3] = {};
   double w[10] = {};
};
constexpr P2ProlongStencil p2ProlongStencil( int rx, int ry, int rz )
{
   P2ProlongStencil ps;
   if ( ( ( rx | ry | rz ) & 1 ) == 0 )
   {
      ps.n       = 1;
      ps.d[0][0] = rx / 2;
      ps.d[0][1] = ry / 2;
      ps.d[0][2] = rz / 2;
      ps.w[0]    = 1.0;
      return ps;
   }
   const int r[3] = { rx, ry, rz };
   for ( int t = 0; t < 6; ++t )
      for ( int o = 0; o < 8; ++o )
      {
         const int oo[3] = { o & 1, ( o >> 1 ) & 1, ( o >> 2 ) & 1 };
         int       l4[4] = {};
         if ( !p2QuarterBary( t, oo, r, l4 ) )
            continue;
         const double lam[4]  = { l4[0] / 4.0, l4[1] / 4.0, l4[2] / 4.0, l4[3] / 4.0 };
         double       phi[10] = {};
         p2BasisAt( lam, phi );
         for ( int j = 0; j < 10; ++j )
            if ( phi[j] != 0.0 )
            {
               for ( int c = 0; c < 3; ++c )
                  ps.d[ps.n][c] = p2NodeOffset( t, j, c ) - 2 * oo[c];
               ps.w[ps.n++] = phi[j];
            }
         return ps;
      }
   return ps;
}

// Restriction (transpose of the interpolation) at a coarse point n of parity class cls away from the cell
// boundary: the fine lattice points 2 n + d where its basis function is nonzero, with those values. The support
// is the union of the coarse micro-cells that have n as a node; each contributes its 35 quarter points.
constexpr int kP2RestrictMax = 128;
struct P2RestrictStencil
{
   int    n = 0;
   int    d[kP2RestrictMax][3] = {};
   double w[kP2RestrictMax] = {};
};
constexpr P2RestrictStencil p2RestrictStencil( int cls )
{
   P2RestrictStencil rs;
   const int         n[3] = { 8 + ( cls & 1 ), 8 + ( ( cls >> 1 ) & 1 ), 8 + ( ( cls >> 2 ) & 1 ) };
   for ( int t = 0; t < 6; ++t )
      for ( int l = 0; l < 10; ++l )
      {
         if ( p2NodeClass( t, l ) != cls )
            continue;
         int a[3] = {};
         for ( int c = 0; c < 3; ++c )
            a[c] = ( n[c] - p2NodeOffset( t, l, c ) ) / 2;
         for ( int i = 0; i <= 4; ++i )
            for ( int j = 0; i + j <= 4; ++j )
               for ( int q = 0; i + j + q <= 4; ++q )
               {
                  const int    l4[4]   = { 4 - i - j - q, i, j, q };
                  const double lam[4]  = { l4[0] / 4.0, l4[1] / 4.0, l4[2] / 4.0, l4[3] / 4.0 };
                  double       phi[10] = {};
                  p2BasisAt( lam, phi );
                  if ( phi[l] == 0.0 )
                     continue;
                  int d[3] = {};
                  for ( int c = 0; c < 3; ++c )
                  {
                     int p = 4 * a[c];
                     for ( int v = 0; v < 4; ++v )
                        p += l4[v] * kMicroCell[t][v][c];
                     d[c] = p - 2 * n[c];
                  }
                  bool seen = false;
                  for ( int m = 0; m < rs.n; ++m )
                     seen = seen || ( rs.d[m][0] == d[0] && rs.d[m][1] == d[1] && rs.d[m][2] == d[2] );
                  if ( seen )
                     continue;
                  for ( int c = 0; c < 3; ++c )
                     rs.d[rs.n][c] = d[c];
                  rs.w[rs.n++] = phi[l];
               }
      }
   return rs;
}


// Xiao-Gimbutas degree-4 rule with 11 points of p2_div_k_grad_affine_q4: barycentric coordinates (vertices 0..3)
// and weights (summing to 1/6, the factors of |det J|)
constexpr double kXG4Lambda[11][4] = {
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
constexpr double kXG4Weight[11] = { 0.006541848487473325,  0.0092122281926561474, 0.032235135341605742, 0.0092322998119293929,
                                    0.0099888641910932524, 0.011578327656272558,  0.012693785874259723, 0.013237780011337548,
                                    0.017744672359248346,  0.018372372071416277,  0.025829352669374347 };

// ------------------------------------------------------------------ host objects
struct Mesh
{
   std::vector< std::array< double, 3 > > X;
   std::vector< std::array< int, 2 > >    E;
   std::vector< std::array< int, 3 > >    F;
   std::vector< std::array< int, 4 > >    Cv; // cell vertices in the reference's local order
   std::vector< std::array< int, 4 > >    cF;
   std::vector< std::array< int, 6 > >    cE;
   std::vector< std::vector< int > >      vC, eC, fC;
   std::vector< uint8_t >                 vB, eB, fB; // mesh boundary flags (0 inner, 1 Dirichlet, 2 Neumann, 3 free-slip)
   int nV() const { return (int) X.size(); }
   int nE() const { return (int) E.size(); }
   int nF() const { return (int) F.size(); }
   int nC() const { return (int) Cv.size(); }
};

// Spherical-shell blending of one macro-cell (the reference's IcosahedralShellMap, geometry/IcosahedralShellMap.hpp):
// lambda(x) = |(x - ray) . n / vt| with n = (thr - ray) x (fo - ray), vt = (ref - ray) . n is the fraction of the way
// from the plane (ray, thr, fo) to ref, and F(x) = x (rRay + lambda (rRef - rRay)) / |x|.
struct ShellMap
{
   double ref[3], ray[3], thr[3], fo[3];
   double rRef, rRay;
   double n[3], vt;
};
__host__ __device__ inline double shellLambda( const ShellMap& s, const double x[3], double g[3] )
{
   const double v  = ( ( x[0] - s.ray[0] ) * s.n[0] + ( x[1] - s.ray[1] ) * s.n[1] + ( x[2] - s.ray[2] ) * s.n[2] ) / s.vt;
   const double sg = v < 0.0 ? -1.0 : 1.0;
   for ( int k = 0; k < 3; ++k )
      g[k] = sg * s.n[k] / s.vt;
   return v < 0.0 ? -v : v;
}
__host__ __device__ inline void shellF( const ShellMap& s, const double x[3], double y[3] )
{
   double       g[3];
   const double r = s.rRay + shellLambda( s, x, g ) * ( s.rRef - s.rRay );
   const double q = r / sqrt( x[0] * x[0] + x[1] * x[1] + x[2] * x[2] );
   for ( int k = 0; k < 3; ++k )
      y[k] = x[k] * q;
}
// Jacobian of shellF: DF = (r / |x|) I + x (r' grad lambda / |x| - r x / |x|^3)^T, r' = rRef - rRay
__host__ __device__ inline void shellDF( const ShellMap& s, const double x[3], double D[3][3] )
{
   double       g[3];
   const double r  = s.rRay + shellLambda( s, x, g ) * ( s.rRef - s.rRay );
   const double nx = sqrt( x[0] * x[0] + x[1] * x[1] + x[2] * x[2] );
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j )
         D[i][j] = ( i == j ? r / nx : 0.0 ) + x[i] * ( ( s.rRef - s.rRay ) * g[j] / nx - r * x[j] / ( nx * nx * nx ) );
}

std::unique_ptr< Mesh > readGmsh( const std::string& path, std::string& err, bool keepTags = false );
// DoF type of a mesh boundary flag under BoundaryCondition::create0123BC (boundary/BoundaryConditions.cpp:26-33):
// 1 Dirichlet, 2 Neumann, 3 free-slip, every other flag Inner
inline int bcOfFlag( int flag ) { return flag == 1 ? 2 : flag == 2 ? 4 : flag == 3 ? 8 : 1; }
std::unique_ptr< Mesh > meshSphericalShell( int ntan, int nrad, double rmin, double rmax, std::string& err );
std::unique_ptr< Mesh > meshSymmetricCuboid( const double lo[3], const double hi[3], int nx, int ny, int nz, std::string& err );
std::unique_ptr< Mesh > meshCuboid( const double lo[3], const double hi[3], int nx, int ny, int nz, std::string& err );
std::unique_ptr< Mesh > meshCubedDomain( const int* cubes, int nCubes, int cubeType, std::string& err );

// element matrices of the six micro-tetrahedron types of a cell with vertices cv at level L, and the
// one-sided 15-point stencil of that cell at lattice point (qx, qy, qz) (w[0..14])
void elementMatrices( const Mesh& m, const std::array< int, 4 >& cv, int L, double K[6][4][4] );
void stencilAt( const double K[6][4][4], int N, int qx, int qy, int qz, double* w );

// canonical DoF count, V + E(N-1) + F(N-1)(N-2)/2 + C(N-1)(N-2)(N-3)/6
i64 numDofs( const Mesh& m, int L );

// Per-level device geometry shared by all functions and operators of a mesh.
struct LevelGeom
{
   int                        L = -1;
   i64                        stride = 0; // per-cell array stride (values)
   std::vector< PlaneTile >   tilesHost;  // all planes z = 0..N, z-major
   PlaneTile*                 tiles = nullptr;
   int                        nTiles = 0;
   int                        interiorTile0 = 0; // tiles of planes 1..N-3 (points off the cell boundary)
   int                        nInteriorTiles = 0;
   // z-marching segments per configuration (same for every cell)
   std::vector< ZSeg >                      zSegsHost[kNumCellCfgs];
   ZSeg*                                    zSegs[kNumCellCfgs]  = {};
   int                                      nZSegs[kNumCellCfgs] = {};
   int                                      cellCfg[2][6]; // chosen configuration per value type (FP64, FP32) and stencil mode, -1 = default
   LevelGeom()
   {
      for ( auto& t : cellCfg )
         for ( int& v : t )
            v = -1;
   }
   std::vector< int4 >                      flatItemsHost; // z = 0 / y = 0 face rows: (cell, slot, line, x0)
   int4*                                    flatItems  = nullptr;
   int                                      nFlatItems = 0;
   // P2 constant operator row-pair segments of this lattice level, one table per segment length kP2Chunks[k], and the
   // chosen row-kernel configuration per value type and mode (-1: p2DefaultConfig)
   std::vector< ZSeg >                      p2ZSegsHost[4];
   ZSeg*                                    p2ZSegs[4]  = {};
   int                                      nP2ZSegs[4] = {};
   int                                      p2Cfg[2][6] = { { -1, -1, -1, -1, -1, -1 }, { -1, -1, -1, -1, -1, -1 } };
   std::vector< LowerChunk >                chunksHost; // macro-vertex, then -edge, then -face work
   LowerChunk*                              chunks  = nullptr;
   int                                      nChunks = 0;
   int                                      firstFaceChunk = 0; // chunks [firstFaceChunk, nChunks) are face rows
};

// Reference-layout halo tables of one level (csrc/halo.cpp): gather[k] = offset in a function's level array of element
// k of the reference's lower-dimensional primitive memory (own values and ghost layers); for the nOwn own values,
// ownSrc = their element index, replica[ownRow[j] .. ownRow[j+1]) the offsets of all replicas, ownPrimBc the DoF type;
// sumOwn lists the own values of points inside their primitive (each point once: edge ends and face borders belong
// to the vertices and edges)
struct HaloMaps
{
   i64  size = 0;
   int  nOwn = 0, nSum = 0;
   int* sumOwn = nullptr;
   i64* gather = nullptr;
   int* ownRow = nullptr;
   i64* ownSrc = nullptr;
   i64* replica = nullptr;
   int* ownPrimBc = nullptr;
   ~HaloMaps();
};
// element offsets of the reference's per-primitive memory (vertices, edges, faces in storage order; size nV+nE+nF+1)
std::vector< i64 > haloLayout( const Mesh& m, int L );

struct MeshCtx
{
   std::unique_ptr< Mesh >     mesh;
   std::vector< LevelGeom >    levels; // indexed by level (lazily built)
   std::vector< CellInfo >     cellInfoHost;
   std::vector< LowerPrim >    primsHost;
   std::vector< LowerEntry >   entriesHost;
   std::vector< CellFaceSlot > cellFacesHost; // 4 per cell
   LowerPrim*                  prims    = nullptr;
   LowerEntry*                 entries  = nullptr;
   CellInfo*                   cellInfo = nullptr;
   CellFaceSlot*               cellFaces = nullptr;
   int                         device   = 0;
   P2Tables*                   p2tab    = nullptr; // device copy, built on first use
   int                         nColours = 0;       // colours of the macro-vertex colouring (multi-colour SOR)
   std::vector< int >          vertexColours;      // colour of every macro-vertex
   std::vector< std::unique_ptr< HaloMaps > > haloMaps; // per level, built on first use
   const HaloMaps&             halo( int L );
   // spherical-shell blending (hyg_mesh_set_shell_map): the map of every cell, host and device; empty: affine cells
   std::vector< ShellMap >     shellHost;
   ShellMap*                   shell = nullptr;
   void                        setShellMap(); // throws when a cell is not a shell cell
   void                        buildTopology(); // prims, entries, cell info (host + device)
   LevelGeom&                  geom( int L );
   const P2Tables*             p2Tables();
   ~MeshCtx();
};

} // namespace hyg

struct hyg_mesh_s
{
   hyg::MeshCtx ctx;
};

struct hyg_func_s
{
   hyg_mesh_s*           mesh;
   int                   minL, maxL, dtype; // lattice levels (a P2 function of level L holds lattice L+1)
   int                   lat = 0;           // 0: P1 function, 1: P2 function
   uint64_t              gen = 0;           // creation number, unique over the process (cache keys)
   std::vector< void* >  data; // per level (nullptr below minL)
   size_t                elem() const { return dtype == 8 ? 8 : 4; }
};

// Tables of the reference-order SOR of one level, built on first use: the lower-dimensional primitives by kind,
// the face points by anti-diagonal i + 2j and the cell-interior points by hyperplane x + 2y + 3z (with offsets), and
// the snapshot buffers of the function before a sweep (FP64, FP32).
struct SorOrderTables
{
   int*               vertices = nullptr;
   int*               edges    = nullptr;
   int*               faces    = nullptr;
   int                nVertices = 0, nEdges = 0, nFaces = 0;
   int*               facePts  = nullptr; // i | j << 16
   int*               faceOff  = nullptr;
   int                nFaceTau = 0;
   int*               cellPts  = nullptr; // cell-array indices
   std::vector< int > cellOff;            // host: offsets per hyperplane
   void*              snap64   = nullptr;
   void*              snap32   = nullptr;
   SorOrderTables()                        = default;
   SorOrderTables( const SorOrderTables& ) = delete;
   SorOrderTables& operator=( const SorOrderTables& ) = delete;
   ~SorOrderTables();
};

struct hyg_op_s
{
   hyg_mesh_s* mesh;
   int         minL, maxL; // lattice levels
   int         lat = 0;    // 0: P1 operator, 1: P2 operator (its functions hold lattice level L+1)
   // 0: constant-coefficient stencils (P1ConstantLaplaceOperator)
   // 1: elementwise -div(k grad) with k = k0 exp( sum_m a_m cos(2 pi kappa_m . x + phi_m) )
   // 2: P2 constant-coefficient -Laplace (P2ConstantLaplaceOperator), merged stencils
   // 3: P2 elementwise -div(k grad) (p2_div_k_grad_affine_q4, 11-point rule), k of P2Coefficient
   int                   kind = 0;
   // kind 1 after hyg_op_store_local_matrices: the coefficient integrals kb of every level are kept (not
   // re-evaluated per apply); kind 3 keeps per-point stencils in Level::p2s instead
   bool                  stored = false;
   struct P2Coefficient
   {
      int      kind  = 0; // 0: k0 exp(alpha (1 - |x|)) (1 + beta xi_e), 1: 1 + x/2 + y^2/4
      double   k0    = 1.0, alpha = 0.0, beta = 0.0;
      uint64_t seed  = 0;
   } p2k;
   double                k0   = 1.0;
   std::vector< double > modesHost; // nModes x 5: kappa_x, kappa_y, kappa_z, phi, a
   double*               modes  = nullptr;
   int                   nModes = 0;
   // kind 1 with a polynomial coefficient (kPolyDeg >= 0): k = sum_i kPoly[i] x^a y^b z^c, monomials of degree <=
   // kPolyDeg in the reference's MonomialBasis3D order (device copy; the Fourier fields above are unused)
   int                   kPolyDeg = -1;
   std::vector< double > kPolyHost;
   double*               kPoly = nullptr;
   // kind 1 after hyg_op_interpolate_stencils (P1SurrogateOperator): the polynomial degree of the cell-interior
   // stencils (Level::surPoly), -1 without
   int                   surDegree = -1;
   // kind 1 as the blended Laplacian of a shell mesh (P1ElementwiseBlendingLaplaceOperator): weights in Level::bw
   bool                  blend = false;
   struct Level
   {
      std::vector< double > cellStHost;  // nC x 16: 15 weights, inverse centre weight
      std::vector< double > entryWHost;  // nEntries x 16: one-sided weights of each (primitive, cell)
      std::vector< double > primInvHost; // nPrims: inverse centre weight (sum over incident cells)
      double*               cellSt  = nullptr;
      float*                cellStF = nullptr;
      double*               entryW  = nullptr;
      float*                entryWF = nullptr;
      double*               primInv = nullptr;
      float*                primInvF = nullptr;
      // one-sided face partial sums (2 sides x (N+1)^2 values per face), FP64 and FP32
      double*               facePart  = nullptr;
      float*                facePartF = nullptr;
      // elementwise operator: cell geometry, per-element coefficient integrals (6 types x S(N) per
      // cell), one-sided sums of edge and vertex points (N+1 per primitive entry), inverse diagonal
      // (a function array with replicas)
      hyg::ElemCell*        elem    = nullptr;
      double*               kb      = nullptr;
      float*                kbF     = nullptr;
      // Fourier coefficient: per cell and mode the phase terms (theta0, gx, gy, gz, amplitude) and the cos/sin of the
      // 30 quadrature offsets of the six micro-cell types (built once at operator setup, read by kKBar)
      double*               kbPhase = nullptr;
      double*               kbTab   = nullptr;
      float*                kbTabF  = nullptr;
      double*               lowPart = nullptr;
      float*                lowPartF = nullptr;
      double*               dinv    = nullptr;
      float*                dinvF   = nullptr;
      // P2: element matrices (6 types x 10 x 10 per cell) of the micro-cells of level lattice-1;
      // variable coefficient: per cell and type the 11 quadrature-point matrices w_q |det J| grad phi grad phi^T
      // (11 x 10 x 10) and per micro-cell the coefficient at its 11 quadrature points
      double*               p2el    = nullptr;
      float*                p2elF   = nullptr;
      double*               p2kq    = nullptr;
      float*                p2kqF   = nullptr;
      // constant P2 operator: merged stencil weights of cell-interior points, per cell and parity class
      // (8 x kP2StencilMax, offsets of p2StencilOffsets)
      double*               p2st    = nullptr;
      float*                p2stF   = nullptr;
      // one-sided stencils of the cell-face points, per cell, face and class (4 x 8 x kP2StencilMax, offsets
      // of p2StencilOffsets( cls, face ))
      double*               p2fst   = nullptr;
      float*                p2fstF  = nullptr;
      // variable-coefficient P2 operator after hyg_op_store_local_matrices: merged stencils of every cell-interior
      // point and one-sided stencils of every cell-face point (per cell p2StoredCellSize(N) values)
      double*               p2s     = nullptr;
      float*                p2sF    = nullptr;
      // surrogate (kind 1 after hyg_op_interpolate_stencils): per cell the polynomial of every stencil direction
      // (15 x monomial count, direction-major) in the cell-local coordinates (x, y, z) / N, and the cell-array indices
      // of the points on a cell's faces, edges and vertices (the points the exact stencils keep)
      double*               surPoly = nullptr;
      int*                  surBnd  = nullptr;
      int                   surNBnd = 0;
      // constant P1 operator: tables of the reference-order SOR (hyg_smooth_sor_ordered), built on first use
      std::unique_ptr< SorOrderTables > sorOrder;
      // blended Laplacian (kind 1 with blend): the 15 one-sided stencil weights of every point of every cell, as
      // summed from the cell's micro-elements (per cell, direction-major: [cell][direction][point]), FP64 and FP32
      double*               bw  = nullptr;
      float*                bwF = nullptr;
   };
   std::vector< Level > levels;
   ~hyg_op_s();
};
