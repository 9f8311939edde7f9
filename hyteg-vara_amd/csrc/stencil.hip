// gfx950 kernels of the 15-point P1 stencil operations (apply, residual, Jacobi, Chebyshev steps).
//
// Cell interior: z-marching wavefronts (cellZ). A wavefront owns R adjacent rows (y0 .. y0+R-1) and
// a 64-point x-window (x0-1 .. x0+62, outputs x0 .. x0+61) of one macro-cell and marches up the
// planes z0 .. z1-1. A five-plane register ring holds rows y0-1 .. y0+R of planes z-1, z, z+1 and
// has planes z+2, z+3 in flight, so each plane step issues R+2 coalesced row loads (plus R rows of the
// epilogue input) and no barrier. Row and plane offsets are wave-uniform and advanced incrementally
// in scalar registers; loads and stores go through per-cell buffer descriptors with a uniform byte
// offset, lanes past a row end take an out-of-range offset (they read 0 and fetch nothing), and the
// stores of lanes without an output are dropped the same way, so a step has no branch around a
// memory operation. The x-1 / x+1 neighbours come from the adjacent lanes: the terms needing lane-1
// (W, NW, TW, BNW) and lane+1 (E, SE, TSE, BE) are summed per lane first and moved with one DPP shift
// each.
//
// Macro-faces. A face value is the sum of one-sided stencils of its (one or two) cells. Each cell
// computes its own one-sided sums where its rows are already streamed: the x = 0 face is lane 0 of
// the windows starting at a row start (per-lane weights), the diagonal face x+y+z = N is the last
// point of each row (a second weight set, evaluated by the wavefronts whose window holds row ends);
// both are gathered in registers over the march and stored once per segment. The z = 0 and y = 0
// faces are whole planes/rows and are computed by faceFlat with coalesced loads. kFaceCombine adds
// the partial sums of the two sides, applies the epilogue and writes every replica.
//
// Macro-edges and -vertices (lower): one lane per point gathers the one-sided stencils of all
// incident cells.

#include "sweep.hpp"

namespace hyg {

// ------------------------------------------------------------------ launch
// cell-interior configuration of a level, value type and stencil mode: set by the tuner, else
// HYG_CELL_CFG, else 2 rows x 1 point per lane
int cellConfig( LevelGeom& g, int dtype, int mode )
{
   int& c = g.cellCfg[dtype == 8 ? 0 : 1][mode];
   if ( c < 0 )
   {
      const char* e = getenv( "HYG_CELL_CFG" );
      c             = e ? std::max( 0, std::min( kNumCellCfgs - 1, atoi( e ) ) ) : 1;
   }
   return c;
}
void setCellConfig( LevelGeom& g, int dtype, int mode, int cfg ) { g.cellCfg[dtype == 8 ? 0 : 1][mode] = cfg; }

template < typename T, int MODE >
static void stencilT( const MeshCtx& m, LevelGeom& g, const hyg_op_s::Level& ol, int L, const StencilCall& c, hipStream_t s )
{
   constexpr bool dbl  = std::is_same< T, double >::value;
   const int      nC   = m.mesh->nC();
   const T*       cst  = dbl ? (const T*) ol.cellSt : (const T*) ol.cellStF;
   const T*       ew   = dbl ? (const T*) ol.entryW : (const T*) ol.entryWF;
   const T*       pi   = dbl ? (const T*) ol.primInv : (const T*) ol.primInvF;
   T*             part = dbl ? (T*) ol.facePart : (T*) ol.facePartF;
   const int      cfg  = cellConfig( g, dbl ? 8 : 4, MODE );
   const T*       x    = (const T*) c.x;
   const T*       in1  = (const T*) c.in1;
   T*             out0 = (T*) c.out0;
   T*             out1 = (T*) c.out1;
   // 1. cell interiors with the one-sided sums of the x = 0 and diagonal faces, the one-sided sums
   // of the z = 0 and y = 0 faces, and macro-edges and -vertices: one launch
   const int nLow = g.firstFaceChunk, nFace = g.nChunks - g.firstFaceChunk;
   SweepArgs< T > p;
   p.segs    = g.zSegs[cfg];
   p.nSegs   = g.nZSegs[cfg];
   p.nC      = nC;
   p.L       = L;
   p.flag    = c.flag;
   p.stride  = g.stride;
   p.cellSt  = cst;
   p.cfs     = m.cellFaces;
   p.entryW  = ew;
   p.primInv = pi;
   p.part    = part;
   p.x       = x;
   p.in1     = in1;
   p.out0    = out0;
   p.out1    = out1;
   p.a       = T( c.a );
   p.items   = g.flatItems;
   p.nItems  = nFace > 0 ? g.nFlatItems : 0;
   p.chunks  = g.chunks;
   p.nChunks = nLow;
   p.prims   = m.prims;
   p.entries = m.entries;
   p.gC      = p.nSegs > 0 ? xcdGrid( ( ( p.nSegs + 3 ) / 4 ) * nC ) : 0;
   p.gF      = p.nItems > 0 ? xcdGrid( ceilDiv( p.nItems, 4 ) ) : 0;
   const int gL = nLow > 0 ? xcdGrid( ceilDiv( nLow, kBlock / 64 ) ) : 0;
   if ( p.gC + p.gF + gL > 0 )
   {
      decltype( &kSweep< T, MODE, 1, 1, 2 > ) kf;
      switch ( cfg )
      {
      case 0: kf = kSweep< T, MODE, 1, 1, 2 >; break;
      case 1: kf = kSweep< T, MODE, 2, 1, 2 >; break;
      case 2: kf = kSweep< T, MODE, 4, 1, 2 >; break;
      case 3: kf = kSweep< T, MODE, 1, 1, 4 >; break;
      case 4: kf = kSweep< T, MODE, 2, 1, 4 >; break;
      case 5: kf = kSweep< T, MODE, 4, 1, 4 >; break;
      default: kf = kSweep< T, MODE, 8, 1, 2 >; break;
      }
      hipLaunchKernelGGL( kf, dim3( p.gC + p.gF + gL ), dim3( kBlock ), 0, s, p );
   }
   // 2. face values from the partial sums
   if ( nFace > 0 )
      hipLaunchKernelGGL( ( kFaceCombine< T, MODE > ), dim3( xcdGrid( ceilDiv( nFace, kBlock / 64 ) ) ), dim3( kBlock ), 0, s,
                          g.chunks + g.firstFaceChunk, nFace, m.prims, m.entries, m.mesh->nV() + m.mesh->nE(), part, pi,
                          (const T*) nullptr, L, g.stride, c.flag, x, in1, out0, out1, T( c.a ) );
}

// Multi-colour SOR sweep (forward: colours 0 .. nColours-1, backward: the reverse order). Each colour pass is the
// Jacobi update restricted to the points of that colour, written in place: the sweep launch (cell interiors,
// one-sided face sums, macro-edges and -vertices) and the face combine. Points of one colour are not coupled, so
// a pass reads only values it does not write (the centre value aside, which each point reads before its own
// store); outputs of other colours are computed and dropped.
template < typename T >
static void sorT( const MeshCtx& m, LevelGeom& g, const hyg_op_s::Level& ol, int L, void* xv, const void* bv, double relax, int flag,
                  bool backwards, hipStream_t s )
{
   constexpr bool dbl = std::is_same< T, double >::value;
   const int      cfg = 1; // 2 rows per wave
   const int      nLow = g.firstFaceChunk, nFace = g.nChunks - g.firstFaceChunk;
   T*             part = dbl ? (T*) ol.facePart : (T*) ol.facePartF;
   const T*       pi   = dbl ? (const T*) ol.primInv : (const T*) ol.primInvF;
   SweepArgs< T > p;
   p.segs    = g.zSegs[cfg];
   p.nSegs   = g.nZSegs[cfg];
   p.nC      = m.mesh->nC();
   p.L       = L;
   p.flag    = flag;
   p.stride  = g.stride;
   p.cellSt  = dbl ? (const T*) ol.cellSt : (const T*) ol.cellStF;
   p.cfs     = m.cellFaces;
   p.entryW  = dbl ? (const T*) ol.entryW : (const T*) ol.entryWF;
   p.primInv = pi;
   p.part    = part;
   p.x       = (const T*) xv;
   p.in1     = (const T*) bv;
   p.out0    = (T*) xv;
   p.out1    = (T*) xv;
   p.a       = T( relax );
   p.items   = g.flatItems;
   p.nItems  = nFace > 0 ? g.nFlatItems : 0;
   p.chunks  = g.chunks;
   p.nChunks = nLow;
   p.prims   = m.prims;
   p.entries = m.entries;
   p.gC      = p.nSegs > 0 ? xcdGrid( ( ( p.nSegs + 3 ) / 4 ) * p.nC ) : 0;
   p.gF      = p.nItems > 0 ? xcdGrid( ceilDiv( p.nItems, 4 ) ) : 0;
   p.cinfo   = m.cellInfo;
   p.sorM    = m.nColours;
   const int gL = nLow > 0 ? xcdGrid( ceilDiv( nLow, kBlock / 64 ) ) : 0;
   for ( int q = 0; q < m.nColours; ++q )
   {
      p.sorK = backwards ? m.nColours - 1 - q : q;
      if ( p.gC + p.gF + gL > 0 )
         hipLaunchKernelGGL( ( kSweep< T, MODE_SOR, 2, 1, 2 > ), dim3( p.gC + p.gF + gL ), dim3( kBlock ), 0, s, p );
      if ( nFace > 0 )
         hipLaunchKernelGGL( ( kFaceCombineSor< T > ), dim3( xcdGrid( ceilDiv( nFace, kBlock / 64 ) ) ), dim3( kBlock ), 0, s,
                             g.chunks + g.firstFaceChunk, nFace, m.prims, m.entries, m.mesh->nV() + m.mesh->nE(), part, pi,
                             m.cellInfo, p.sorK, p.sorM, L, g.stride, flag, (T*) xv, (const T*) bv, T( relax ) );
   }
}

void launchSor( const MeshCtx& m, LevelGeom& g, const hyg_op_s::Level& ol, int L, int dtype, void* x, const void* b, double relax,
                int flag, bool backwards, hipStream_t s )
{
   if ( dtype == 8 )
      sorT< double >( m, g, ol, L, x, b, relax, flag, backwards, s );
   else
      sorT< float >( m, g, ol, L, x, b, relax, flag, backwards, s );
}

template < typename T >
static void stencilMode( const MeshCtx& m, LevelGeom& g, const hyg_op_s::Level& ol, int L, const StencilCall& c, hipStream_t s )
{
   switch ( c.mode )
   {
   case MODE_APPLY:
      stencilT< T, MODE_APPLY >( m, g, ol, L, c, s );
      break;
   case MODE_APPLY_ADD:
      stencilT< T, MODE_APPLY_ADD >( m, g, ol, L, c, s );
      break;
   case MODE_RESIDUAL:
      stencilT< T, MODE_RESIDUAL >( m, g, ol, L, c, s );
      break;
   case MODE_JACOBI:
      stencilT< T, MODE_JACOBI >( m, g, ol, L, c, s );
      break;
   case MODE_CHEB0:
      stencilT< T, MODE_CHEB0 >( m, g, ol, L, c, s );
      break;
   default:
      stencilT< T, MODE_CHEBK >( m, g, ol, L, c, s );
      break;
   }
}

void launchStencil( const MeshCtx& m, LevelGeom& g, const hyg_op_s::Level& ol, int L, int dtype, const StencilCall& c,
                    hipStream_t s )
{
   if ( dtype == 8 )
      stencilMode< double >( m, g, ol, L, c, s );
   else
      stencilMode< float >( m, g, ol, L, c, s );
}

} // namespace hyg
