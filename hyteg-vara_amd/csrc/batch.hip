// Batched independent problems: one persistent launch per grid operation for all problems of a batch.
//
// A batch step is one grid operation (a stencil sweep, its face/edge/vertex combine, a restriction or
// prolongation, a pointwise update, a dot product or a value-type conversion) applied to every problem of
// the batch, each at its own level, on its own mesh, with its own operator and functions. The step's task
// table holds one entry per problem (two for operations made of two independent kernels): the kernel
// arguments of that problem and the range of virtual blocks it takes. Each block finds the task of its virtual
// block by binary search on the range starts and runs the same kernel body the single-problem launch runs with
// the same block index, so every problem's arithmetic and order of operations are those of the unbatched call.
// The stencil sweeps launch one block per virtual block; the lighter operations run a fixed grid of kBatchGrid
// blocks that walks the virtual blocks (block b takes b, b + grid, ...).
#include "elementwise.hpp"
#include "kernels_dev.hpp"
#include "sweep.hpp"

#include <cstring>

namespace hyg {

constexpr int kBatchGrid = 2048; // blocks of a persistent batched launch (8 per CU)

// --------------------------------------------------------------------------------------------- task tables
template < typename T >
struct ElemArgs
{
   int                 L, tilesPerCell;
   i64                 stride;
   const ElemCell*     cells;
   const T*            kb;
   const CellInfo*     cinfo;
   const CellFaceSlot* cfs;
   const LowerPrim*    prims;
   const LowerEntry*   entries;
   T*                  part;
   T*                  lowPart;
   const T*            dinv;
   int                 flag;
   const T*            x;
   const T*            in1;
   T*                  out0;
   T*                  out1;
   T                   a;
};

// constant-coefficient sweep (cell interiors, flat faces, edges and vertices) of one problem
template < typename T >
struct SweepTask
{
   int            kind, first;
   SweepArgs< T > sw;
};
// elementwise operator (stored coefficient integrals) of one problem
template < typename T >
struct ElemTask
{
   int           kind, first;
   ElemArgs< T > el;
};

// kind 0: macro-face points from the one-sided sums; kind 1: macro-edge and -vertex points (elementwise operator)
template < typename T >
struct CombineTask
{
   int               kind, first;
   const LowerChunk* chunks;
   int               nChunks, firstFacePrim, L, flag;
   const LowerPrim*  prims;
   const LowerEntry* entries;
   const T*          part;
   const T*          primInv;
   const T*          dinv;
   i64               stride;
   const T*          x;
   const T*          in1;
   T*                out0;
   T*                out1;
   T                 a;
};

// kind 0: restriction of cell-interior coarse points; 1: restriction of face, edge and vertex points;
// 2: prolongation (replace or add)
template < typename T >
struct TransferTask
{
   int               kind, first;
   const PlaneTile*  tiles;
   int               tile0, nTiles, nC, L, flag, add;
   i64               strideF, strideC;
   const LowerChunk* chunks;
   int               nChunks;
   const LowerPrim*  prims;
   const LowerEntry* entries;
   const CellInfo*   cinfo;
   T*                dst;
   const T*          src;
};

// kind 0: linear combination dst = (acc ? dst : 0) + sum a_k s_k (or the constant `value`); kind 1: block partial
// sums of a dot product into `partials`
template < typename T >
struct VecTask
{
   int              kind, first;
   const PlaneTile* tiles;
   int              nTiles, nC, L, flag, nsrc, acc;
   i64              stride;
   const CellInfo*  cinfo;
   T*               dst;
   const T*         s0;
   const T*         s1;
   const T*         s2;
   T                a0, a1, a2, value;
   double*          partials;
};

struct DotFinalTask
{
   const double* partials;
   int           n;
   double*       result;
};

struct ConvertTask
{
   int         first;
   i64         n;
   void*       dst;
   const void* src;
};
constexpr int kConvertPerBlock = kBlock * 16;

// ------------------------------------------------------------------------------------------- batched kernels
template < typename Task >
__device__ __forceinline__ int findTask( const Task* __restrict__ t, int n, int vb )
{
   int lo = 0, hi = n - 1;
   while ( lo < hi )
   {
      const int mid = ( lo + hi + 1 ) >> 1;
      if ( t[mid].first <= vb )
         lo = mid;
      else
         hi = mid - 1;
   }
   return lo;
}

// The two stencil kernels are separate launches (a batch with both operator kinds issues both): one kernel
// holding both bodies needs the register budget of both at once (253 VGPRs against 128-214 and 120) and runs
// at a quarter of the occupancy. Their grids take one block per work item of every problem, as the
// single-problem launches do; the lighter operations below walk their work with a fixed grid.
template < typename T, int MODE, int R, int P, int D >
__global__ __launch_bounds__( kBlock ) void kBatchSweep( const SweepTask< T >* __restrict__ tasks, int nTasks, int total )
{
   const int vb = int( blockIdx.x );
   if ( vb >= total )
      return;
   const SweepTask< T >& t = tasks[findTask( tasks, nTasks, vb )];
   bSweep< T, MODE, R, P, D >( vb - t.first, t.sw );
}

template < typename T, int MODE >
__global__ __launch_bounds__( kBlock ) void kBatchElem( const ElemTask< T >* __restrict__ tasks, int nTasks, int total )
{
   const int vb = int( blockIdx.x );
   if ( vb >= total )
      return;
   const ElemTask< T >& t = tasks[findTask( tasks, nTasks, vb )];
   const ElemArgs< T >& e = t.el;
   bElemApply< T, MODE, false >( vb - t.first, e.L, e.tilesPerCell, e.stride, e.cells, e.kb, e.cinfo, e.cfs, e.prims, e.entries, e.part,
                                 e.lowPart, e.dinv, e.flag, e.x, e.in1, e.out0, e.out1, e.a );
}

template < typename T, int MODE >
__global__ __launch_bounds__( kBlock ) void kBatchCombine( const CombineTask< T >* __restrict__ tasks, int nTasks, int total )
{
   for ( int vb = int( blockIdx.x ); vb < total; vb += int( gridDim.x ) )
   {
      const CombineTask< T >& t = tasks[findTask( tasks, nTasks, vb )];
      const int               b = vb - t.first;
      if ( t.kind == 0 )
         bFaceCombine< T, MODE >( b, t.chunks, t.nChunks, t.prims, t.entries, t.firstFacePrim, t.part, t.primInv, t.dinv, t.L,
                                  t.stride, t.flag, t.x, t.in1, t.out0, t.out1, t.a );
      else
         bLowCombine< T, MODE >( b, t.chunks, t.nChunks, t.prims, t.entries, t.part, t.dinv, t.L, t.stride, t.flag, t.x, t.in1,
                                 t.out0, t.out1, t.a );
   }
}

template < typename T >
__global__ __launch_bounds__( kBlock ) void kBatchTransfer( const TransferTask< T >* __restrict__ tasks, int nTasks, int total )
{
   for ( int vb = int( blockIdx.x ); vb < total; vb += int( gridDim.x ) )
   {
      const TransferTask< T >& t = tasks[findTask( tasks, nTasks, vb )];
      const int                b = vb - t.first;
      if ( t.kind == 0 )
         bRestrictCell< T >( b, t.tiles, t.tile0, t.nTiles, t.nC, t.L, t.strideC, t.strideF, t.dst, t.src );
      else if ( t.kind == 1 )
         bRestrictLower< T >( b, t.chunks, t.nChunks, t.prims, t.entries, t.cinfo, t.L, t.strideC, t.strideF, t.flag, t.dst, t.src );
      else
         bProlongate< T >( b, t.tiles, t.nTiles, t.nC, t.L, t.strideF, t.strideC, t.cinfo, t.flag, t.dst, t.src, t.add );
   }
}

template < typename T >
__global__ __launch_bounds__( kBlock ) void kBatchVec( const VecTask< T >* __restrict__ tasks, int nTasks, int total )
{
   for ( int vb = int( blockIdx.x ); vb < total; vb += int( gridDim.x ) )
   {
      const VecTask< T >& t = tasks[findTask( tasks, nTasks, vb )];
      const int           b = vb - t.first;
      if ( t.kind == 0 )
         bLincomb< T >( b, t.tiles, t.nTiles, t.nC, t.L, t.stride, t.cinfo, t.flag, t.dst, t.nsrc, t.s0, t.s1, t.s2, t.a0, t.a1, t.a2,
                        t.acc, t.value );
      else
      {
         bDot< T >( b, t.tiles, t.nTiles, t.nC, t.L, t.stride, t.cinfo, t.flag, t.s0, t.s1, t.partials );
         // the block's reduction buffer is reused by its next virtual block
         __syncthreads();
      }
   }
}

constexpr int kBatchFinalBlock = 1024;

__global__ __launch_bounds__( kBatchFinalBlock ) void kBatchDotFinal( const DotFinalTask* __restrict__ tasks )
{
   const DotFinalTask t = tasks[blockIdx.x];
   double             v = 0.0;
   for ( int k = int( threadIdx.x ); k < t.n; k += kBatchFinalBlock )
      v += t.partials[k];
   v = blockSum< kBatchFinalBlock >( v );
   if ( threadIdx.x == 0 )
      t.result[0] = v;
}

template < typename D, typename S >
__global__ __launch_bounds__( kBlock ) void kBatchConvert( const ConvertTask* __restrict__ tasks, int nTasks, int total )
{
   for ( int vb = int( blockIdx.x ); vb < total; vb += int( gridDim.x ) )
   {
      const ConvertTask& t   = tasks[findTask( tasks, nTasks, vb )];
      const i64          k0  = i64( vb - t.first ) * kConvertPerBlock;
      D*                 dst = static_cast< D* >( t.dst );
      const S*           src = static_cast< const S* >( t.src );
      for ( i64 k = k0 + threadIdx.x; k < k0 + kConvertPerBlock && k < t.n; k += kBlock )
         dst[k] = D( src[k] );
   }
}

// ---------------------------------------------------------------------------------------------- plan (host)
namespace {

template < typename Task >
void appendTask( BatchPlan::Step& s, Task t, int blocks )
{
   if ( blocks <= 0 )
      return;
   t.first = s.total;
   s.total += blocks;
   const size_t o = s.host.size();
   s.host.resize( o + sizeof( Task ) );
   std::memcpy( s.host.data() + o, &t, sizeof( Task ) );
   ++s.nTasks;
}

template < typename T >
SweepArgs< T > sweepArgs( const MeshCtx& m, LevelGeom& g, const hyg_op_s::Level& ol, int L, const StencilCall& c, int cfg )
{
   constexpr bool dbl = std::is_same< T, double >::value;
   const int      nC  = m.mesh->nC();
   const int      nLow = g.firstFaceChunk, nFace = g.nChunks - g.firstFaceChunk;
   SweepArgs< T > p;
   p.segs    = g.zSegs[cfg];
   p.nSegs   = g.nZSegs[cfg];
   p.nC      = nC;
   p.L       = L;
   p.flag    = c.flag;
   p.stride  = g.stride;
   p.cellSt  = dbl ? (const T*) ol.cellSt : (const T*) ol.cellStF;
   p.cfs     = m.cellFaces;
   p.entryW  = dbl ? (const T*) ol.entryW : (const T*) ol.entryWF;
   p.primInv = dbl ? (const T*) ol.primInv : (const T*) ol.primInvF;
   p.part    = dbl ? (T*) ol.facePart : (T*) ol.facePartF;
   p.x       = (const T*) c.x;
   p.in1     = (const T*) c.in1;
   p.out0    = (T*) c.out0;
   p.out1    = (T*) c.out1;
   p.a       = T( c.a );
   p.items   = g.flatItems;
   p.nItems  = nFace > 0 ? g.nFlatItems : 0;
   p.chunks  = g.chunks;
   p.nChunks = nLow;
   p.prims   = m.prims;
   p.entries = m.entries;
   p.gC      = p.nSegs > 0 ? xcdGrid( ( ( p.nSegs + 3 ) / 4 ) * nC ) : 0;
   p.gF      = p.nItems > 0 ? xcdGrid( ceilDiv( p.nItems, 4 ) ) : 0;
   return p;
}

template < typename T >
void addStencilT( BatchPlan& plan, int mode, const std::vector< BatchPlan::Item >& items, int cfg )
{
   constexpr bool   dbl = std::is_same< T, double >::value;
   BatchPlan::Step  st{}, el{}, cb{};
   st.phase = BatchPlan::STENCIL;
   st.cfg   = cfg;
   el.phase = BatchPlan::ELEM;
   cb.phase = BatchPlan::COMBINE;
   st.mode = el.mode = cb.mode = mode;
   st.dtype = el.dtype = cb.dtype = dbl ? 8 : 4;
   for ( const auto& it : items )
   {
      const MeshCtx&          m  = *it.m;
      LevelGeom&              g  = *it.g;
      const hyg_op_s&         op = *it.op;
      const hyg_op_s::Level&  ol = op.levels[it.L];
      const StencilCall&      c  = it.c;
      const int               nC = m.mesh->nC(), N = 1 << it.L, w = N + 1;
      const int               nLow = g.firstFaceChunk, nFace = g.nChunks - g.firstFaceChunk;
      SweepTask< T >          t{};
      ElemTask< T >           u{};
      CombineTask< T >        f{};
      f.chunks        = g.chunks + g.firstFaceChunk;
      f.nChunks       = nFace;
      f.firstFacePrim = m.mesh->nV() + m.mesh->nE();
      f.L             = it.L;
      f.flag          = c.flag;
      f.prims         = m.prims;
      f.entries       = m.entries;
      f.stride        = g.stride;
      f.x             = (const T*) c.x;
      f.in1           = (const T*) c.in1;
      f.out0          = (T*) c.out0;
      f.out1          = (T*) c.out1;
      f.a             = T( c.a );
      if ( op.kind == 0 )
      {
         t.kind          = 0;
         t.sw            = sweepArgs< T >( m, g, ol, it.L, c, cfg );
         const int gL    = nLow > 0 ? xcdGrid( ceilDiv( nLow, kBlock / 64 ) ) : 0;
         appendTask( st, t, t.sw.gC + t.sw.gF + gL );
         f.kind    = 0;
         f.part    = dbl ? (const T*) ol.facePart : (const T*) ol.facePartF;
         f.primInv = dbl ? (const T*) ol.primInv : (const T*) ol.primInvF;
         f.dinv    = nullptr;
         appendTask( cb, f, nFace > 0 ? xcdGrid( ceilDiv( nFace, kBlock / 64 ) ) : 0 );
      }
      else
      {
         ElemArgs< T >& e = u.el;
         u.kind           = 1;
         e.L              = it.L;
         e.tilesPerCell   = ceilDiv( S3i( w ), kBlock );
         e.stride         = g.stride;
         e.cells          = ol.elem;
         e.kb             = dbl ? (const T*) ol.kb : (const T*) ol.kbF;
         e.cinfo          = m.cellInfo;
         e.cfs            = m.cellFaces;
         e.prims          = m.prims;
         e.entries        = m.entries;
         e.part           = dbl ? (T*) ol.facePart : (T*) ol.facePartF;
         e.lowPart        = dbl ? (T*) ol.lowPart : (T*) ol.lowPartF;
         e.dinv           = dbl ? (const T*) ol.dinv : (const T*) ol.dinvF;
         e.flag           = c.flag;
         e.x              = (const T*) c.x;
         e.in1            = (const T*) c.in1;
         e.out0           = (T*) c.out0;
         e.out1           = (T*) c.out1;
         e.a              = T( c.a );
         appendTask( el, u, nC * e.tilesPerCell );
         f.kind    = 0;
         f.part    = e.part;
         f.primInv = nullptr;
         f.dinv    = e.dinv;
         appendTask( cb, f, nFace > 0 ? xcdGrid( ceilDiv( nFace, kBlock / 64 ) ) : 0 );
         CombineTask< T > lo = f;
         lo.kind             = 1;
         lo.chunks           = g.chunks;
         lo.nChunks          = nLow;
         lo.part             = e.lowPart;
         appendTask( cb, lo, nLow > 0 ? ceilDiv( nLow, kBlock / 64 ) : 0 );
      }
   }
   plan.steps.push_back( std::move( st ) );
   plan.steps.push_back( std::move( el ) );
   plan.steps.push_back( std::move( cb ) );
}

template < typename T >
void addRestrictT( BatchPlan& plan, const std::vector< BatchPlan::Transfer >& items )
{
   BatchPlan::Step s{};
   s.phase = BatchPlan::TRANSFER;
   s.dtype = sizeof( T );
   for ( const auto& it : items )
   {
      const MeshCtx&    m  = *it.m;
      const LevelGeom&  gc = *it.g;
      const int         nC = m.mesh->nC();
      TransferTask< T > t{};
      t.tiles   = gc.tiles;
      t.tile0   = gc.interiorTile0;
      t.nTiles  = gc.nInteriorTiles;
      t.nC      = nC;
      t.L       = it.L;
      t.flag    = it.flag;
      t.strideC = gc.stride;
      t.strideF = cellStride( it.L + 1 );
      t.chunks  = gc.chunks;
      t.nChunks = gc.nChunks;
      t.prims   = m.prims;
      t.entries = m.entries;
      t.cinfo   = m.cellInfo;
      t.dst     = (T*) it.dst;
      t.src     = (const T*) it.src;
      t.kind    = 0;
      appendTask( s, t, gc.nInteriorTiles > 0 ? xcdGrid( gc.nInteriorTiles * nC ) : 0 );
      t.kind = 1;
      appendTask( s, t, gc.nChunks > 0 ? ceilDiv( gc.nChunks, kBlock / 64 ) : 0 );
   }
   plan.steps.push_back( std::move( s ) );
}

template < typename T >
void addProlongateT( BatchPlan& plan, const std::vector< BatchPlan::Transfer >& items )
{
   BatchPlan::Step s{};
   s.phase = BatchPlan::TRANSFER;
   s.dtype = sizeof( T );
   for ( const auto& it : items )
   {
      const MeshCtx&    m  = *it.m;
      const LevelGeom&  gf = *it.g;
      const int         nC = m.mesh->nC();
      TransferTask< T > t{};
      t.kind    = 2;
      t.tiles   = gf.tiles;
      t.nTiles  = gf.nTiles;
      t.nC      = nC;
      t.L       = it.L;
      t.flag    = it.flag;
      t.add     = it.add ? 1 : 0;
      t.strideF = gf.stride;
      t.strideC = cellStride( it.L - 1 );
      t.cinfo   = m.cellInfo;
      t.dst     = (T*) it.dst;
      t.src     = (const T*) it.src;
      appendTask( s, t, xcdGrid( gf.nTiles * nC ) );
   }
   plan.steps.push_back( std::move( s ) );
}

template < typename T >
void addVecT( BatchPlan& plan, const std::vector< BatchPlan::Vec >& items )
{
   BatchPlan::Step s{};
   s.phase = BatchPlan::VEC;
   s.dtype = sizeof( T );
   for ( const auto& it : items )
   {
      const MeshCtx&   m  = *it.m;
      const LevelGeom& g  = *it.g;
      const int        nC = m.mesh->nC();
      VecTask< T >     t{};
      t.kind     = it.dot ? 1 : 0;
      t.tiles    = g.tiles;
      t.nTiles   = g.nTiles;
      t.nC       = nC;
      t.L        = it.L;
      t.flag     = it.flag;
      t.nsrc     = it.nsrc;
      t.acc      = it.acc ? 1 : 0;
      t.stride   = g.stride;
      t.cinfo    = m.cellInfo;
      t.dst      = (T*) it.dst;
      t.s0       = (const T*) it.src[0];
      t.s1       = (const T*) it.src[1];
      t.s2       = (const T*) it.src[2];
      t.a0       = T( it.a[0] );
      t.a1       = T( it.a[1] );
      t.a2       = T( it.a[2] );
      t.value    = T( it.value );
      t.partials = it.partials;
      // the dot's partial sums are indexed by the block, so its grid is exact
      appendTask( s, t, it.dot ? g.nTiles * nC : xcdGrid( g.nTiles * nC ) );
   }
   plan.steps.push_back( std::move( s ) );
}

template < typename T, int MODE, int R, int D >
void launchStencilCfg( const BatchPlan::Step& s, hipStream_t st )
{
   hipLaunchKernelGGL( ( kBatchSweep< T, MODE, R, 1, D > ), dim3( s.total ), dim3( kBlock ), 0, st, (const SweepTask< T >*) s.dev,
                       s.nTasks, s.total );
}

// the seven cell-interior configurations (rows per wavefront R, planes in flight D) of kSweep
template < typename T, int MODE >
void launchStencilMode( const BatchPlan::Step& s, hipStream_t st )
{
   switch ( s.cfg )
   {
   case 0: launchStencilCfg< T, MODE, 1, 2 >( s, st ); break;
   case 1: launchStencilCfg< T, MODE, 2, 2 >( s, st ); break;
   case 2: launchStencilCfg< T, MODE, 4, 2 >( s, st ); break;
   case 3: launchStencilCfg< T, MODE, 1, 4 >( s, st ); break;
   case 4: launchStencilCfg< T, MODE, 2, 4 >( s, st ); break;
   case 5: launchStencilCfg< T, MODE, 4, 4 >( s, st ); break;
   default: launchStencilCfg< T, MODE, 8, 2 >( s, st ); break;
   }
}

template < typename T >
void launchElemStep( const BatchPlan::Step& s, hipStream_t st )
{
   const dim3 g( s.total ), b( kBlock );
   switch ( s.mode )
   {
   case MODE_APPLY: hipLaunchKernelGGL( ( kBatchElem< T, MODE_APPLY > ), g, b, 0, st, (const ElemTask< T >*) s.dev, s.nTasks, s.total ); break;
   case MODE_RESIDUAL:
      hipLaunchKernelGGL( ( kBatchElem< T, MODE_RESIDUAL > ), g, b, 0, st, (const ElemTask< T >*) s.dev, s.nTasks, s.total );
      break;
   case MODE_JACOBI: hipLaunchKernelGGL( ( kBatchElem< T, MODE_JACOBI > ), g, b, 0, st, (const ElemTask< T >*) s.dev, s.nTasks, s.total ); break;
   case MODE_CHEB0: hipLaunchKernelGGL( ( kBatchElem< T, MODE_CHEB0 > ), g, b, 0, st, (const ElemTask< T >*) s.dev, s.nTasks, s.total ); break;
   default: hipLaunchKernelGGL( ( kBatchElem< T, MODE_CHEBK > ), g, b, 0, st, (const ElemTask< T >*) s.dev, s.nTasks, s.total ); break;
   }
}

template < typename T >
void launchStencilStep( const BatchPlan::Step& s, hipStream_t st )
{
   switch ( s.mode )
   {
   case MODE_APPLY: launchStencilMode< T, MODE_APPLY >( s, st ); break;
   case MODE_RESIDUAL: launchStencilMode< T, MODE_RESIDUAL >( s, st ); break;
   case MODE_JACOBI: launchStencilMode< T, MODE_JACOBI >( s, st ); break;
   case MODE_CHEB0: launchStencilMode< T, MODE_CHEB0 >( s, st ); break;
   default: launchStencilMode< T, MODE_CHEBK >( s, st ); break;
   }
}

template < typename T >
void launchCombineStep( const BatchPlan::Step& s, hipStream_t st )
{
   const int grid = std::min( s.total, kBatchGrid );
   auto*     tk   = (const CombineTask< T >*) s.dev;
   switch ( s.mode )
   {
   case MODE_APPLY: hipLaunchKernelGGL( ( kBatchCombine< T, MODE_APPLY > ), dim3( grid ), dim3( kBlock ), 0, st, tk, s.nTasks, s.total ); break;
   case MODE_RESIDUAL:
      hipLaunchKernelGGL( ( kBatchCombine< T, MODE_RESIDUAL > ), dim3( grid ), dim3( kBlock ), 0, st, tk, s.nTasks, s.total );
      break;
   case MODE_JACOBI: hipLaunchKernelGGL( ( kBatchCombine< T, MODE_JACOBI > ), dim3( grid ), dim3( kBlock ), 0, st, tk, s.nTasks, s.total ); break;
   case MODE_CHEB0: hipLaunchKernelGGL( ( kBatchCombine< T, MODE_CHEB0 > ), dim3( grid ), dim3( kBlock ), 0, st, tk, s.nTasks, s.total ); break;
   default: hipLaunchKernelGGL( ( kBatchCombine< T, MODE_CHEBK > ), dim3( grid ), dim3( kBlock ), 0, st, tk, s.nTasks, s.total ); break;
   }
}

} // namespace

bool BatchPlan::modeSupported( int mode )
{
   return mode == MODE_APPLY || mode == MODE_RESIDUAL || mode == MODE_JACOBI || mode == MODE_CHEB0 || mode == MODE_CHEBK;
}

void BatchPlan::stencil( int dtype, int mode, const std::vector< Item >& items, int cfg )
{
   if ( dtype == 8 )
      addStencilT< double >( *this, mode, items, cfg );
   else
      addStencilT< float >( *this, mode, items, cfg );
}

void BatchPlan::restrictLevels( int dtype, const std::vector< Transfer >& items )
{
   if ( dtype == 8 )
      addRestrictT< double >( *this, items );
   else
      addRestrictT< float >( *this, items );
}

void BatchPlan::prolongate( int dtype, const std::vector< Transfer >& items )
{
   if ( dtype == 8 )
      addProlongateT< double >( *this, items );
   else
      addProlongateT< float >( *this, items );
}

void BatchPlan::vec( int dtype, const std::vector< Vec >& items )
{
   if ( dtype == 8 )
      addVecT< double >( *this, items );
   else
      addVecT< float >( *this, items );
}

void BatchPlan::dotFinal( const std::vector< DotFinal >& items )
{
   Step s{};
   s.phase = DOTFINAL;
   for ( const auto& it : items )
   {
      DotFinalTask t{ it.partials, it.n, it.result };
      const size_t o = s.host.size();
      s.host.resize( o + sizeof( t ) );
      std::memcpy( s.host.data() + o, &t, sizeof( t ) );
      ++s.nTasks;
   }
   s.total = s.nTasks;
   steps.push_back( std::move( s ) );
}

void BatchPlan::convert( int dstType, int srcType, const std::vector< Convert >& items )
{
   Step s{};
   s.phase = CONVERT;
   s.dtype = dstType;
   s.mode  = srcType;
   for ( const auto& it : items )
   {
      ConvertTask t{ 0, it.n, it.dst, it.src };
      appendTask( s, t, int( ( it.n + kConvertPerBlock - 1 ) / kConvertPerBlock ) );
   }
   steps.push_back( std::move( s ) );
}

hipError_t BatchPlan::upload()
{
   size_t bytes = 0;
   for ( auto& s : steps )
      bytes += ( s.host.size() + 255 ) & ~size_t( 255 );
   if ( bytes == 0 )
      return hipSuccess;
   hipError_t e = hipMalloc( &dev, bytes );
   if ( e != hipSuccess )
   {
      dev = nullptr;
      return e;
   }
   size_t o = 0;
   for ( auto& s : steps )
   {
      s.dev = static_cast< char* >( dev ) + o;
      if ( !s.host.empty() )
      {
         e = hipMemcpy( s.dev, s.host.data(), s.host.size(), hipMemcpyHostToDevice );
         if ( e != hipSuccess )
            return e;
      }
      o += ( s.host.size() + 255 ) & ~size_t( 255 );
   }
   return hipSuccess;
}

void BatchPlan::launch( hipStream_t st, size_t from, size_t to ) const
{
   to = std::min( to, steps.size() );
   for ( size_t k = from; k < to; ++k )
   {
      const Step& s = steps[k];
      if ( s.nTasks == 0 || s.total == 0 )
         continue;
      const int grid = std::min( s.total, kBatchGrid );
      switch ( s.phase )
      {
      case STENCIL:
         if ( s.dtype == 8 )
            launchStencilStep< double >( s, st );
         else
            launchStencilStep< float >( s, st );
         break;
      case ELEM:
         if ( s.dtype == 8 )
            launchElemStep< double >( s, st );
         else
            launchElemStep< float >( s, st );
         break;
      case COMBINE:
         if ( s.dtype == 8 )
            launchCombineStep< double >( s, st );
         else
            launchCombineStep< float >( s, st );
         break;
      case TRANSFER:
         if ( s.dtype == 8 )
            hipLaunchKernelGGL( kBatchTransfer< double >, dim3( grid ), dim3( kBlock ), 0, st, (const TransferTask< double >*) s.dev,
                                s.nTasks, s.total );
         else
            hipLaunchKernelGGL( kBatchTransfer< float >, dim3( grid ), dim3( kBlock ), 0, st, (const TransferTask< float >*) s.dev,
                                s.nTasks, s.total );
         break;
      case VEC:
         if ( s.dtype == 8 )
            hipLaunchKernelGGL( kBatchVec< double >, dim3( grid ), dim3( kBlock ), 0, st, (const VecTask< double >*) s.dev, s.nTasks,
                                s.total );
         else
            hipLaunchKernelGGL( kBatchVec< float >, dim3( grid ), dim3( kBlock ), 0, st, (const VecTask< float >*) s.dev, s.nTasks,
                                s.total );
         break;
      case DOTFINAL:
         hipLaunchKernelGGL( kBatchDotFinal, dim3( s.nTasks ), dim3( kBatchFinalBlock ), 0, st, (const DotFinalTask*) s.dev );
         break;
      default: // CONVERT: dtype = destination, mode = source value size
         if ( s.dtype == 4 && s.mode == 8 )
            hipLaunchKernelGGL( ( kBatchConvert< float, double > ), dim3( grid ), dim3( kBlock ), 0, st, (const ConvertTask*) s.dev,
                                s.nTasks, s.total );
         else if ( s.dtype == 8 && s.mode == 4 )
            hipLaunchKernelGGL( ( kBatchConvert< double, float > ), dim3( grid ), dim3( kBlock ), 0, st, (const ConvertTask*) s.dev,
                                s.nTasks, s.total );
         else if ( s.dtype == 8 )
            hipLaunchKernelGGL( ( kBatchConvert< double, double > ), dim3( grid ), dim3( kBlock ), 0, st, (const ConvertTask*) s.dev,
                                s.nTasks, s.total );
         else
            hipLaunchKernelGGL( ( kBatchConvert< float, float > ), dim3( grid ), dim3( kBlock ), 0, st, (const ConvertTask*) s.dev,
                                s.nTasks, s.total );
         break;
      }
   }
}

BatchPlan::~BatchPlan() { (void) hipFree( dev ); }

} // namespace hyg
