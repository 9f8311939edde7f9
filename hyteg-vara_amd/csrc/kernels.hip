// gfx950 kernels of the P1 matrix-free multigrid path.
//
// Work decomposition. Cell arrays are processed in plane tiles: 256 consecutive values of one
// z-plane (contiguous in memory), one thread per value, so loads and stores of a wave are
// contiguous. The tile list is z-major within a cell and a block's logical index is remapped
// so that each of the 8 XCDs walks a contiguous range of (cell, plane) tiles: the planes
// z-1, z, z+1 read by a stencil tile were just streamed through the same XCD's L2.
// Macro-face, -edge and -vertex points are processed in chunks of up to 64 points of one
// primitive (one wavefront per chunk); their results are written to every incident cell.

#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "kernels_dev.hpp"

namespace hyg {

constexpr int kFinalBlock = 1024;

__global__ __launch_bounds__( kFinalBlock ) void kDotFinal( const double* __restrict__ partials, int n, double* __restrict__ result )
{
   double v = 0.0;
   for ( int k = int( threadIdx.x ); k < n; k += kFinalBlock )
      v += partials[k];
   v = blockSum< kFinalBlock >( v );
   if ( threadIdx.x == 0 )
      result[0] = v;
}

// ------------------------------------------------------------------ reference-layout halos
template < typename T >
__global__ __launch_bounds__( kBlock ) void kHaloGather( const i64* __restrict__ map, i64 n, const T* __restrict__ f, T* __restrict__ out )
{
   for ( i64 k = i64( blockIdx.x ) * kBlock + threadIdx.x; k < n; k += i64( gridDim.x ) * kBlock )
      out[k] = f[map[k]];
}

// each point takes the own value of the primitive it lies inside (vertex, edge interior, face interior), as after the
// reference's downward chain vertex -> edge -> face -> cell, where lower-dimensional values overwrite the borders
template < typename T >
__global__ __launch_bounds__( kBlock ) void kHaloScatter( const int* __restrict__ sumOwn, int nSum, const int* __restrict__ row,
                                                          const i64* __restrict__ src, const i64* __restrict__ rep,
                                                          const T* __restrict__ in, T* __restrict__ f )
{
   for ( int s = int( blockIdx.x ) * kBlock + int( threadIdx.x ); s < nSum; s += int( gridDim.x ) * kBlock )
   {
      const int j = sumOwn[s];
      const T   v = in[src[j]];
      for ( int r = row[j]; r < row[j + 1]; ++r )
         f[rep[r]] = v;
   }
}

// communicateAdditively<Cell, Face/Edge/Vertex> followed by the downward copy: each replica set holds per-cell
// contributions, their sum becomes the value of every replica
template < typename T >
__global__ __launch_bounds__( kBlock ) void kHaloSum( const int* __restrict__ sumOwn, int nSum, const int* __restrict__ row,
                                                      const i64* __restrict__ rep, const int* __restrict__ bc, int flag,
                                                      T* __restrict__ f )
{
   for ( int s = int( blockIdx.x ) * kBlock + int( threadIdx.x ); s < nSum; s += int( gridDim.x ) * kBlock )
   {
      const int j = sumOwn[s];
      if ( !( bc[j] & flag ) )
         continue;
      T v = T( 0 );
      for ( int r = row[j]; r < row[j + 1]; ++r )
         v += f[rep[r]];
      for ( int r = row[j]; r < row[j + 1]; ++r )
         f[rep[r]] = v;
   }
}

void launchHaloGather( const HaloMaps& h, int dtype, const void* f, void* out, hipStream_t st )
{
   const int grid = int( std::min< i64 >( ( h.size + kBlock - 1 ) / kBlock, 16384 ) );
   if ( h.size == 0 )
      return;
   if ( dtype == 8 )
      hipLaunchKernelGGL( kHaloGather< double >, dim3( grid ), dim3( kBlock ), 0, st, h.gather, h.size, (const double*) f, (double*) out );
   else
      hipLaunchKernelGGL( kHaloGather< float >, dim3( grid ), dim3( kBlock ), 0, st, h.gather, h.size, (const float*) f, (float*) out );
}

void launchHaloScatter( const HaloMaps& h, int dtype, const void* in, void* f, hipStream_t st )
{
   const int grid = std::min( ( h.nSum + kBlock - 1 ) / kBlock, 16384 );
   if ( h.nSum == 0 )
      return;
   if ( dtype == 8 )
      hipLaunchKernelGGL( kHaloScatter< double >, dim3( grid ), dim3( kBlock ), 0, st, h.sumOwn, h.nSum, h.ownRow, h.ownSrc, h.replica,
                          (const double*) in, (double*) f );
   else
      hipLaunchKernelGGL( kHaloScatter< float >, dim3( grid ), dim3( kBlock ), 0, st, h.sumOwn, h.nSum, h.ownRow, h.ownSrc, h.replica,
                          (const float*) in, (float*) f );
}

void launchHaloSum( const HaloMaps& h, int dtype, int flag, void* f, hipStream_t st )
{
   const int grid = std::min( ( h.nSum + kBlock - 1 ) / kBlock, 16384 );
   if ( h.nSum == 0 )
      return;
   if ( dtype == 8 )
      hipLaunchKernelGGL( kHaloSum< double >, dim3( grid ), dim3( kBlock ), 0, st, h.sumOwn, h.nSum, h.ownRow, h.replica, h.ownPrimBc,
                          flag, (double*) f );
   else
      hipLaunchKernelGGL( kHaloSum< float >, dim3( grid ), dim3( kBlock ), 0, st, h.sumOwn, h.nSum, h.ownRow, h.replica, h.ownPrimBc,
                          flag, (float*) f );
}

// ------------------------------------------------------------------ launchers
template < typename T >
static void lincombT( const MeshCtx& m, const LevelGeom& g, int L, void* dst, int nsrc, const double* s, const void* const* src,
                      bool accumulate, double value, int flag, hipStream_t st )
{
   const int nC = m.mesh->nC();
   const int nb = g.nTiles * nC;
   hipLaunchKernelGGL( ( kLincomb< T > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride,
                       m.cellInfo, flag, (T*) dst, nsrc, nsrc > 0 ? (const T*) src[0] : nullptr,
                       nsrc > 1 ? (const T*) src[1] : nullptr, nsrc > 2 ? (const T*) src[2] : nullptr, T( nsrc > 0 ? s[0] : 0 ),
                       T( nsrc > 1 ? s[1] : 0 ), T( nsrc > 2 ? s[2] : 0 ), accumulate ? 1 : 0, T( value ) );
}

void launchLincomb( const MeshCtx& m, const LevelGeom& g, int L, int dtype, void* dst, int nsrc, const double* s,
                    const void* const* src, bool accumulate, double value, int flag, hipStream_t st )
{
   if ( dtype == 8 )
      lincombT< double >( m, g, L, dst, nsrc, s, src, accumulate, value, flag, st );
   else
      lincombT< float >( m, g, L, dst, nsrc, s, src, accumulate, value, flag, st );
}

void launchMult( const MeshCtx& m, const LevelGeom& g, int L, int dtype, void* dst, const void* a, const void* b, int flag,
                 hipStream_t st )
{
   const int nC = m.mesh->nC();
   const int nb = g.nTiles * nC;
   if ( dtype == 8 )
      hipLaunchKernelGGL( ( kMult< double > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride,
                          m.cellInfo, flag, (double*) dst, (const double*) a, (const double*) b );
   else
      hipLaunchKernelGGL( ( kMult< float > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride,
                          m.cellInfo, flag, (float*) dst, (const float*) a, (const float*) b );
}

template < typename T >
static void cgUpdT( const MeshCtx& m, const LevelGeom& g, int L, void* x, void* r, const void* p, const void* ap,
                    const double* sc, int rsOld, int pAp, double tol2, int flag, hipStream_t st )
{
   const int nC = m.mesh->nC(), nb = g.nTiles * nC;
   hipLaunchKernelGGL( ( kCgUpdate< T > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride,
                       m.cellInfo, flag, (T*) x, (T*) r, (const T*) p, (const T*) ap, sc, rsOld, pAp, tol2 );
}
template < typename T >
static void cgDirT( const MeshCtx& m, const LevelGeom& g, int L, void* p, const void* r, const double* sc, int rsOld, int rsNew,
                    double tol2, int flag, hipStream_t st )
{
   const int nC = m.mesh->nC(), nb = g.nTiles * nC;
   hipLaunchKernelGGL( ( kCgDirection< T > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride,
                       m.cellInfo, flag, (T*) p, (const T*) r, sc, rsOld, rsNew, tol2 );
}
void launchCgUpdate( const MeshCtx& m, const LevelGeom& g, int L, int dtype, void* x, void* r, const void* p, const void* ap,
                     const double* sc, int rsOld, int pAp, double tol2, int flag, hipStream_t st )
{
   if ( dtype == 8 )
      cgUpdT< double >( m, g, L, x, r, p, ap, sc, rsOld, pAp, tol2, flag, st );
   else
      cgUpdT< float >( m, g, L, x, r, p, ap, sc, rsOld, pAp, tol2, flag, st );
}
void launchCgDirection( const MeshCtx& m, const LevelGeom& g, int L, int dtype, void* p, const void* r, const double* sc, int rsOld,
                        int rsNew, double tol2, int flag, hipStream_t st )
{
   if ( dtype == 8 )
      cgDirT< double >( m, g, L, p, r, sc, rsOld, rsNew, tol2, flag, st );
   else
      cgDirT< float >( m, g, L, p, r, sc, rsOld, rsNew, tol2, flag, st );
}

int dotPartialsNeeded( const MeshCtx& m, const LevelGeom& g ) { return g.nTiles * m.mesh->nC(); }

void launchDot( const MeshCtx& m, const LevelGeom& g, int L, int dtype, const void* a, const void* b, int flag, double* partials,
                double* result, hipStream_t st )
{
   const int nC = m.mesh->nC();
   const int nb = g.nTiles * nC;
   if ( dtype == 8 )
      hipLaunchKernelGGL( ( kDot< double > ), dim3( nb ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride, m.cellInfo,
                          flag, (const double*) a, (const double*) b, partials );
   else
      hipLaunchKernelGGL( ( kDot< float > ), dim3( nb ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride, m.cellInfo,
                          flag, (const float*) a, (const float*) b, partials );
   hipLaunchKernelGGL( kDotFinal, dim3( 1 ), dim3( kFinalBlock ), 0, st, partials, nb, result );
}

void launchConvert( i64 n, int dstType, void* dst, int srcType, const void* src, hipStream_t st )
{
   const int grid = int( std::min< i64 >( ( n + kBlock - 1 ) / kBlock, 65536 ) );
   if ( dstType == 8 && srcType == 4 )
      hipLaunchKernelGGL( ( kConvert< double, float > ), dim3( grid ), dim3( kBlock ), 0, st, n, (double*) dst, (const float*) src );
   else if ( dstType == 4 && srcType == 8 )
      hipLaunchKernelGGL( ( kConvert< float, double > ), dim3( grid ), dim3( kBlock ), 0, st, n, (float*) dst, (const double*) src );
   else if ( dstType == 8 )
      hipLaunchKernelGGL( ( kConvert< double, double > ), dim3( grid ), dim3( kBlock ), 0, st, n, (double*) dst, (const double*) src );
   else
      hipLaunchKernelGGL( ( kConvert< float, float > ), dim3( grid ), dim3( kBlock ), 0, st, n, (float*) dst, (const float*) src );
}

void launchCopy( i64 n, int dtype, void* dst, const void* src, hipStream_t st ) { launchConvert( n, dtype, dst, dtype, src, st ); }

void launchInvDiag( const MeshCtx& m, const LevelGeom& g, const hyg_op_s::Level& ol, int L, int dtype, void* dst, hipStream_t st )
{
   const int nC = m.mesh->nC();
   const int nb = g.nTiles * nC;
   if ( dtype == 8 )
      hipLaunchKernelGGL( ( kInvDiag< double > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride,
                          m.cellInfo, ol.cellSt, ol.primInv, (double*) dst );
   else
      hipLaunchKernelGGL( ( kInvDiag< float > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, g.tiles, g.nTiles, nC, L, g.stride,
                          m.cellInfo, ol.cellStF, ol.primInvF, (float*) dst );
}

template < typename T >
static void restrictT( const MeshCtx& m, const LevelGeom& gc, int Lc, void* coarse, const void* fine, int flag, hipStream_t st )
{
   const int nC      = m.mesh->nC();
   const i64 strideF = cellStride( Lc + 1 );
   // cell-interior coarse points are restricted whatever the flag (the reference's cell loop has no flag test,
   // P1toP1LinearRestriction.cpp:181-260)
   if ( gc.nInteriorTiles > 0 )
   {
      const int nb = gc.nInteriorTiles * nC;
      hipLaunchKernelGGL( ( kRestrictCell< T > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, gc.tiles, gc.interiorTile0,
                          gc.nInteriorTiles, nC, Lc, gc.stride, strideF, (T*) coarse, (const T*) fine );
   }
   if ( gc.nChunks > 0 )
      hipLaunchKernelGGL( ( kRestrictLower< T > ), dim3( ceilDiv( gc.nChunks, kBlock / 64 ) ), dim3( kBlock ), 0, st, gc.chunks,
                          gc.nChunks, m.prims, m.entries, m.cellInfo, Lc, gc.stride, strideF, flag, (T*) coarse,
                          (const T*) fine );
}

void launchRestrict( const MeshCtx& m, const LevelGeom& gc, int Lc, int dtype, void* coarse, const void* fine, int flag,
                     hipStream_t st )
{
   if ( dtype == 8 )
      restrictT< double >( m, gc, Lc, coarse, fine, flag, st );
   else
      restrictT< float >( m, gc, Lc, coarse, fine, flag, st );
}

void launchProlongate( const MeshCtx& m, const LevelGeom& gf, int Lf, int dtype, void* fine, const void* coarse, int flag,
                       bool add, hipStream_t st )
{
   const int nC      = m.mesh->nC();
   const int nb      = gf.nTiles * nC;
   const i64 strideC = cellStride( Lf - 1 );
   if ( dtype == 8 )
      hipLaunchKernelGGL( ( kProlongate< double > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, gf.tiles, gf.nTiles, nC, Lf,
                          gf.stride, strideC, m.cellInfo, flag, (double*) fine, (const double*) coarse, add ? 1 : 0 );
   else
      hipLaunchKernelGGL( ( kProlongate< float > ), dim3( xcdGrid( nb ) ), dim3( kBlock ), 0, st, gf.tiles, gf.nTiles, nC, Lf,
                          gf.stride, strideC, m.cellInfo, flag, (float*) fine, (const float*) coarse, add ? 1 : 0 );
}

} // namespace hyg
