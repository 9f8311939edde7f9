// Device helpers shared by the gfx950 kernels.
#pragma once

#include "kernels.hpp"

#include <algorithm>
#include <cstdlib>
#include <type_traits>

namespace hyg {

constexpr int kBlock = 256;

inline int ceilDiv( i64 a, i64 b ) { return int( ( a + b - 1 ) / b ); }
inline int xcdGrid( int nb ) { return ( ( nb + 7 ) / 8 ) * 8; }

// hardware dispatches consecutive workgroups round-robin over the 8 XCDs
__device__ __forceinline__ int remapXcd( int b, int nb )
{
   const int per = ( nb + 7 ) >> 3;
   return ( b & 7 ) * per + ( b >> 3 );
}

// (x, y) of the s-th point of a plane whose row 0 holds W points (row y holds W - y points)
__device__ __forceinline__ void planeXY( int W, int s, int& x, int& y )
{
   const float b  = 2.0f * float( W ) + 1.0f;
   int         yy = int( ( b - sqrtf( b * b - 8.0f * float( s ) ) ) * 0.5f );
   yy             = yy < 0 ? 0 : yy;
   if ( rowStart( W, yy + 1 ) <= s )
      ++yy;
   if ( rowStart( W, yy ) > s )
      --yy;
   y = yy;
   x = s - rowStart( W, yy );
}


// cell coordinates of point (i, j) of a lower-dimensional primitive in an incident cell
__device__ __forceinline__ void entryPos( const LowerEntry& e, int N, int i, int j, int& qx, int& qy, int& qz )
{
   const int u0x = e.l0 == 1, u0y = e.l0 == 2, u0z = e.l0 == 3;
   const int u1x = e.l1 == 1, u1y = e.l1 == 2, u1z = e.l1 == 3;
   const int u2x = e.l2 == 1, u2y = e.l2 == 2, u2z = e.l2 == 3;
   qx = N * u0x + i * ( u1x - u0x ) + j * ( u2x - u0x );
   qy = N * u0y + i * ( u1y - u0y ) + j * ( u2y - u0y );
   qz = N * u0z + i * ( u1z - u0z ) + j * ( u2z - u0z );
}

__device__ __forceinline__ int slotType( const CellInfo& ci, int slot ) { return int( ( ci.bcword >> ( 4 * slot ) ) & 15u ); }

// streaming store: results are not re-read by this kernel, keep them from displacing inputs in L2
template < typename T >
__device__ __forceinline__ void storeStream( T* p, T v )
{
   __builtin_nontemporal_store( v, p );
}

// ------------------------------------------------------------------ single-lane moves
// returns `old` with lane `dst` replaced by lane `src` of v (src, dst wave-uniform)
__device__ __forceinline__ double moveLane( double v, int src, int dst, double old )
{
   const int    lo = __builtin_amdgcn_readlane( __double2loint( v ), src );
   const int    hi = __builtin_amdgcn_readlane( __double2hiint( v ), src );
   const double s  = __hiloint2double( hi, lo );
   return int( threadIdx.x & 63u ) == dst ? s : old;
}
__device__ __forceinline__ float moveLane( float v, int src, int dst, float old )
{
   const float s = __int_as_float( __builtin_amdgcn_readlane( __float_as_int( v ), src ) );
   return int( threadIdx.x & 63u ) == dst ? s : old;
}

// ------------------------------------------------------------------ raw buffer access
// A buffer descriptor over `bytes` bytes from p (gfx9 family: word 3 = 0x00020000, untyped access).
// Accesses give `voff` (per lane) and `soff` (wave-uniform) byte offsets; a store whose voff is
// kDropOffset lies outside every descriptor's range and is dropped by the hardware, so masked
// lanes need no branch around the store.
constexpr unsigned kDropOffset = 0x80000000u;
typedef unsigned int u32x2 __attribute__( ( ext_vector_type( 2 ) ) );

__device__ __forceinline__ __amdgpu_buffer_rsrc_t makeRsrc( const void* p, unsigned bytes )
{
   return __builtin_amdgcn_make_buffer_rsrc( const_cast< void* >( p ), 0, int( bytes ), 0x00020000 );
}
__device__ __forceinline__ double bufLoad( __amdgpu_buffer_rsrc_t r, unsigned voff, int soff, double* )
{
   return __builtin_bit_cast( double, __builtin_amdgcn_raw_buffer_load_b64( r, voff, soff, 0 ) );
}
__device__ __forceinline__ float bufLoad( __amdgpu_buffer_rsrc_t r, unsigned voff, int soff, float* )
{
   return __builtin_bit_cast( float, __builtin_amdgcn_raw_buffer_load_b32( r, voff, soff, 0 ) );
}
template < typename T >
__device__ __forceinline__ T bufLoad( __amdgpu_buffer_rsrc_t r, unsigned voff, int soff )
{
   return bufLoad( r, voff, soff, static_cast< T* >( nullptr ) );
}
// two consecutive values per lane
typedef unsigned int u32x4 __attribute__( ( ext_vector_type( 4 ) ) );
__device__ __forceinline__ void bufLoad2( __amdgpu_buffer_rsrc_t r, unsigned voff, int soff, double& a, double& b )
{
   const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128( r, voff, soff, 0 );
   a             = __hiloint2double( int( v.y ), int( v.x ) );
   b             = __hiloint2double( int( v.w ), int( v.z ) );
}
__device__ __forceinline__ void bufLoad2( __amdgpu_buffer_rsrc_t r, unsigned voff, int soff, float& a, float& b )
{
   const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64( r, voff, soff, 0 );
   a             = __uint_as_float( v.x );
   b             = __uint_as_float( v.y );
}
// streaming (non-temporal) stores
__device__ __forceinline__ void bufStore( double v, __amdgpu_buffer_rsrc_t r, unsigned voff, int soff )
{
   __builtin_amdgcn_raw_buffer_store_b64( __builtin_bit_cast( u32x2, v ), r, voff, soff, 2 );
}
__device__ __forceinline__ void bufStore( float v, __amdgpu_buffer_rsrc_t r, unsigned voff, int soff )
{
   __builtin_amdgcn_raw_buffer_store_b32( __builtin_bit_cast( unsigned, v ), r, voff, soff, 2 );
}

// ------------------------------------------------------------------ stencil epilogue
template < typename T, int MODE >
__device__ __forceinline__ void epilogue( T Ax, T xc, T in1v, T out0v, T dinv, T a, T& o0, T& o1 )
{
   if constexpr ( MODE == MODE_APPLY )
      o0 = Ax;
   else if constexpr ( MODE == MODE_APPLY_ADD )
      o0 = out0v + Ax;
   else if constexpr ( MODE == MODE_RESIDUAL )
      o0 = in1v - Ax;
   else if constexpr ( MODE == MODE_JACOBI || MODE == MODE_SOR )
      o0 = xc + a * dinv * ( in1v - Ax );
   else if constexpr ( MODE == MODE_CHEB0 )
   {
      const T t = dinv * ( in1v - Ax );
      o0        = t;
      o1        = xc + a * t;
   }
   else
   {
      const T t = dinv * Ax;
      o0        = t;
      o1        = in1v + a * t;
   }
}

// ------------------------------------------------------------------ cross-lane shifts
// wave_shr:1 -> lane i receives lane i-1, wave_shl:1 -> lane i receives lane i+1 (whole 64-lane wave)
constexpr int kDppWaveShr1 = 0x138;
constexpr int kDppWaveShl1 = 0x130;

template < int CTRL >
__device__ __forceinline__ double laneShift( double v )
{
   const int lo = __builtin_amdgcn_update_dpp( 0, __double2loint( v ), CTRL, 0xf, 0xf, false );
   const int hi = __builtin_amdgcn_update_dpp( 0, __double2hiint( v ), CTRL, 0xf, 0xf, false );
   return __hiloint2double( hi, lo );
}
template < int CTRL >
__device__ __forceinline__ float laneShift( float v )
{
   return __int_as_float( __builtin_amdgcn_update_dpp( 0, __float_as_int( v ), CTRL, 0xf, 0xf, false ) );
}
// the same shifts where the lane without a source (lane 0 for shr, lane 63 for shl) may receive 0
template < int CTRL >
__device__ __forceinline__ double laneMove( double v )
{
   const int lo = __builtin_amdgcn_mov_dpp( __double2loint( v ), CTRL, 0xf, 0xf, true );
   const int hi = __builtin_amdgcn_mov_dpp( __double2hiint( v ), CTRL, 0xf, 0xf, true );
   return __hiloint2double( hi, lo );
}
template < int CTRL >
__device__ __forceinline__ float laneMove( float v )
{
   return __int_as_float( __builtin_amdgcn_mov_dpp( __float_as_int( v ), CTRL, 0xf, 0xf, true ) );
}


} // namespace hyg
