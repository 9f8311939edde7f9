// Surrogate stencils of the variable-coefficient P1 operator (the reference's P1SurrogateOperator<
// p1_div_k_grad_affine_q3 >, p1functionspace/P1SurrogateOperator.hpp).
//
// interpolateStencils( degree, interpolationLevel ) (:232-335) replaces, on every level L, the 15 stencil weights of
// the interior points of each macro-cell by polynomials of the given degree in the cell-local coordinates
// (x, y, z) / N, fitted by least squares (polynomial/LSQPInterpolator.hpp:124-183, monomial basis
// polynomial/MonomialBasis3D.hpp) to the exact element-assembled stencils at the interior points of the
// interpolation level Li = min(L, interpolationLevel), scaled to level L (point (x, y, z) of Li is point
// 2^(L - Li) (x, y, z) of L). Face, edge and vertex points keep the exact stencils (:404-469), and so does
// everything at levels without interior points.
//
// Here:
//   * the sample stencils come from the device (kSurSample: the element assembly of the elementwise operator from
//     the stored coefficient integrals), the fit runs on the host (one column-pivoted Householder QR of the
//     sample matrix per level, shared by all cells and directions);
//   * kSurrogate applies the polynomial stencils, one block per interior z-plane of a macro-cell: the polynomials are
//     restricted to the plane, then to pairs of rows, and every lane evaluates its row's polynomials in x by Horner
//     and applies the stencil with the mode's epilogue. The reference evaluates by forward differences along the
//     row (polynomial/PolynomialEvaluator.hpp); both evaluate the same polynomial, in another rounding;
//   * the inverse diagonal of the interior points is 1 / the polynomial centre weight (P1Operator.hpp:645-915 with
//     assemble_stencil_cell).
#include "elementwise.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <vector>

namespace hyg {

namespace {

constexpr int kSurMaxDeg = 12;
constexpr int kSurWaves  = 4;

// number of monomials x^a y^b z^c of degree <= q, and the index of x^a y^b z^c in the order of MonomialBasis3D
// (by total degree; within a degree by decreasing a, then decreasing b)
__host__ __device__ inline int surCoefs( int q ) { return ( q + 1 ) * ( q + 2 ) * ( q + 3 ) / 6; }
__host__ __device__ inline int surMonomial( int a, int b, int c )
{
   const int d = a + b + c;
   return d * ( d + 1 ) * ( d + 2 ) / 6 + ( d - a ) * ( d - a + 1 ) / 2 + ( d - a - b );
}

// exact stencils at sample points (3 lattice coordinates each) of every cell: out[(cell, sample)][15]
__global__ __launch_bounds__( kBlock ) void kSurSample( int L, int nPts, int tiles, const int* __restrict__ pts,
                                                        const ElemCell* __restrict__ cells, const double* __restrict__ kb,
                                                        double* __restrict__ out )
{
   const int c = int( blockIdx.x ) / tiles;
   const int s = ( int( blockIdx.x ) - c * tiles ) * kBlock + int( threadIdx.x );
   if ( s >= nPts )
      return;
   const int N = 1 << L;
   double    W[NDIR];
   elemWeights< double, false >( cells[c], kb + i64( c ) * 6 * S3i( N ), N, pts[3 * s], pts[3 * s + 1], pts[3 * s + 2], W );
   double* o = out + ( i64( c ) * nPts + s ) * NDIR;
#pragma unroll
   for ( int d = 0; d < NDIR; ++d )
      o[d] = W[d];
}

// Cell-interior points of the surrogate operator. Block = one interior z-plane (z = 1 .. N-3) of a macro-cell:
//   1. the block restricts the 15 direction polynomials to the plane, pz[d][(a, b)] = sum_c P_d[a, b, c] z^c (LDS);
//   2. the plane's interior rows y = 1 .. M (M = N-2-z, row y has M+1-y points) are taken in folded pairs
//      (y, M+1-y), whose lengths add up to M+1, so that a wave's lanes cover both rows of a pair without gaps; per
//      pair the wave restricts the plane polynomials to its two rows, c1[row][d][a] = sum_b pz[d][(a, b)] y^b (LDS),
//      and every lane evaluates its row's polynomials at its point by Horner and applies the stencil.
// DIAG: writes 1 / the centre weight into diagOut instead. The LDS of each wave is written and read by that wave
// only (its LDS instructions execute in order), so no barrier is needed after step 1.
// DEG >= 0: the degree at compile time (all polynomial loops unrolled, so the LDS reads of the 15 Horner chains
// are issued together); DEG = -1: any degree up to kSurMaxDeg, given by degRun.
template < typename T, int MODE, bool DIAG, int DEG >
__global__ __launch_bounds__( 64 * kSurWaves ) void kSurrogate( int L, int degRun, int planes, const double* __restrict__ poly, i64 stride,
                                                               int flag, const T* __restrict__ dinv, const T* __restrict__ x,
                                                               const T* __restrict__ in1, T* __restrict__ out0, T* __restrict__ out1, T a,
                                                               double* __restrict__ diagOut )
{
   const int deg = DEG >= 0 ? DEG : degRun;
   __shared__ double pz[NDIR * ( kSurMaxDeg + 1 ) * ( kSurMaxDeg + 2 ) / 2];
   __shared__ double c1[kSurWaves][2][NDIR * ( kSurMaxDeg + 1 )];
   const int    c  = int( blockIdx.x ) / planes;
   const int    z  = 1 + int( blockIdx.x ) - c * planes;
   const int    wv = int( threadIdx.x ) >> 6, lane = int( threadIdx.x ) & 63;
   const int    N = 1 << L, w = N + 1, nc = surCoefs( deg ), q1 = deg + 1, n2 = q1 * ( q1 + 1 ) / 2;
   const double h  = 1.0 / double( N );
   {
      const double  Z = z * h;
      const double* P = poly + i64( c ) * NDIR * nc;
      for ( int it = int( threadIdx.x ); it < NDIR * n2; it += 64 * kSurWaves )
      {
         const int d = it / n2;
         int       j = it - d * n2, aa = 0; // j = index of (aa, b) with aa + b <= deg, aa-major
         while ( j > deg - aa )
         {
            j -= deg - aa + 1;
            ++aa;
         }
         const int b = j;
         double    t = 0.0;
#pragma unroll
         for ( int cc = kSurMaxDeg; cc >= 0; --cc )
            if ( cc <= deg - aa - b )
               t = t * Z + P[d * nc + surMonomial( aa, b, cc )];
         pz[it] = t;
      }
   }
   __syncthreads();
   if ( !( flag & 1 ) )
      return;
   const int M = N - 2 - z; // interior rows y = 1 .. M of the plane, row y has M + 1 - y points
   const int nPairs = ( M + 1 ) / 2;
   const int nIt    = NDIR * q1;
   for ( int k = wv; k < nPairs; k += kSurWaves )
   {
      const int y1 = 1 + k, y2 = M - k;
      const int len1 = M - k, total = y1 == y2 ? len1 : M + 1;
      // restriction of the plane polynomials to rows y1 and y2 (Horner in y)
      for ( int it = lane; it < 2 * nIt; it += 64 )
      {
         const int    row = it >= nIt ? 1 : 0, rem = it - row * nIt;
         const int    d = rem / q1, aa = rem - d * q1;
         const double Y = ( row ? y2 : y1 ) * h;
         const double* pd = pz + d * n2 + aa * ( deg + 1 ) - aa * ( aa - 1 ) / 2;
         double        t  = 0.0;
#pragma unroll
         for ( int b = ( DEG >= 0 ? DEG : kSurMaxDeg ); b >= 0; --b )
            if ( b <= deg - aa )
               t = t * Y + pd[b];
         c1[wv][row][rem] = t;
      }
      __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
      for ( int p0 = 0; p0 < total; p0 += 64 )
      {
         const int p = p0 + lane;
         if ( p >= total )
            break;
         const int    row = p < len1 ? 0 : 1;
         const int    px = 1 + ( row ? p - len1 : p ), py = row ? y2 : y1;
         const double X  = px * h;
         const double* cr = c1[wv][row];
         double        Wd[NDIR];
#pragma unroll
         for ( int d = 0; d < NDIR; ++d )
         {
            double v = 0.0;
            if constexpr ( DEG >= 0 )
            {
#pragma unroll
               for ( int aa = DEG; aa >= 0; --aa )
                  v = v * X + cr[d * q1 + aa];
            }
            else
               for ( int aa = deg; aa >= 0; --aa )
                  v = v * X + cr[d * q1 + aa];
            Wd[d] = v;
         }
         const i64 q = i64( c ) * stride + cidx( w, px, py, z );
         if constexpr ( DIAG )
            diagOut[q] = 1.0 / Wd[0];
         else
         {
            const int Wz = w - z, tc = T2i( Wz ) - py, bc = T2i( Wz + 1 ) - py, nn = Wz - py;
            const int off[NDIR] = { 0, -1, 1, nn, -( nn + 1 ), nn - 1, -nn, tc, tc - 1, tc - nn, tc - nn + 1, -bc, -bc + nn + 1, -bc + 1, -bc + nn };
            T         val = T( 0 );
#pragma unroll
            for ( int d = 0; d < NDIR; ++d )
               val += T( Wd[d] ) * x[q + off[d]];
            T i1 = T( 0 ), r0, r1;
            if constexpr ( MODE == MODE_RESIDUAL || MODE == MODE_JACOBI || MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
               i1 = in1[q];
            if constexpr ( MODE == MODE_APPLY_ADD )
               i1 = out0[q];
            epilogue< T, MODE >( val, x[q], i1, i1, dinv[q], a, r0, r1 );
            out0[q] = r0;
            if constexpr ( MODE == MODE_CHEB0 || MODE == MODE_CHEBK )
               out1[q] = r1;
         }
      }
      __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
   }
}

// Least squares by Householder QR with column pivoting (the fit of LSQPInterpolator3D::interpolate, which calls
// Eigen's colPivHouseholderQr().solve): a is m x n row-major, b holds nr right-hand sides of length m one after the
// other and receives the nr solutions of length n. The pivot is the remaining column of largest norm (the first of
// equals); columns beyond the numerical rank (|R_kk| <= eps min(m, n) |R_00|) get zero.
void lsqColumnPivoted( std::vector< double > a, int m, int n, std::vector< double >& b, int nr )
{
   std::vector< int >    piv( n );
   std::vector< double > tau( n, 0.0 );
   for ( int j = 0; j < n; ++j )
      piv[j] = j;
   const int kmax = std::min( m, n );
   int       rank = 0;
   double    r00  = 0.0;
   auto      A    = [&]( int i, int j ) -> double& { return a[size_t( i ) * n + j]; };
   for ( int k = 0; k < kmax; ++k )
   {
      int    jb = k;
      double nb = -1.0;
      for ( int j = k; j < n; ++j )
      {
         double s = 0.0;
         for ( int i = k; i < m; ++i )
            s += A( i, j ) * A( i, j );
         if ( s > nb )
         {
            nb = s;
            jb = j;
         }
      }
      if ( jb != k )
      {
         for ( int i = 0; i < m; ++i )
            std::swap( A( i, k ), A( i, jb ) );
         std::swap( piv[k], piv[jb] );
      }
      const double nx = std::sqrt( nb );
      if ( k == 0 )
         r00 = nx;
      if ( nx == 0.0 || nx <= std::numeric_limits< double >::epsilon() * kmax * r00 )
         break;
      ++rank;
      const double alpha = A( k, k ), beta = alpha >= 0.0 ? -nx : nx;
      tau[k]             = ( beta - alpha ) / beta;
      const double sc    = 1.0 / ( alpha - beta );
      for ( int i = k + 1; i < m; ++i )
         A( i, k ) *= sc;
      A( k, k ) = beta;
      for ( int j = k + 1; j < n; ++j )
      {
         double v = A( k, j );
         for ( int i = k + 1; i < m; ++i )
            v += A( i, k ) * A( i, j );
         v *= tau[k];
         A( k, j ) -= v;
         for ( int i = k + 1; i < m; ++i )
            A( i, j ) -= v * A( i, k );
      }
   }
   std::vector< double > out( size_t( n ) * nr, 0.0 ), z( n );
   for ( int r = 0; r < nr; ++r )
   {
      double* br = b.data() + size_t( r ) * m;
      for ( int k = 0; k < rank; ++k )
      {
         double v = br[k];
         for ( int i = k + 1; i < m; ++i )
            v += A( i, k ) * br[i];
         v *= tau[k];
         br[k] -= v;
         for ( int i = k + 1; i < m; ++i )
            br[i] -= v * A( i, k );
      }
      std::fill( z.begin(), z.end(), 0.0 );
      for ( int k = rank - 1; k >= 0; --k )
      {
         double v = br[k];
         for ( int j = k + 1; j < rank; ++j )
            v -= A( k, j ) * z[j];
         z[k] = v / A( k, k );
      }
      for ( int j = 0; j < n; ++j )
         out[size_t( r ) * n + piv[j]] = z[j];
   }
   b.swap( out );
}

template < typename T >
T* toDeviceArray( const std::vector< T >& h )
{
   void* d = nullptr;
   if ( hipMalloc( &d, std::max< size_t >( h.size(), 1 ) * sizeof( T ) ) != hipSuccess )
      throw std::runtime_error( "hipMalloc failed for surrogate tables" );
   if ( !h.empty() && hipMemcpy( d, h.data(), h.size() * sizeof( T ), hipMemcpyHostToDevice ) != hipSuccess )
      throw std::runtime_error( "hipMemcpy failed for surrogate tables" );
   return static_cast< T* >( d );
}

template < typename T, int MODE, bool DIAG, int DEG >
void surLaunchDeg( const MeshCtx& m, const LevelGeom& g, const hyg_op_s& op, int L, const StencilCall& c, double* diagOut, hipStream_t s )
{
   const hyg_op_s::Level& ol     = op.levels[L];
   const int              planes = ( 1 << L ) - 3;
   constexpr bool         dbl    = std::is_same< T, double >::value;
   const T*               dv     = dbl ? (const T*) ol.dinv : (const T*) ol.dinvF;
   hipLaunchKernelGGL( ( kSurrogate< T, MODE, DIAG, DEG > ), dim3( m.mesh->nC() * planes ), dim3( 64 * kSurWaves ), 0, s, L,
                       op.surDegree, planes, ol.surPoly, g.stride, c.flag, dv, (const T*) c.x, (const T*) c.in1, (T*) c.out0, (T*) c.out1,
                       T( c.a ), diagOut );
}

template < typename T, int MODE, bool DIAG >
void surLaunch( const MeshCtx& m, const LevelGeom& g, const hyg_op_s& op, int L, const StencilCall& c, double* diagOut, hipStream_t s )
{
   switch ( op.surDegree )
   {
   case 0: surLaunchDeg< T, MODE, DIAG, 0 >( m, g, op, L, c, diagOut, s ); break;
   case 1: surLaunchDeg< T, MODE, DIAG, 1 >( m, g, op, L, c, diagOut, s ); break;
   case 2: surLaunchDeg< T, MODE, DIAG, 2 >( m, g, op, L, c, diagOut, s ); break;
   case 3: surLaunchDeg< T, MODE, DIAG, 3 >( m, g, op, L, c, diagOut, s ); break;
   case 4: surLaunchDeg< T, MODE, DIAG, 4 >( m, g, op, L, c, diagOut, s ); break;
   case 5: surLaunchDeg< T, MODE, DIAG, 5 >( m, g, op, L, c, diagOut, s ); break;
   case 6: surLaunchDeg< T, MODE, DIAG, 6 >( m, g, op, L, c, diagOut, s ); break;
   default: surLaunchDeg< T, MODE, DIAG, -1 >( m, g, op, L, c, diagOut, s ); break;
   }
}

template < typename T >
void surMode( const MeshCtx& m, const LevelGeom& g, const hyg_op_s& op, int L, const StencilCall& c, hipStream_t s )
{
   switch ( c.mode )
   {
   case MODE_APPLY: surLaunch< T, MODE_APPLY, false >( m, g, op, L, c, nullptr, s ); break;
   case MODE_APPLY_ADD: surLaunch< T, MODE_APPLY_ADD, false >( m, g, op, L, c, nullptr, s ); break;
   case MODE_RESIDUAL: surLaunch< T, MODE_RESIDUAL, false >( m, g, op, L, c, nullptr, s ); break;
   case MODE_JACOBI: surLaunch< T, MODE_JACOBI, false >( m, g, op, L, c, nullptr, s ); break;
   case MODE_CHEB0: surLaunch< T, MODE_CHEB0, false >( m, g, op, L, c, nullptr, s ); break;
   default: surLaunch< T, MODE_CHEBK, false >( m, g, op, L, c, nullptr, s ); break;
   }
}

} // namespace

bool surrogateActive( const hyg_op_s& op, int L )
{
   return op.kind == 1 && op.surDegree >= 0 && L < (int) op.levels.size() && op.levels[L].surPoly != nullptr;
}

void launchSurrogateInterior( const MeshCtx& m, const LevelGeom& g, const hyg_op_s& op, int L, int dtype, const StencilCall& c, hipStream_t s )
{
   if ( dtype == 8 )
      surMode< double >( m, g, op, L, c, s );
   else
      surMode< float >( m, g, op, L, c, s );
}

void launchSurrogateInverseDiagonal( const MeshCtx& m, const LevelGeom& g, const hyg_op_s& op, int L, hipStream_t s )
{
   StencilCall c;
   c.flag = 15;
   surLaunch< double, MODE_APPLY, true >( m, g, op, L, c, op.levels[L].dinv, s );
}

void surrogateInterpolate( const MeshCtx& mc, hyg_op_s& op, int deg, int interpolationLevel )
{
   const Mesh& m  = *mc.mesh;
   const int   nC = m.nC();
   const int   nc = surCoefs( deg );
   for ( int L = op.minL; L <= op.maxL; ++L )
   {
      hyg_op_s::Level& ol = op.levels[L];
      (void) hipFree( ol.surPoly );
      (void) hipFree( ol.surBnd );
      ol.surPoly = nullptr;
      ol.surBnd  = nullptr;
      ol.surNBnd = 0;
      const int N = 1 << L, w = N + 1;
      if ( N < 4 )
         continue; // no interior points
      std::vector< int > bnd;
      for ( int z = 0; z <= N; ++z )
         for ( int y = 0; y + z <= N; ++y )
            for ( int x = 0; x + y + z <= N; ++x )
               if ( x == 0 || y == 0 || z == 0 || x + y + z == N )
                  bnd.push_back( cidx( w, x, y, z ) );
      // sample points: the interior points of the interpolation level, scaled to level L
      const int Li = std::min( L, interpolationLevel ), Ni = 1 << Li, sh = L - Li;
      std::vector< int >    pts;
      std::vector< double > vand;
      for ( int z = 1; z <= Ni - 3; ++z )
         for ( int y = 1; y <= Ni - 2 - z; ++y )
            for ( int x = 1; x <= Ni - 1 - y - z; ++x )
            {
               pts.insert( pts.end(), { x << sh, y << sh, z << sh } );
               const double X[3] = { double( x ) / Ni, double( y ) / Ni, double( z ) / Ni };
               for ( int i = 0; i < nc; ++i )
                  vand.push_back( 0.0 );
               for ( int d = 0; d <= deg; ++d )
                  for ( int ea = d; ea >= 0; --ea )
                     for ( int eb = d - ea; eb >= 0; --eb )
                        vand[vand.size() - nc + surMonomial( ea, eb, d - ea - eb )] =
                            std::pow( X[0], ea ) * std::pow( X[1], eb ) * std::pow( X[2], d - ea - eb );
            }
      const int np = int( pts.size() / 3 );
      if ( np == 0 )
         throw std::invalid_argument( "the interpolation level has no interior points" );
      // exact stencils at the sample points, from the device
      std::vector< double > st( size_t( nC ) * np * NDIR );
      int*                  dPts = toDeviceArray( pts );
      double*               dOut = nullptr;
      if ( hipMalloc( &dOut, st.size() * sizeof( double ) ) != hipSuccess )
      {
         (void) hipFree( dPts );
         throw std::runtime_error( "hipMalloc failed for surrogate samples" );
      }
      const int tiles = ceilDiv( np, kBlock );
      hipLaunchKernelGGL( kSurSample, dim3( nC * tiles ), dim3( kBlock ), 0, nullptr, L, np, tiles, dPts, ol.elem, ol.kb, dOut );
      const hipError_t e = hipMemcpy( st.data(), dOut, st.size() * sizeof( double ), hipMemcpyDeviceToHost );
      (void) hipFree( dPts );
      (void) hipFree( dOut );
      if ( e != hipSuccess )
         throw std::runtime_error( "surrogate sampling failed" );
      // one fit per cell and direction, all with the same sample matrix
      std::vector< double > rhs( size_t( nC ) * NDIR * np );
      for ( int c = 0; c < nC; ++c )
         for ( int d = 0; d < NDIR; ++d )
            for ( int p = 0; p < np; ++p )
               rhs[( size_t( c ) * NDIR + d ) * np + p] = st[( size_t( c ) * np + p ) * NDIR + d];
      lsqColumnPivoted( vand, np, nc, rhs, nC * NDIR );
      ol.surPoly = toDeviceArray( rhs ); // [cell][direction][monomial]
      ol.surBnd  = toDeviceArray( bnd );
      ol.surNBnd = int( bnd.size() );
   }
   op.surDegree = deg;
}

} // namespace hyg
