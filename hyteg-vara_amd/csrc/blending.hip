// Spherical-shell blending: the reference's IcosahedralShellMap (geometry/IcosahedralShellMap.hpp) and the blended
// P1 Laplacian P1ElementwiseBlendingLaplaceOperator = P1ElementwiseOperator< p1_diffusion_blending_q3 >
// (elementwiseoperators/P1ElementwiseOperator.hpp:332, forms/form_hyteg_generated/p1/p1_diffusion_blending_q3.cpp).
//
// Classification of a shell cell (classifyTet :281-317, classifyVertices :346-470): two vertices lie on one ray from
// the origin; when one vertex lies on one sphere and three on the other, the lone vertex is `ref`, its ray partner
// `ray`, the remaining two `thr` (the later in local order) and `fo`; a cell with two vertices on each sphere takes
// the classification of the cell across its first local face that does not contain both ray vertices
// (findNonSkewTetInPrism :473-515).
//
// Element matrix of a micro-cell with affine vertices X (the reference integrates over the affine micro-cell with
// the map's Jacobian at the quadrature points): a_ij = |det J| sum_q w_q |det DF_q| (DF_q^-T g_i) . (DF_q^-T g_j),
// g_i the gradients of the affine barycentric coordinates, the 5-point degree-3 rule. Here the 15 one-sided stencil
// weights of every point of every macro-cell (the rows of its incident micro-cells' matrices) are summed once, on the
// device, at operator creation (kBlendWeights); applies, smoothers and the inverse diagonal then run the elementwise
// operator's kernels with those weights (kElemApply< ..., STW >), so an apply streams 15 weights per point.
#include "elementwise.hpp"
#include "kernels.hpp"

#include <cmath>
#include <stdexcept>

namespace hyg {

namespace {

// classification of cell c (depth guards the skew-cell hop)
ShellMap classifyShellCell( const Mesh& m, int c, int depth )
{
   if ( depth > 3 )
      throw std::runtime_error( "no classifiable shell cell near a skew cell" );
   const auto& cv = m.Cv[c];
   double      v[4][3], rad[4], lo = 1e300, hi = 0.0;
   for ( int k = 0; k < 4; ++k )
   {
      for ( int r = 0; r < 3; ++r )
         v[k][r] = m.X[cv[k]][r];
      rad[k] = std::sqrt( v[k][0] * v[k][0] + v[k][1] * v[k][1] + v[k][2] * v[k][2] );
      lo     = std::min( lo, rad[k] );
      hi     = std::max( hi, rad[k] );
   }
   int outer = 0;
   for ( int k = 0; k < 4; ++k )
      outer += hi - rad[k] < 1e-3 * ( hi - lo ) ? 1 : 0;
   int a = -1, b = -1;
   for ( int k = 0; k < 4 && a < 0; ++k )
      for ( int j = k + 1; j < 4 && a < 0; ++j )
      {
         const double s     = rad[k] * rad[j];
         const double cr[3] = { v[k][1] * v[j][2] - v[k][2] * v[j][1], v[k][2] * v[j][0] - v[k][0] * v[j][2],
                                v[k][0] * v[j][1] - v[k][1] * v[j][0] };
         if ( std::fabs( cr[0] ) <= 1e-14 * s && std::fabs( cr[1] ) <= 1e-14 * s && std::fabs( cr[2] ) <= 1e-14 * s )
            a = k, b = j;
      }
   if ( a < 0 || outer < 1 || outer > 3 )
      throw std::runtime_error( "cell is not a spherical-shell cell (no two vertices on one ray, or not between two spheres)" );
   if ( outer == 2 )
   {
      static const int faces[4][3] = { { 0, 1, 2 }, { 0, 1, 3 }, { 0, 2, 3 }, { 1, 2, 3 } };
      for ( int f = 0; f < 4; ++f )
      {
         const bool hasA = faces[f][0] == a || faces[f][1] == a || faces[f][2] == a;
         const bool hasB = faces[f][0] == b || faces[f][1] == b || faces[f][2] == b;
         if ( hasA && hasB )
            continue;
         for ( int o : m.fC[m.cF[c][f]] )
            if ( o != c )
               return classifyShellCell( m, o, depth + 1 );
      }
      throw std::runtime_error( "no neighbour of a skew shell cell" );
   }
   int ref = a, ray = b;
   if ( ( outer == 1 && rad[ref] < rad[ray] ) || ( outer == 3 && rad[ref] > rad[ray] ) )
      std::swap( ref, ray );
   int thr = -1, fo = -1;
   for ( int k = 0; k < 4; ++k )
      if ( k != ref && k != ray )
         thr = k; // the later of the two
   for ( int k = 0; k < 4; ++k )
      if ( k != ref && k != ray && k != thr )
         fo = k;
   ShellMap sm{};
   for ( int r = 0; r < 3; ++r )
   {
      sm.ref[r] = v[ref][r];
      sm.ray[r] = v[ray][r];
      sm.thr[r] = v[thr][r];
      sm.fo[r]  = v[fo][r];
   }
   sm.rRef = rad[ref];
   sm.rRay = rad[ray];
   const double p[3] = { sm.thr[0] - sm.ray[0], sm.thr[1] - sm.ray[1], sm.thr[2] - sm.ray[2] };
   const double q[3] = { sm.fo[0] - sm.ray[0], sm.fo[1] - sm.ray[1], sm.fo[2] - sm.ray[2] };
   sm.n[0]           = p[1] * q[2] - p[2] * q[1];
   sm.n[1]           = p[2] * q[0] - p[0] * q[2];
   sm.n[2]           = p[0] * q[1] - p[1] * q[0];
   sm.vt = ( sm.ref[0] - sm.ray[0] ) * sm.n[0] + ( sm.ref[1] - sm.ray[1] ) * sm.n[1] + ( sm.ref[2] - sm.ray[2] ) * sm.n[2];
   return sm;
}

__device__ __forceinline__ double inv3( const double D[3][3], double Di[3][3] )
{
   const double d = D[0][0] * ( D[1][1] * D[2][2] - D[1][2] * D[2][1] ) - D[0][1] * ( D[1][0] * D[2][2] - D[1][2] * D[2][0] ) +
                    D[0][2] * ( D[1][0] * D[2][1] - D[1][1] * D[2][0] );
   const double r = 1.0 / d;
   Di[0][0]       = ( D[1][1] * D[2][2] - D[1][2] * D[2][1] ) * r;
   Di[0][1]       = ( D[0][2] * D[2][1] - D[0][1] * D[2][2] ) * r;
   Di[0][2]       = ( D[0][1] * D[1][2] - D[0][2] * D[1][1] ) * r;
   Di[1][0]       = ( D[1][2] * D[2][0] - D[1][0] * D[2][2] ) * r;
   Di[1][1]       = ( D[0][0] * D[2][2] - D[0][2] * D[2][0] ) * r;
   Di[1][2]       = ( D[0][2] * D[1][0] - D[0][0] * D[1][2] ) * r;
   Di[2][0]       = ( D[1][0] * D[2][1] - D[1][1] * D[2][0] ) * r;
   Di[2][1]       = ( D[0][1] * D[2][0] - D[0][0] * D[2][1] ) * r;
   Di[2][2]       = ( D[0][0] * D[1][1] - D[0][1] * D[1][0] ) * r;
   return d;
}

// One thread per point of a macro-cell: the rows of the blended element matrices of its incident micro-cells
// (those inside the cell), summed per stencil direction into W[cell][direction][point].
__global__ __launch_bounds__( kBlock ) void kBlendWeights( int L, int tilesPerCell, const ElemCell* __restrict__ cells,
                                                           const ShellMap* __restrict__ maps, double* __restrict__ W )
{
   const int c = int( blockIdx.x ) / tilesPerCell;
   const int s = ( int( blockIdx.x ) - c * tilesPerCell ) * kBlock + int( threadIdx.x );
   const int N = 1 << L, w = N + 1, S = S3i( w );
   if ( s >= S )
      return;
   int px, py, pz;
   tetXYZ( w, s, px, py, pz );
   const ElemCell& ec = cells[c];
   const ShellMap  sm = maps[c];
   double          acc[NDIR];
   for ( int d = 0; d < NDIR; ++d )
      acc[d] = 0.0;
   constexpr double lamQ[5][4] = { { 0.25, 0.25, 0.25, 0.25 },
                                   { 1.0 / 6, 1.0 / 6, 1.0 / 6, 0.5 },
                                   { 1.0 / 6, 1.0 / 6, 0.5, 1.0 / 6 },
                                   { 1.0 / 6, 0.5, 1.0 / 6, 1.0 / 6 },
                                   { 0.5, 1.0 / 6, 1.0 / 6, 1.0 / 6 } };
   constexpr double wQ[5]      = { -2.0 / 15.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0 };
   for ( int t = 0; t < 6; ++t )
      for ( int k = 0; k < 4; ++k )
      {
         // the micro-cell of type t whose vertex k is this point
         const int ax = px - kMC[t][k][0], ay = py - kMC[t][k][1], az = pz - kMC[t][k][2];
         if ( ax < 0 || ay < 0 || az < 0 || ax + ay + az + kMCMax[t] > N )
            continue;
         double X[4][3];
         for ( int u = 0; u < 4; ++u )
         {
            const double l[3] = { double( ax + kMC[t][u][0] ), double( ay + kMC[t][u][1] ), double( az + kMC[t][u][2] ) };
            for ( int r = 0; r < 3; ++r )
               X[u][r] = ec.v0[r] + ec.a[r][0] * l[0] + ec.a[r][1] * l[1] + ec.a[r][2] * l[2];
         }
         double J[3][3], Ji[3][3];
         for ( int r = 0; r < 3; ++r )
            for ( int q = 0; q < 3; ++q )
               J[r][q] = X[q + 1][r] - X[0][r];
         const double detJ = fabs( inv3( J, Ji ) );
         // gradients of the barycentric coordinates: rows of J^-1, and minus their sum
         double G[4][3];
         for ( int q = 0; q < 3; ++q )
         {
            G[1][q] = Ji[0][q];
            G[2][q] = Ji[1][q];
            G[3][q] = Ji[2][q];
            G[0][q] = -( Ji[0][q] + Ji[1][q] + Ji[2][q] );
         }
         double row[4] = { 0.0, 0.0, 0.0, 0.0 };
         for ( int q = 0; q < 5; ++q )
         {
            double xq[3] = { 0.0, 0.0, 0.0 };
            for ( int u = 0; u < 4; ++u )
               for ( int r = 0; r < 3; ++r )
                  xq[r] += lamQ[q][u] * X[u][r];
            double D[3][3], Di[3][3];
            shellDF( sm, xq, D );
            const double detD = fabs( inv3( D, Di ) );
            double       P[4][3];
            for ( int i = 0; i < 4; ++i )
               for ( int r = 0; r < 3; ++r )
                  P[i][r] = G[i][0] * Di[0][r] + G[i][1] * Di[1][r] + G[i][2] * Di[2][r];
            const double f = wQ[q] * detD;
            for ( int j = 0; j < 4; ++j )
               row[j] += f * ( P[k][0] * P[j][0] + P[k][1] * P[j][1] + P[k][2] * P[j][2] );
         }
         for ( int j = 0; j < 4; ++j )
            acc[kMCDir[t][k][j]] += detJ * row[j];
      }
   double* out = W + i64( c ) * NDIR * S + s;
   for ( int d = 0; d < NDIR; ++d )
      out[i64( d ) * S] = acc[d];
}

} // namespace

void MeshCtx::setShellMap()
{
   std::vector< ShellMap > v;
   for ( int c = 0; c < mesh->nC(); ++c )
      v.push_back( classifyShellCell( *mesh, c, 0 ) );
   ShellMap* d = nullptr;
   if ( hipMalloc( &d, v.size() * sizeof( ShellMap ) ) != hipSuccess ||
        hipMemcpy( d, v.data(), v.size() * sizeof( ShellMap ), hipMemcpyHostToDevice ) != hipSuccess )
   {
      (void) hipFree( d );
      throw std::runtime_error( "uploading the shell map failed" );
   }
   (void) hipFree( shell );
   shell     = d;
   shellHost = v;
}

void assembleBlendWeights( const MeshCtx& m, hyg_op_s& op, hipStream_t s )
{
   for ( int L = op.minL; L <= op.maxL; ++L )
   {
      hyg_op_s::Level& ol = op.levels[L];
      const int        w  = ( 1 << L ) + 1;
      const size_t     n  = size_t( m.mesh->nC() ) * NDIR * size_t( S3i( w ) );
      if ( hipMalloc( &ol.bw, n * sizeof( double ) ) != hipSuccess || hipMalloc( &ol.bwF, n * sizeof( float ) ) != hipSuccess )
         throw std::runtime_error( "hipMalloc failed for the blended stencil weights" );
      const int tiles = ceilDiv( S3i( w ), kBlock );
      hipLaunchKernelGGL( kBlendWeights, dim3( m.mesh->nC() * tiles ), dim3( kBlock ), 0, s, L, tiles, ol.elem, m.shell, ol.bw );
      launchConvert( i64( n ), 4, ol.bwF, 8, ol.bw, s );
   }
}

} // namespace hyg
