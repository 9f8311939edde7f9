// C ABI of the library (declared in include/hyteg_amd.h): handles, host<->device transfer,
// operators, grid transfer and the multigrid / defect-correction drivers.
#include "../../include/hyteg_amd.h"
#include "kernels.hpp"

#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>

namespace hyg {
void assembleP1Const( hyg_op_s& op );
void assembleP1DivKGrad( hyg_op_s& op );
void assembleP2Const( hyg_op_s& op );
void assembleP2DivKGrad( hyg_op_s& op );
void primEntriesPos( const LowerEntry& e, int N, int i, int j, int& qx, int& qy, int& qz );
} // namespace hyg

using namespace hyg;

// partial sums of hyg_dot_noalloc (HYG_NOALLOC_DOT_PARTIALS values, 2 MiB per device)
static __device__ __align__( 256 ) double scratch[HYG_NOALLOC_DOT_PARTIALS];

namespace {

thread_local std::string g_err;
int                      g_device = -1;

// Fixed reduction scratch of hyg_dot_noalloc: one partial sum per dot block. Every device the library runs on
// holds its own copy until the process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (kept in g_scratchDev by hyg_init).
constexpr int kMaxDevices      = 64;
double*       g_scratchDev[kMaxDevices] = {};

// per-process reduction scratch (dot products); calls that overlap on one device race on it
struct DotScratch
{
   double* partials = nullptr;
   size_t  cap      = 0;
   double* result   = nullptr;
   double* host     = nullptr; // pinned
} g_dot;

int fail( int code, const std::string& msg )
{
   g_err = msg;
   return code;
}

#define HYG_HIP( call )                                                                                                         \
   do                                                                                                                          \
   {                                                                                                                           \
      hipError_t e_ = ( call );                                                                                                \
      if ( e_ != hipSuccess )                                                                                                  \
         return fail( HYG_ERR_HIP, std::string( #call ) + ": " + hipGetErrorString( e_ ) );                                   \
   } while ( 0 )

#define HYG_TRY( expr )                                                                                                         \
   do                                                                                                                          \
   {                                                                                                                           \
      int rc_ = ( expr );                                                                                                      \
      if ( rc_ != HYG_OK )                                                                                                     \
         return rc_;                                                                                                           \
   } while ( 0 )

inline hipStream_t S( void* s ) { return static_cast< hipStream_t >( s ); }

int checkInit()
{
   if ( g_device < 0 )
      return fail( HYG_ERR_STATE, "hyg_init has not been called" );
   return HYG_OK;
}

int launched( const char* what )
{
   hipError_t e = hipGetLastError();
   if ( e != hipSuccess )
      return fail( HYG_ERR_HIP, std::string( what ) + ": kernel launch failed: " + hipGetErrorString( e ) );
   return HYG_OK;
}

int checkLevel( const hyg_func_s* f, int L )
{
   if ( !f )
      return fail( HYG_ERR_ARG, "null function handle" );
   if ( L < f->minL || L > f->maxL )
      return fail( HYG_ERR_ARG, "level " + std::to_string( L ) + " outside the function's range [" + std::to_string( f->minL ) +
                                     ", " + std::to_string( f->maxL ) + "]" );
   return HYG_OK;
}

int checkOpLevel( const hyg_op_s* op, int L )
{
   if ( !op )
      return fail( HYG_ERR_ARG, "null operator handle" );
   if ( L < op->minL || L > op->maxL )
      return fail( HYG_ERR_ARG, "level " + std::to_string( L ) + " outside the operator's range" );
   return HYG_OK;
}

int checkPair( const hyg_func_s* a, const hyg_func_s* b, int L )
{
   HYG_TRY( checkLevel( a, L ) );
   HYG_TRY( checkLevel( b, L ) );
   if ( a->mesh != b->mesh )
      return fail( HYG_ERR_ARG, "functions live on different meshes" );
   if ( a->dtype != b->dtype )
      return fail( HYG_ERR_ARG, "functions have different value types" );
   if ( a->lat != b->lat )
      return fail( HYG_ERR_ARG, "P1 and P2 functions cannot be combined" );
   return HYG_OK;
}

// Level arguments of the C ABI are levels of the handle's space; a P2 function or operator of level L
// works on the vertex lattice of level L+1, which every internal routine takes.
inline int lat( const hyg_func_s* f, int level ) { return f ? level + f->lat : level; }
inline int lat( const hyg_op_s* op, int level ) { return op ? level + op->lat : level; }

int checkSpace( const hyg_op_s* op, const hyg_func_s* f )
{
   if ( op && f && op->lat != f->lat )
      return fail( HYG_ERR_ARG, "P1 and P2 operators and functions cannot be combined" );
   return HYG_OK;
}

// an operator and a function it is applied to: same mesh (the kernels walk the operator's cells over the
// function's arrays) and same space
int checkOpFunc( const hyg_op_s* op, const hyg_func_s* f )
{
   if ( !op || !f )
      return fail( HYG_ERR_ARG, "null operator or function handle" );
   if ( f->mesh != op->mesh )
      return fail( HYG_ERR_ARG, "function and operator live on different meshes" );
   return checkSpace( op, f );
}

i64 levelValues( const hyg_func_s* f, int L ) { return i64( f->mesh->ctx.mesh->nC() ) * cellStride( L ); }

int createFunc( hyg_mesh_s* mesh, int minL, int maxL, int dtype, hyg_func_s** out, int space = 0 )
{
   auto* f  = new ( std::nothrow ) hyg_func_s;
   if ( !f )
      return fail( HYG_ERR_ARG, "out of host memory" );
   static uint64_t nextGen = 0;
   f->gen   = ++nextGen;
   f->lat   = space;
   f->mesh  = mesh;
   f->minL  = minL;
   f->maxL  = maxL;
   f->dtype = dtype;
   f->data.assign( maxL + 1, nullptr );
   for ( int L = minL; L <= maxL; ++L )
   {
      mesh->ctx.geom( L );
      const size_t bytes = size_t( levelValues( f, L ) + kTailSlack ) * f->elem();
      hipError_t   e     = hipMalloc( &f->data[L], bytes );
      if ( e == hipSuccess )
         e = hipMemset( f->data[L], 0, bytes );
      if ( e != hipSuccess )
      {
         for ( void* p : f->data )
            (void) hipFree( p );
         delete f;
         return fail( HYG_ERR_HIP, std::string( "allocating function memory: " ) + hipGetErrorString( e ) );
      }
   }
   *out = f;
   return HYG_OK;
}

void destroyFunc( hyg_func_s* f )
{
   if ( !f )
      return;
   for ( void* p : f->data )
      (void) hipFree( p );
   delete f;
}

struct FuncGuard
{
   hyg_func_s* f = nullptr;
   ~FuncGuard() { destroyFunc( f ); }
};

// visit every canonical DoF: fn(canonicalIndex, prim or -1, entry list / cell, i, j, k)
template < typename Fn >
void forEachCanonical( const MeshCtx& mc, int L, Fn&& fn )
{
   const Mesh& m = *mc.mesh;
   const int   N = 1 << L;
   i64         n = 0;
   for ( size_t p = 0; p < mc.primsHost.size(); ++p )
   {
      const LowerPrim& pr = mc.primsHost[p];
      if ( pr.kind == 0 )
         fn( n++, int( p ), 0, 0, 0 );
      else if ( pr.kind == 1 )
         for ( int i = 1; i <= N - 1; ++i )
            fn( n++, int( p ), i, 0, 0 );
      else
         for ( int j = 1; j <= N - 2; ++j )
            for ( int i = 1; i + j <= N - 1; ++i )
               fn( n++, int( p ), i, j, 0 );
   }
   for ( int c = 0; c < m.nC(); ++c )
      for ( int z = 1; z <= N - 3; ++z )
         for ( int y = 1; y + z <= N - 2; ++y )
            for ( int x = 1; x + y + z <= N - 1; ++x )
               fn( n++, -1 - c, x, y, z );
}

int ensureDotScratch( size_t n )
{
   if ( g_dot.cap >= n )
      return HYG_OK;
   (void) hipFree( g_dot.partials );
   g_dot.partials = nullptr;
   HYG_HIP( hipMalloc( &g_dot.partials, n * sizeof( double ) ) );
   g_dot.cap = n;
   if ( !g_dot.result )
      HYG_HIP( hipMalloc( &g_dot.result, sizeof( double ) ) );
   if ( !g_dot.host )
      HYG_HIP( hipHostMalloc( (void**) &g_dot.host, sizeof( double ), hipHostMallocDefault ) );
   return HYG_OK;
}

int dotImpl( hyg_func_s* a, hyg_func_s* b, int L, int flag, double* out, hipStream_t st )
{
   MeshCtx&   mc = a->mesh->ctx;
   LevelGeom& g  = mc.geom( L );
   HYG_TRY( ensureDotScratch( size_t( dotPartialsNeeded( mc, g ) ) ) );
   launchDot( mc, g, L, a->dtype, a->data[L], b->data[L], flag, g_dot.partials, g_dot.result, st );
   HYG_TRY( launched( "dot" ) );
   HYG_HIP( hipMemcpyAsync( g_dot.host, g_dot.result, sizeof( double ), hipMemcpyDeviceToHost, st ) );
   HYG_HIP( hipStreamSynchronize( st ) );
   *out = *g_dot.host;
   return HYG_OK;
}

int lincombImpl( hyg_func_s* dst, int n, const double* s, hyg_func_s* const* src, bool acc, double value, int L, int flag,
                 hipStream_t st )
{
   MeshCtx&    mc = dst->mesh->ctx;
   const void* p[3] = { nullptr, nullptr, nullptr };
   for ( int k = 0; k < n; ++k )
      p[k] = src[k]->data[L];
   launchLincomb( mc, mc.geom( L ), L, dst->dtype, dst->data[L], n, s, p, acc, value, flag, st );
   return launched( "lincomb" );
}

int stencilImpl( hyg_op_s* op, int L, int dtype, const StencilCall& c, hipStream_t st )
{
   MeshCtx& mc = op->mesh->ctx;
   if ( op->kind == 1 )
      launchElementwise( mc, mc.geom( L ), *op, L, dtype, c, st );
   else if ( op->kind == 2 || op->kind == 3 )
      launchP2( mc, mc.geom( L ), *op, L, dtype, c, st );
   else
      launchStencil( mc, mc.geom( L ), op->levels[L], L, dtype, c, st );
   return launched( "stencil" );
}

} // namespace

// ================================================================== GMG object
struct hyg_gmg_s
{
   hyg_op_s*             op;
   int                   dtype;
   int                   minL, maxL, pre, post, inc, smoother, chebOrder, cgMaxIter;
   double                relax, chebLo, chebHi, cgTol;
   double                cheb[5];
   hyg_func_s*           tmp = nullptr, *jt = nullptr, *t1 = nullptr, *t2 = nullptr, *cy = nullptr;
   hyg_func_s*           cgp = nullptr, *cgz = nullptr, *cgap = nullptr, *cgr = nullptr;
   hyg_func_s*           x32 = nullptr, *b32 = nullptr, *r64 = nullptr, *t64 = nullptr;
   // captured V-cycle: device-side coarse CG scalars and partial sums, the graph and its arguments
   int                   cgGraphIters = 0;
   double*               cgDev        = nullptr; // rs (two slots), pAp
   double*               cgPartials   = nullptr;
   hipGraph_t            graph        = nullptr;
   hipGraphExec_t        graphExec    = nullptr;
   uint64_t              graphX = 0, graphB = 0; // creation numbers of the captured x and b
   int                   graphLevel   = -1;
   ~hyg_gmg_s()
   {
      for ( hyg_func_s* f : { tmp, jt, t1, t2, cy, cgp, cgz, cgap, cgr, x32, b32, r64, t64 } )
         destroyFunc( f );
      if ( graphExec )
         (void) hipGraphExecDestroy( graphExec );
      if ( graph )
         (void) hipGraphDestroy( graph );
      (void) hipFree( cgDev );
      (void) hipFree( cgPartials );
   }
};

namespace {

const int kGmgFlag = HYG_INNER | HYG_NEUMANN;

int smoothImpl( hyg_gmg_s* g, hyg_func_s* x, hyg_func_s* b, int L, hipStream_t st )
{
   StencilCall c;
   c.flag = kGmgFlag;
   if ( g->smoother == 0 )
   {
      // WeightedJacobiSmoother::solve: tmp := x (All); smooth_jac(x, b, tmp)
      launchCopy( levelValues( x, L ), g->dtype, g->jt->data[L], x->data[L], st );
      HYG_TRY( launched( "copy" ) );
      c.mode = MODE_JACOBI;
      c.x    = g->jt->data[L];
      c.in1  = b->data[L];
      c.out0 = x->data[L];
      c.a    = g->relax;
      return stencilImpl( g->op, L, g->dtype, c, st );
   }
   if ( g->smoother >= 4 )
   {
      // the same in the reference's own order (smooth_sor_ordered)
      MeshCtx& mc = g->op->mesh->ctx;
      launchSorOrdered( mc, mc.geom( L ), g->op->levels[L], L, g->dtype, x->data[L], b->data[L], g->relax, kGmgFlag, false, st );
      if ( g->smoother == 5 )
         launchSorOrdered( mc, mc.geom( L ), g->op->levels[L], L, g->dtype, x->data[L], b->data[L], g->relax, kGmgFlag, true, st );
      return launched( "ordered sor" );
   }
   if ( g->smoother >= 2 )
   {
      // SORSmoother::solve: smooth_sor(x, b, relax, level, flag), multi-colour order; SymmetricSORSmoother::solve
      // follows it with smooth_sor_backwards
      MeshCtx& mc = g->op->mesh->ctx;
      launchSor( mc, mc.geom( L ), g->op->levels[L], L, g->dtype, x->data[L], b->data[L], g->relax, kGmgFlag, false, st );
      if ( g->smoother == 3 )
         launchSor( mc, mc.geom( L ), g->op->levels[L], L, g->dtype, x->data[L], b->data[L], g->relax, kGmgFlag, true, st );
      return launched( "sor" );
   }
   // ChebyshevSmoother::solve, fused: t0 = D^-1 (b - A x), y = x + c0 t0; t_k = D^-1 A t_{k-1}, y += c_k t_k; x = y
   const int K = g->chebOrder;
   c.mode      = MODE_CHEB0;
   c.x         = x->data[L];
   c.in1       = b->data[L];
   c.out0      = g->t1->data[L];
   c.out1      = g->cy->data[L];
   c.a         = g->cheb[0];
   HYG_TRY( stencilImpl( g->op, L, g->dtype, c, st ) );
   hyg_func_s* tin  = g->t1;
   hyg_func_s* tout = g->t2;
   for ( int k = 1; k < K; ++k )
   {
      c.mode = MODE_CHEBK;
      c.x    = tin->data[L];
      c.in1  = g->cy->data[L];
      c.out0 = tout->data[L];
      c.out1 = k == K - 1 ? x->data[L] : g->cy->data[L];
      c.a    = g->cheb[k];
      HYG_TRY( stencilImpl( g->op, L, g->dtype, c, st ) );
      std::swap( tin, tout );
   }
   if ( K == 1 )
   {
      const double one = 1.0;
      hyg_func_s*  src = g->cy;
      return lincombImpl( x, 1, &one, &src, false, 0.0, L, kGmgFlag, st );
   }
   return HYG_OK;
}

// CGSolver::solve with the identity preconditioner (z = r)
int cgImpl( hyg_gmg_s* g, hyg_func_s* x, hyg_func_s* b, int L, hipStream_t st )
{
   const int   flag = HYG_INNER | HYG_NEUMANN | HYG_FREESLIP;
   StencilCall c;
   c.flag = flag;
   c.mode = MODE_RESIDUAL;
   c.x    = x->data[L];
   c.in1  = b->data[L];
   c.out0 = g->cgr->data[L]; // r = b - A x
   HYG_TRY( stencilImpl( g->op, L, g->dtype, c, st ) );
   const double one = 1.0;
   hyg_func_s*  r   = g->cgr;
   HYG_TRY( lincombImpl( g->cgp, 1, &one, &r, false, 0.0, L, flag, st ) ); // p = z = r
   double prsold = 0.0;
   HYG_TRY( dotImpl( g->cgr, g->cgr, L, flag, &prsold, st ) );
   if ( std::sqrt( prsold ) < g->cgTol )
      return HYG_OK;
   for ( int it = 0; it < g->cgMaxIter; ++it )
   {
      c.mode = MODE_APPLY;
      c.x    = g->cgp->data[L];
      c.out0 = g->cgap->data[L];
      HYG_TRY( stencilImpl( g->op, L, g->dtype, c, st ) );
      double pAp = 0.0;
      HYG_TRY( dotImpl( g->cgp, g->cgap, L, flag, &pAp, st ) );
      const double alpha = prsold / pAp;
      hyg_func_s*  p     = g->cgp;
      hyg_func_s*  ap    = g->cgap;
      HYG_TRY( lincombImpl( x, 1, &alpha, &p, true, 0.0, L, flag, st ) );
      const double malpha = -alpha;
      HYG_TRY( lincombImpl( g->cgr, 1, &malpha, &ap, true, 0.0, L, flag, st ) );
      double rsnew = 0.0;
      HYG_TRY( dotImpl( g->cgr, g->cgr, L, flag, &rsnew, st ) );
      if ( std::sqrt( rsnew ) < g->cgTol )
         break;
      const double beta = rsnew / prsold;
      // p = z + beta p with z = r
      const double sc[2]  = { 1.0, beta };
      hyg_func_s*  srcs[2] = { g->cgr, g->cgp };
      HYG_TRY( lincombImpl( g->cgp, 2, sc, srcs, false, 0.0, L, flag, st ) );
      prsold = rsnew;
   }
   return HYG_OK;
}

// CGSolver::solve for a captured V-cycle: a fixed number of iterations (cgGraphIters) whose step
// lengths come from device memory; once the residual norm is below the tolerance the remaining
// iterations take no step, which reproduces the early exit of cgImpl.
int cgDeviceImpl( hyg_gmg_s* g, hyg_func_s* x, hyg_func_s* b, int L, hipStream_t st )
{
   MeshCtx&     mc   = g->op->mesh->ctx;
   LevelGeom&   gl   = mc.geom( L );
   const int    flag = HYG_INNER | HYG_NEUMANN | HYG_FREESLIP;
   const double tol2 = g->cgTol * g->cgTol;
   StencilCall  c;
   c.flag = flag;
   c.mode = MODE_RESIDUAL;
   c.x    = x->data[L];
   c.in1  = b->data[L];
   c.out0 = g->cgr->data[L];
   HYG_TRY( stencilImpl( g->op, L, g->dtype, c, st ) );
   const double one = 1.0;
   hyg_func_s*  r   = g->cgr;
   HYG_TRY( lincombImpl( g->cgp, 1, &one, &r, false, 0.0, L, flag, st ) );
   launchDot( mc, gl, L, g->dtype, g->cgr->data[L], g->cgr->data[L], flag, g->cgPartials, g->cgDev + 0, st );
   for ( int it = 0; it < g->cgGraphIters; ++it )
   {
      const int o = it & 1, nw = 1 - o;
      c.mode      = MODE_APPLY;
      c.x         = g->cgp->data[L];
      c.out0      = g->cgap->data[L];
      HYG_TRY( stencilImpl( g->op, L, g->dtype, c, st ) );
      launchDot( mc, gl, L, g->dtype, g->cgp->data[L], g->cgap->data[L], flag, g->cgPartials, g->cgDev + 2, st );
      launchCgUpdate( mc, gl, L, g->dtype, x->data[L], g->cgr->data[L], g->cgp->data[L], g->cgap->data[L], g->cgDev, o, 2, tol2,
                      flag, st );
      launchDot( mc, gl, L, g->dtype, g->cgr->data[L], g->cgr->data[L], flag, g->cgPartials, g->cgDev + nw, st );
      launchCgDirection( mc, gl, L, g->dtype, g->cgp->data[L], g->cgr->data[L], g->cgDev, o, nw, tol2, flag, st );
      HYG_TRY( launched( "coarse CG" ) );
   }
   return HYG_OK;
}

// GeometricMultigridSolver::solveRecursively, V-cycle (deviceCg: coarse CG without host round trips)
int vcycleImpl( hyg_gmg_s* g, hyg_func_s* x, hyg_func_s* b, int L, int invoked, hipStream_t st, bool deviceCg = false )
{
   if ( L == g->minL )
      return deviceCg ? cgDeviceImpl( g, x, b, L, st ) : cgImpl( g, x, b, L, st );
   MeshCtx&  mc   = g->op->mesh->ctx;
   const int preS = g->pre + g->inc * ( invoked - L );
   for ( int i = 0; i < preS; ++i )
      HYG_TRY( smoothImpl( g, x, b, L, st ) );
   StencilCall c;
   c.flag = kGmgFlag;
   c.mode = MODE_RESIDUAL;
   c.x    = x->data[L];
   c.in1  = b->data[L];
   c.out0 = g->tmp->data[L];
   HYG_TRY( stencilImpl( g->op, L, g->dtype, c, st ) );
   if ( g->op->lat )
      launchP2Restrict( mc, mc.geom( L - 1 ), mc.p2Tables(), L - 1, g->dtype, g->tmp->data[L - 1], g->tmp->data[L], kGmgFlag, st );
   else
      launchRestrict( mc, mc.geom( L - 1 ), L - 1, g->dtype, g->tmp->data[L - 1], g->tmp->data[L], kGmgFlag, st );
   HYG_TRY( launched( "restrict" ) );
   const double one = 1.0;
   hyg_func_s*  t   = g->tmp;
   HYG_TRY( lincombImpl( b, 1, &one, &t, false, 0.0, L - 1, kGmgFlag, st ) );
   HYG_TRY( lincombImpl( x, 0, nullptr, nullptr, false, 0.0, L - 1, HYG_ALL, st ) );
   HYG_TRY( vcycleImpl( g, x, b, L - 1, invoked, st, deviceCg ) );
   if ( g->op->lat )
      launchP2Prolongate( mc, mc.geom( L ), mc.p2Tables(), L, g->dtype, x->data[L], x->data[L - 1], kGmgFlag, true, st );
   else
      launchProlongate( mc, mc.geom( L ), L, g->dtype, x->data[L], x->data[L - 1], kGmgFlag, true, st );
   HYG_TRY( launched( "prolongate" ) );
   const int postS = g->post + g->inc * ( invoked - L );
   for ( int i = 0; i < postS; ++i )
      HYG_TRY( smoothImpl( g, x, b, L, st ) );
   return HYG_OK;
}

int residualNorm( hyg_gmg_s* g, hyg_func_s* x, hyg_func_s* b, hyg_func_s* r, int L, double* out, hipStream_t st )
{
   StencilCall c;
   c.flag = kGmgFlag;
   c.mode = MODE_RESIDUAL;
   c.x    = x->data[L];
   c.in1  = b->data[L];
   c.out0 = r->data[L];
   HYG_TRY( stencilImpl( g->op, L, r->dtype, c, st ) );
   double d = 0.0;
   HYG_TRY( dotImpl( r, r, L, kGmgFlag, &d, st ) );
   *out = std::sqrt( d );
   if ( !std::isfinite( *out ) )
      return fail( HYG_ERR_NUMERIC, "non-finite residual norm" );
   return HYG_OK;
}

int checkGmgArgs( hyg_gmg_s* g, hyg_func_s* x, hyg_func_s* b, int L )
{
   if ( !g )
      return fail( HYG_ERR_ARG, "null solver handle" );
   HYG_TRY( checkPair( x, b, L ) );
   if ( x->dtype != g->dtype )
      return fail( HYG_ERR_ARG, "function value type differs from the solver's" );
   if ( L != g->maxL || x->minL > g->minL || b->minL > g->minL )
      return fail( HYG_ERR_ARG, "solve level must be the solver's max level and functions must cover its levels" );
   if ( x->mesh != g->op->mesh )
      return fail( HYG_ERR_ARG, "functions and solver live on different meshes" );
   HYG_TRY( checkSpace( g->op, x ) );
   HYG_TRY( checkSpace( g->op, b ) );
   return HYG_OK;
}

} // namespace

// ================================================================== C ABI
extern "C" {

const char* hyg_last_error( void ) { return g_err.c_str(); }

int hyg_init( int device )
{
   int n = 0;
   HYG_HIP( hipGetDeviceCount( &n ) );
   if ( device < 0 || device >= n )
      return fail( HYG_ERR_ARG, "device index out of range" );
   HYG_HIP( hipSetDevice( device ) );
   if ( device < kMaxDevices && !g_scratchDev[device] )
   {
      void* p = nullptr;
      HYG_HIP( hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) ) );
      g_scratchDev[device] = static_cast< double* >( p );
   }
   g_device = device;
   return HYG_OK;
}

int hyg_mesh_create_from_msh( const char* path, hyg_mesh* out ) { return hyg_mesh_create_from_msh_tags( path, 0, out ); }

int hyg_mesh_create_from_msh_tags( const char* path, int keep_tags, hyg_mesh* out )
{
   HYG_TRY( checkInit() );
   if ( !path || !out )
      return fail( HYG_ERR_ARG, "null argument" );
   std::string err;
   auto        m = readGmsh( path, err, keep_tags != 0 );
   if ( !m )
      return fail( HYG_ERR_IO, err );
   auto* h = new hyg_mesh_s;
   h->ctx.mesh   = std::move( m );
   h->ctx.device = g_device;
   try
   {
      h->ctx.buildTopology();
   } catch ( const std::exception& e )
   {
      delete h;
      return fail( HYG_ERR_HIP, e.what() );
   }
   *out = h;
   return HYG_OK;
}

int hyg_mesh_create_spherical_shell( int ntan, int nrad, double rmin, double rmax, hyg_mesh* out )
{
   HYG_TRY( checkInit() );
   if ( !out )
      return fail( HYG_ERR_ARG, "null argument" );
   std::string err;
   auto        m = meshSphericalShell( ntan, nrad, rmin, rmax, err );
   if ( !m )
      return fail( HYG_ERR_ARG, err );
   auto* h = new hyg_mesh_s;
   h->ctx.mesh   = std::move( m );
   h->ctx.device = g_device;
   try
   {
      h->ctx.buildTopology();
   } catch ( const std::exception& e )
   {
      delete h;
      return fail( HYG_ERR_HIP, e.what() );
   }
   *out = h;
   return HYG_OK;
}

int hyg_mesh_create_symmetric_cuboid( const double lower[3], const double upper[3], int nx, int ny, int nz, hyg_mesh* out )
{
   HYG_TRY( checkInit() );
   if ( !out || !lower || !upper )
      return fail( HYG_ERR_ARG, "null argument" );
   std::string err;
   auto        m = meshSymmetricCuboid( lower, upper, nx, ny, nz, err );
   if ( !m )
      return fail( HYG_ERR_ARG, err );
   auto* h = new hyg_mesh_s;
   h->ctx.mesh   = std::move( m );
   h->ctx.device = g_device;
   try
   {
      h->ctx.buildTopology();
   } catch ( const std::exception& e )
   {
      delete h;
      return fail( HYG_ERR_HIP, e.what() );
   }
   *out = h;
   return HYG_OK;
}

int hyg_mesh_create_cuboid( const double lower[3], const double upper[3], int nx, int ny, int nz, hyg_mesh* out )
{
   HYG_TRY( checkInit() );
   if ( !out || !lower || !upper )
      return fail( HYG_ERR_ARG, "null argument" );
   std::string err;
   auto        m = meshCuboid( lower, upper, nx, ny, nz, err );
   if ( !m )
      return fail( HYG_ERR_ARG, err );
   auto* h = new hyg_mesh_s;
   h->ctx.mesh   = std::move( m );
   h->ctx.device = g_device;
   try
   {
      h->ctx.buildTopology();
   } catch ( const std::exception& e )
   {
      delete h;
      return fail( HYG_ERR_HIP, e.what() );
   }
   *out = h;
   return HYG_OK;
}

int hyg_mesh_create_cubed_domain( const int* cubes, int ncubes, int cube_type, hyg_mesh* out )
{
   HYG_TRY( checkInit() );
   if ( !out || !cubes )
      return fail( HYG_ERR_ARG, "null argument" );
   std::string err;
   auto        m = meshCubedDomain( cubes, ncubes, cube_type, err );
   if ( !m )
      return fail( HYG_ERR_ARG, err );
   auto* h = new hyg_mesh_s;
   h->ctx.mesh   = std::move( m );
   h->ctx.device = g_device;
   try
   {
      h->ctx.buildTopology();
   } catch ( const std::exception& e )
   {
      delete h;
      return fail( HYG_ERR_HIP, e.what() );
   }
   *out = h;
   return HYG_OK;
}

int hyg_mesh_destroy( hyg_mesh mesh )
{
   delete mesh;
   return HYG_OK;
}

int hyg_mesh_info( hyg_mesh mesh, int64_t* out )
{
   if ( !mesh || !out )
      return fail( HYG_ERR_ARG, "null argument" );
   const Mesh& m = *mesh->ctx.mesh;
   out[0]        = m.nV();
   out[1]        = m.nE();
   out[2]        = m.nF();
   out[3]        = m.nC();
   out[4]        = std::count( m.vB.begin(), m.vB.end(), 1 );
   out[5]        = std::count( m.eB.begin(), m.eB.end(), 1 );
   out[6]        = std::count( m.fB.begin(), m.fB.end(), 1 );
   return HYG_OK;
}

int hyg_mesh_num_dofs( hyg_mesh mesh, int level, int64_t* out )
{
   if ( !mesh || !out || level < 0 || level > 10 )
      return fail( HYG_ERR_ARG, "bad argument" );
   *out = numDofs( *mesh->ctx.mesh, level );
   return HYG_OK;
}

int hyg_mesh_dof_coords( hyg_mesh mesh, int level, double* xyz, uint8_t* boundary )
{
   if ( !mesh || ( !xyz && !boundary ) || level < 0 || level > 10 )
      return fail( HYG_ERR_ARG, "bad argument" );
   const MeshCtx& mc = mesh->ctx;
   const Mesh&    m  = *mc.mesh;
   const int      N  = 1 << level;
   const double   h  = 1.0 / double( N );
   forEachCanonical( mc, level, [&]( i64 n, int p, int i, int j, int k ) {
      double r[3];
      int    bnd = 0;
      if ( p >= 0 )
      {
         const LowerPrim& pr = mc.primsHost[p];
         int              v[3];
         if ( pr.kind == 0 )
            v[0] = v[1] = v[2] = p;
         else if ( pr.kind == 1 )
         {
            const auto& e = m.E[p - m.nV()];
            v[0] = e[0], v[1] = e[1], v[2] = e[0];
         }
         else
         {
            const auto& f = m.F[p - m.nV() - m.nE()];
            v[0] = f[0], v[1] = f[1], v[2] = f[2];
         }
         for ( int a = 0; a < 3; ++a )
            r[a] = m.X[v[0]][a] + ( m.X[v[1]][a] - m.X[v[0]][a] ) * h * i + ( m.X[v[2]][a] - m.X[v[0]][a] ) * h * j;
         bnd = pr.bc == HYG_DIRICHLET;
      }
      else
      {
         const auto& cv = m.Cv[-1 - p];
         for ( int a = 0; a < 3; ++a )
            r[a] = m.X[cv[0]][a] + ( m.X[cv[1]][a] - m.X[cv[0]][a] ) * h * i + ( m.X[cv[2]][a] - m.X[cv[0]][a] ) * h * j +
                   ( m.X[cv[3]][a] - m.X[cv[0]][a] ) * h * k;
      }
      // blended coordinates (hyg_mesh_set_shell_map): the map of the edge's / face's first cell (the reference's
      // IcosahedralShellMap( Edge / Face ) takes neighborCells[0]) or of the cell itself; macro-vertices stay
      if ( !mc.shellHost.empty() && !( p >= 0 && mc.primsHost[p].kind == 0 ) )
      {
         int cell = -1 - p;
         if ( p >= 0 )
            cell = mc.primsHost[p].kind == 1 ? m.eC[p - m.nV()][0] : m.fC[p - m.nV() - m.nE()][0];
         const double xa[3] = { r[0], r[1], r[2] };
         shellF( mc.shellHost[cell], xa, r );
      }
      if ( xyz )
         xyz[3 * n] = r[0], xyz[3 * n + 1] = r[1], xyz[3 * n + 2] = r[2];
      if ( boundary )
         boundary[n] = uint8_t( bnd );
   } );
   return HYG_OK;
}

int hyg_func_create( hyg_mesh mesh, int min_level, int max_level, int dtype, hyg_func* out )
{
   HYG_TRY( checkInit() );
   if ( !mesh || !out )
      return fail( HYG_ERR_ARG, "null argument" );
   if ( min_level < 0 || max_level < min_level || max_level > 10 )
      return fail( HYG_ERR_ARG, "bad level range" );
   if ( dtype != HYG_F64 && dtype != HYG_F32 )
      return fail( HYG_ERR_ARG, "dtype must be HYG_F64 or HYG_F32" );
   return createFunc( mesh, min_level, max_level, dtype, out );
}

int hyg_func_create_p2( hyg_mesh mesh, int min_level, int max_level, int dtype, hyg_func* out )
{
   HYG_TRY( checkInit() );
   if ( !mesh || !out )
      return fail( HYG_ERR_ARG, "null argument" );
   if ( min_level < 0 || max_level < min_level || max_level > 9 )
      return fail( HYG_ERR_ARG, "bad level range" );
   if ( dtype != HYG_F64 && dtype != HYG_F32 )
      return fail( HYG_ERR_ARG, "dtype must be HYG_F64 or HYG_F32" );
   return createFunc( mesh, min_level + 1, max_level + 1, dtype, out, 1 );
}

int hyg_func_destroy( hyg_func f )
{
   destroyFunc( f );
   return HYG_OK;
}

int hyg_func_upload( hyg_func f, int level, const double* canonical, void* stream )
{
   level = lat( f, level );
   HYG_TRY( checkLevel( f, level ) );
   if ( !canonical )
      return fail( HYG_ERR_ARG, "null data" );
   const MeshCtx&        mc = f->mesh->ctx;
   const int             N = 1 << level, w = N + 1;
   const i64             stride = cellStride( level );
   std::vector< double > h( size_t( levelValues( f, level ) ), 0.0 );
   forEachCanonical( mc, level, [&]( i64 n, int p, int i, int j, int k ) {
      if ( p >= 0 )
      {
         const LowerPrim& pr = mc.primsHost[p];
         for ( int e = 0; e < pr.nCells; ++e )
         {
            const LowerEntry& en = mc.entriesHost[pr.entry0 + e];
            int               qx, qy, qz;
            primEntriesPos( en, N, i, j, qx, qy, qz );
            h[size_t( en.cell * stride + cidx( w, qx, qy, qz ) )] = canonical[n];
         }
      }
      else
         h[size_t( ( -1 - p ) * stride + cidx( w, i, j, k ) )] = canonical[n];
   } );
   if ( f->dtype == HYG_F64 )
   {
      HYG_HIP( hipMemcpyAsync( f->data[level], h.data(), h.size() * 8, hipMemcpyHostToDevice, S( stream ) ) );
      HYG_HIP( hipStreamSynchronize( S( stream ) ) );
   }
   else
   {
      std::vector< float > hf( h.begin(), h.end() );
      HYG_HIP( hipMemcpyAsync( f->data[level], hf.data(), hf.size() * 4, hipMemcpyHostToDevice, S( stream ) ) );
      HYG_HIP( hipStreamSynchronize( S( stream ) ) );
   }
   return HYG_OK;
}

int hyg_func_download( hyg_func f, int level, double* canonical, void* stream )
{
   level = lat( f, level );
   HYG_TRY( checkLevel( f, level ) );
   if ( !canonical )
      return fail( HYG_ERR_ARG, "null data" );
   const MeshCtx&        mc = f->mesh->ctx;
   const int             N = 1 << level, w = N + 1;
   const i64             stride = cellStride( level );
   const size_t          nv     = size_t( levelValues( f, level ) );
   std::vector< double > h( nv );
   if ( f->dtype == HYG_F64 )
   {
      HYG_HIP( hipMemcpyAsync( h.data(), f->data[level], nv * 8, hipMemcpyDeviceToHost, S( stream ) ) );
      HYG_HIP( hipStreamSynchronize( S( stream ) ) );
   }
   else
   {
      std::vector< float > hf( nv );
      HYG_HIP( hipMemcpyAsync( hf.data(), f->data[level], nv * 4, hipMemcpyDeviceToHost, S( stream ) ) );
      HYG_HIP( hipStreamSynchronize( S( stream ) ) );
      std::copy( hf.begin(), hf.end(), h.begin() );
   }
   forEachCanonical( mc, level, [&]( i64 n, int p, int i, int j, int k ) {
      if ( p >= 0 )
      {
         const LowerPrim&  pr = mc.primsHost[p];
         const LowerEntry& en = mc.entriesHost[pr.entry0];
         int               qx, qy, qz;
         primEntriesPos( en, N, i, j, qx, qy, qz );
         canonical[n] = h[size_t( en.cell * stride + cidx( w, qx, qy, qz ) )];
      }
      else
         canonical[n] = h[size_t( ( -1 - p ) * stride + cidx( w, i, j, k ) )];
   } );
   return HYG_OK;
}

int hyg_interpolate_const( hyg_func f, double value, int level, int flag, void* stream )
{
   level = lat( f, level );
   HYG_TRY( checkLevel( f, level ) );
   return lincombImpl( f, 0, nullptr, nullptr, false, value, level, flag, S( stream ) );
}

int hyg_assign( hyg_func dst, int n, const double* scalars, const hyg_func* srcs, int level, int flag, void* stream )
{
   level = lat( dst, level );
   HYG_TRY( checkLevel( dst, level ) );
   if ( n < 1 || n > 3 || !scalars || !srcs )
      return fail( HYG_ERR_ARG, "assign takes 1 to 3 sources" );
   for ( int k = 0; k < n; ++k )
      HYG_TRY( checkPair( dst, srcs[k], level ) );
   return lincombImpl( dst, n, scalars, srcs, false, 0.0, level, flag, S( stream ) );
}

int hyg_add( hyg_func dst, int n, const double* scalars, const hyg_func* srcs, int level, int flag, void* stream )
{
   level = lat( dst, level );
   HYG_TRY( checkLevel( dst, level ) );
   if ( n < 1 || n > 3 || !scalars || !srcs )
      return fail( HYG_ERR_ARG, "add takes 1 to 3 sources" );
   for ( int k = 0; k < n; ++k )
      HYG_TRY( checkPair( dst, srcs[k], level ) );
   return lincombImpl( dst, n, scalars, srcs, true, 0.0, level, flag, S( stream ) );
}

int hyg_mult_elementwise( hyg_func dst, hyg_func a, hyg_func b, int level, int flag, void* stream )
{
   level = lat( dst, level );
   HYG_TRY( checkPair( dst, a, level ) );
   HYG_TRY( checkPair( dst, b, level ) );
   MeshCtx& mc = dst->mesh->ctx;
   launchMult( mc, mc.geom( level ), level, dst->dtype, dst->data[level], a->data[level], b->data[level], flag, S( stream ) );
   return launched( "multElementwise" );
}

int hyg_dot( hyg_func a, hyg_func b, int level, int flag, double* out, void* stream )
{
   level = lat( a, level );
   HYG_TRY( checkPair( a, b, level ) );
   if ( !out )
      return fail( HYG_ERR_ARG, "null output" );
   return dotImpl( a, b, level, flag, out, S( stream ) );
}

int hyg_dot_noalloc( hyg_func a, hyg_func b, int level, int flag, double* d_out, void* stream )
{
   level = lat( a, level );
   HYG_TRY( checkPair( a, b, level ) );
   if ( !d_out )
      return fail( HYG_ERR_ARG, "null output" );
   MeshCtx& mc = a->mesh->ctx;
   // the level tables exist (built when the functions were created), so geom() allocates nothing here
   LevelGeom& g      = mc.geom( level );
   const int  needed = dotPartialsNeeded( mc, g );
   if ( needed > HYG_NOALLOC_DOT_PARTIALS )
      return fail( HYG_ERR_ARG, "hyg_dot_noalloc: this level needs " + std::to_string( needed ) + " partial sums, more than the " +
                                    std::to_string( HYG_NOALLOC_DOT_PARTIALS ) + " of the fixed scratch (use hyg_dot)" );
   double* part = mc.device >= 0 && mc.device < kMaxDevices ? g_scratchDev[mc.device] : nullptr;
   if ( !part )
      return fail( HYG_ERR_STATE, "hyg_dot_noalloc: hyg_init was not called for the device of this mesh" );
   // every block writes its own partial sum, so nothing needs clearing between calls
   launchDot( mc, g, level, a->dtype, a->data[level], b->data[level], flag, part, d_out, S( stream ) );
   return launched( "dot_noalloc" );
}

int hyg_halo_layout( hyg_mesh mesh, int level, int64_t* offsets )
{
   if ( !mesh || !offsets || level < 0 || level > 10 )
      return fail( HYG_ERR_ARG, "bad argument" );
   const auto off = haloLayout( *mesh->ctx.mesh, level );
   std::copy( off.begin(), off.end(), offsets );
   return HYG_OK;
}

namespace {
int haloArgs( hyg_func f, int level, const void* buf )
{
   HYG_TRY( checkLevel( f, level ) );
   if ( f->lat != 0 )
      return fail( HYG_ERR_ARG, "reference-layout halos are defined for P1 functions" );
   if ( !buf )
      return fail( HYG_ERR_ARG, "null buffer" );
   return HYG_OK;
}
const HaloMaps* haloOf( hyg_func f, int level )
{
   try
   {
      return &f->mesh->ctx.halo( level );
   } catch ( const std::exception& e )
   {
      fail( HYG_ERR_HIP, std::string( "building the halo tables: " ) + e.what() );
      return nullptr;
   }
}
} // namespace

int hyg_halo_pack( hyg_func f, int level, void* out, void* stream )
{
   HYG_TRY( haloArgs( f, level, out ) );
   const HaloMaps* h = haloOf( f, level );
   if ( !h )
      return HYG_ERR_HIP;
   launchHaloGather( *h, f->dtype, f->data[level], out, S( stream ) );
   return launched( "halo pack" );
}

int hyg_halo_unpack( hyg_func f, int level, const void* in, void* stream )
{
   HYG_TRY( haloArgs( f, level, in ) );
   const HaloMaps* h = haloOf( f, level );
   if ( !h )
      return HYG_ERR_HIP;
   launchHaloScatter( *h, f->dtype, in, f->data[level], S( stream ) );
   return launched( "halo unpack" );
}

int hyg_halo_sum_replicas( hyg_func f, int level, int flag, void* stream )
{
   HYG_TRY( haloArgs( f, level, f ) );
   const HaloMaps* h = haloOf( f, level );
   if ( !h )
      return HYG_ERR_HIP;
   launchHaloSum( *h, f->dtype, flag, f->data[level], S( stream ) );
   return launched( "halo sum" );
}

int hyg_convert( hyg_func dst, hyg_func src, int level, void* stream )
{
   level = lat( dst, level );
   HYG_TRY( checkLevel( dst, level ) );
   HYG_TRY( checkLevel( src, level ) );
   if ( dst->mesh != src->mesh )
      return fail( HYG_ERR_ARG, "functions live on different meshes" );
   if ( dst->lat != src->lat )
      return fail( HYG_ERR_ARG, "P1 and P2 functions cannot be combined" );
   launchConvert( levelValues( dst, level ), dst->dtype, dst->data[level], src->dtype, src->data[level], S( stream ) );
   return launched( "convert" );
}

int hyg_op_create_p1_const( hyg_mesh mesh, int min_level, int max_level, hyg_op* out )
{
   HYG_TRY( checkInit() );
   if ( !mesh || !out || min_level < 0 || max_level < min_level || max_level > 10 )
      return fail( HYG_ERR_ARG, "bad argument" );
   auto* op = new hyg_op_s;
   op->mesh = mesh;
   op->minL = min_level;
   op->maxL = max_level;
   try
   {
      assembleP1Const( *op );
      for ( int L = min_level; L <= max_level; ++L )
         mesh->ctx.geom( L );
   } catch ( const std::exception& e )
   {
      delete op;
      return fail( HYG_ERR_HIP, e.what() );
   }
   for ( int L = min_level; L <= max_level; ++L )
   {
      const auto& ol = op->levels[L];
      for ( int c = 0; L >= 2 && c < mesh->ctx.mesh->nC(); ++c )
         if ( !std::isfinite( ol.cellStHost[16 * c + 15] ) )
         {
            delete op;
            return fail( HYG_ERR_NUMERIC, "zero diagonal in cell stencil (degenerate macro-cell)" );
         }
   }
   *out = op;
   return HYG_OK;
}

int hyg_op_create_p1_divkgrad( hyg_mesh mesh, int min_level, int max_level, double k0, const double* modes, int n_modes,
                                hyg_op* out )
{
   HYG_TRY( checkInit() );
   if ( !mesh || !out || min_level < 0 || max_level < min_level || max_level > 10 || n_modes < 0 || n_modes > 64 ||
        ( n_modes > 0 && !modes ) || !std::isfinite( k0 ) )
      return fail( HYG_ERR_ARG, "bad argument" );
   auto* op = new hyg_op_s;
   op->mesh = mesh;
   op->minL = min_level;
   op->maxL = max_level;
   op->kind = 1;
   op->k0   = k0;
   op->modesHost.assign( modes, modes + 5 * size_t( n_modes ) );
   try
   {
      for ( int L = min_level; L <= max_level; ++L )
         mesh->ctx.geom( L );
      assembleP1DivKGrad( *op );
      for ( int L = min_level; L <= max_level; ++L )
         launchElementwiseInverseDiagonal( mesh->ctx, mesh->ctx.geom( L ), *op, L, nullptr );
   } catch ( const std::exception& e )
   {
      delete op;
      return fail( HYG_ERR_HIP, e.what() );
   }
   if ( int rc = launched( "elementwise operator setup" ) )
   {
      delete op;
      return rc;
   }
   if ( hipDeviceSynchronize() != hipSuccess )
   {
      delete op;
      return fail( HYG_ERR_HIP, "elementwise operator setup failed" );
   }
   *out = op;
   return HYG_OK;
}

int hyg_mesh_set_shell_map( hyg_mesh mesh )
{
   HYG_TRY( checkInit() );
   if ( !mesh )
      return fail( HYG_ERR_ARG, "null mesh" );
   try
   {
      mesh->ctx.setShellMap();
   } catch ( const std::runtime_error& e )
   {
      return fail( HYG_ERR_ARG, e.what() );
   }
   return HYG_OK;
}

int hyg_op_create_p1_blending_laplace( hyg_mesh mesh, int min_level, int max_level, hyg_op* out )
{
   HYG_TRY( checkInit() );
   if ( !mesh || !out || min_level < 0 || max_level < min_level || max_level > 10 )
      return fail( HYG_ERR_ARG, "bad argument" );
   if ( mesh->ctx.shellHost.empty() )
      return fail( HYG_ERR_ARG, "the blended operator needs the shell map (hyg_mesh_set_shell_map)" );
   auto* op   = new hyg_op_s;
   op->mesh   = mesh;
   op->minL   = min_level;
   op->maxL   = max_level;
   op->kind   = 1;
   op->blend  = true;
   op->stored = true; // no coefficient integrals: the weights are stored
   try
   {
      for ( int L = min_level; L <= max_level; ++L )
         mesh->ctx.geom( L );
      assembleP1DivKGrad( *op );
      assembleBlendWeights( mesh->ctx, *op, nullptr );
      for ( int L = min_level; L <= max_level; ++L )
         launchElementwiseInverseDiagonal( mesh->ctx, mesh->ctx.geom( L ), *op, L, nullptr );
   } catch ( const std::exception& e )
   {
      delete op;
      return fail( HYG_ERR_HIP, e.what() );
   }
   if ( int rc = launched( "blended operator setup" ) )
   {
      delete op;
      return rc;
   }
   if ( hipDeviceSynchronize() != hipSuccess )
   {
      delete op;
      return fail( HYG_ERR_HIP, "blended operator setup failed" );
   }
   *out = op;
   return HYG_OK;
}

int hyg_op_create_p1_divkgrad_poly( hyg_mesh mesh, int min_level, int max_level, int degree, const double* coeffs, hyg_op* out )
{
   HYG_TRY( checkInit() );
   if ( !mesh || !out || min_level < 0 || max_level < min_level || max_level > 10 || degree < 0 || degree > 12 || !coeffs )
      return fail( HYG_ERR_ARG, "bad argument" );
   const int nc = ( degree + 1 ) * ( degree + 2 ) * ( degree + 3 ) / 6;
   for ( int i = 0; i < nc; ++i )
      if ( !std::isfinite( coeffs[i] ) )
         return fail( HYG_ERR_ARG, "non-finite polynomial coefficient" );
   auto* op     = new hyg_op_s;
   op->mesh     = mesh;
   op->minL     = min_level;
   op->maxL     = max_level;
   op->kind     = 1;
   op->kPolyDeg = degree;
   op->kPolyHost.assign( coeffs, coeffs + nc );
   try
   {
      for ( int L = min_level; L <= max_level; ++L )
         mesh->ctx.geom( L );
      assembleP1DivKGrad( *op );
      for ( int L = min_level; L <= max_level; ++L )
         launchElementwiseInverseDiagonal( mesh->ctx, mesh->ctx.geom( L ), *op, L, nullptr );
   } catch ( const std::exception& e )
   {
      delete op;
      return fail( HYG_ERR_HIP, e.what() );
   }
   if ( int rc = launched( "elementwise operator setup" ) )
   {
      delete op;
      return rc;
   }
   if ( hipDeviceSynchronize() != hipSuccess )
   {
      delete op;
      return fail( HYG_ERR_HIP, "elementwise operator setup failed" );
   }
   *out = op;
   return HYG_OK;
}

int hyg_op_interpolate_stencils( hyg_op op, int degree, int interpolation_level )
{
   HYG_TRY( checkInit() );
   if ( !op || op->kind != 1 || op->blend )
      return fail( HYG_ERR_ARG, "surrogate stencils are for the variable-coefficient P1 operator" );
   if ( degree < 0 || degree > 12 )
      return fail( HYG_ERR_ARG, "surrogate polynomial degree must be in 0..12" );
   if ( interpolation_level < 2 )
      return fail( HYG_ERR_ARG, "the interpolation level must have interior points (level >= 2)" );
   MeshCtx& mc = op->mesh->ctx;
   try
   {
      if ( !op->stored )
         launchElementwiseStore( mc, *op, nullptr );
      if ( int rc = launched( "coefficient integrals" ) )
         return rc;
      surrogateInterpolate( mc, *op, degree, interpolation_level );
      for ( int L = op->minL; L <= op->maxL; ++L )
         launchElementwiseInverseDiagonal( mc, mc.geom( L ), *op, L, nullptr );
   } catch ( const std::exception& e )
   {
      return fail( HYG_ERR_HIP, e.what() );
   }
   if ( int rc = launched( "surrogate setup" ) )
      return rc;
   if ( hipDeviceSynchronize() != hipSuccess )
      return fail( HYG_ERR_HIP, "surrogate setup failed" );
   return HYG_OK;
}

int hyg_op_create_p2_const( hyg_mesh mesh, int min_level, int max_level, hyg_op* out )
{
   HYG_TRY( checkInit() );
   if ( !mesh || !out || min_level < 0 || max_level < min_level || max_level > 9 )
      return fail( HYG_ERR_ARG, "bad argument" );
   auto* op = new hyg_op_s;
   op->mesh = mesh;
   op->lat  = 1;
   op->minL = min_level + 1;
   op->maxL = max_level + 1;
   op->kind = 2;
   try
   {
      mesh->ctx.p2Tables();
      for ( int L = op->minL; L <= op->maxL; ++L )
         mesh->ctx.geom( L );
      assembleP2Const( *op );
      for ( int L = op->minL; L <= op->maxL; ++L )
         launchP2InverseDiagonal( mesh->ctx, mesh->ctx.geom( L ), *op, L, nullptr );
   } catch ( const std::exception& e )
   {
      delete op;
      return fail( HYG_ERR_HIP, e.what() );
   }
   if ( int rc = launched( "P2 operator setup" ) )
   {
      delete op;
      return rc;
   }
   if ( hipDeviceSynchronize() != hipSuccess )
   {
      delete op;
      return fail( HYG_ERR_HIP, "P2 operator setup failed" );
   }
   *out = op;
   return HYG_OK;
}

int hyg_op_create_p2_divkgrad( hyg_mesh mesh, int min_level, int max_level, int kind, double k0, double alpha, double beta,
                               uint64_t seed, hyg_op* out )
{
   HYG_TRY( checkInit() );
   if ( !mesh || !out || min_level < 0 || max_level < min_level || max_level > 9 || ( kind != 0 && kind != 1 ) ||
        !std::isfinite( k0 ) || !std::isfinite( alpha ) || !std::isfinite( beta ) )
      return fail( HYG_ERR_ARG, "bad argument" );
   auto* op = new hyg_op_s;
   op->mesh     = mesh;
   op->lat      = 1;
   op->minL     = min_level + 1;
   op->maxL     = max_level + 1;
   op->kind     = 3;
   op->p2k.kind = kind;
   op->p2k.k0   = k0;
   op->p2k.alpha = alpha;
   op->p2k.beta = beta;
   op->p2k.seed = seed;
   try
   {
      mesh->ctx.p2Tables();
      for ( int L = op->minL; L <= op->maxL; ++L )
         mesh->ctx.geom( L );
      assembleP2DivKGrad( *op );
      for ( int L = op->minL; L <= op->maxL; ++L )
      {
         launchP2Coefficients( mesh->ctx, *op, L, nullptr );
         launchP2InverseDiagonal( mesh->ctx, mesh->ctx.geom( L ), *op, L, nullptr );
      }
   } catch ( const std::exception& e )
   {
      delete op;
      return fail( HYG_ERR_HIP, e.what() );
   }
   if ( int rc = launched( "P2 variable-coefficient operator setup" ) )
   {
      delete op;
      return rc;
   }
   if ( hipDeviceSynchronize() != hipSuccess )
   {
      delete op;
      return fail( HYG_ERR_HIP, "P2 variable-coefficient operator setup failed" );
   }
   *out = op;
   return HYG_OK;
}

int hyg_op_store_local_matrices( hyg_op op )
{
   HYG_TRY( checkInit() );
   if ( !op )
      return fail( HYG_ERR_ARG, "null operator" );
   if ( op->kind != 1 && op->kind != 3 )
      return fail( HYG_ERR_ARG, "stored local matrices are for the variable-coefficient operators" );
   try
   {
      if ( op->kind == 1 )
         launchElementwiseStore( op->mesh->ctx, *op, nullptr );
      else
      {
         for ( int L = op->minL; L <= op->maxL; ++L )
            launchP2StoreStencils( op->mesh->ctx, *op, L, nullptr );
         op->stored = true;
      }
   } catch ( const std::exception& e )
   {
      return fail( HYG_ERR_HIP, e.what() );
   }
   if ( int rc = launched( "stored P2 stencils" ) )
      return rc;
   if ( hipDeviceSynchronize() != hipSuccess )
      return fail( HYG_ERR_HIP, "stored P2 stencil setup failed" );
   return HYG_OK;
}

int hyg_op_destroy( hyg_op op )
{
   delete op;
   return HYG_OK;
}

int hyg_op_cell_stencil( hyg_op op, int level, int cell, double* out15 )
{
   level = lat( op, level );
   HYG_TRY( checkOpLevel( op, level ) );
   if ( cell < 0 || cell >= op->mesh->ctx.mesh->nC() || !out15 )
      return fail( HYG_ERR_ARG, "bad cell index" );
   if ( op->kind != 0 )
      return fail( HYG_ERR_ARG, "the operator has no constant stencil" );
   std::memcpy( out15, &op->levels[level].cellStHost[16 * cell], 15 * sizeof( double ) );
   return HYG_OK;
}

int hyg_op_tune( hyg_op op, int level, int dtype, int reps, int* chosen, void* stream )
{
   level = lat( op, level );
   HYG_TRY( checkOpLevel( op, level ) );
   if ( op->kind != 0 && op->kind != 2 )
      return fail( HYG_ERR_ARG, "configurations exist for the constant-coefficient operators only" );
   if ( dtype != HYG_F64 && dtype != HYG_F32 )
      return fail( HYG_ERR_ARG, "dtype must be HYG_F64 or HYG_F32" );
   if ( reps < 1 )
      return fail( HYG_ERR_ARG, "reps must be positive" );
   MeshCtx&    mc = op->mesh->ctx;
   LevelGeom&  g  = mc.geom( level );
   hipStream_t st = S( stream );
   hyg_func_s *a = nullptr, *b = nullptr, *r = nullptr, *t = nullptr;
   auto        release = [&]() {
      for ( hyg_func_s* f : { a, b, r, t } )
         destroyFunc( f );
   };
   int rc = createFunc( op->mesh, level, level, dtype, &a, op->lat );
   if ( rc == HYG_OK )
      rc = createFunc( op->mesh, level, level, dtype, &b, op->lat );
   if ( rc == HYG_OK )
      rc = createFunc( op->mesh, level, level, dtype, &r, op->lat );
   if ( rc == HYG_OK )
      rc = createFunc( op->mesh, level, level, dtype, &t, op->lat );
   if ( rc != HYG_OK )
   {
      release();
      return rc;
   }
   hipEvent_t e0 = nullptr, e1 = nullptr;
   if ( hipEventCreate( &e0 ) != hipSuccess || hipEventCreate( &e1 ) != hipSuccess )
   {
      release();
      return fail( HYG_ERR_HIP, "hipEventCreate failed" );
   }
   // inputs: 1 everywhere (the timing does not depend on the values)
   const double one = 1.0;
   for ( hyg_func_s* f : { a, b, r, t } )
      lincombImpl( f, 0, nullptr, nullptr, false, one, level, HYG_ALL, st );
   int status = HYG_OK;
   for ( int mode = 0; mode < 6 && status == HYG_OK; ++mode )
   {
      StencilCall c;
      c.mode = mode;
      c.flag = HYG_INNER;
      c.a    = 0.5;
      c.x    = a->data[level];
      c.in1  = b->data[level];
      c.out0 = r->data[level];
      c.out1 = t->data[level];
      float best = 0.0f;
      int   bestCfg = 0;
      const bool p2   = op->kind == 2;
      const int  nCfg = p2 ? kNumP2Cfgs : kNumCellCfgs;
      auto       run  = [&]() {
         if ( p2 )
            launchP2( mc, g, *op, level, dtype, c, st );
         else
            launchStencil( mc, g, op->levels[level], level, dtype, c, st );
      };
      for ( int cfg = 0; cfg < nCfg && status == HYG_OK; ++cfg )
      {
         if ( p2 )
            setP2Config( g, dtype, mode, cfg );
         else
            setCellConfig( g, dtype, mode, cfg );
         for ( int k = 0; k < 2; ++k )
            run();
         (void) hipEventRecord( e0, st );
         for ( int k = 0; k < reps; ++k )
            run();
         (void) hipEventRecord( e1, st );
         float ms = 0.0f;
         if ( hipEventSynchronize( e1 ) != hipSuccess || hipEventElapsedTime( &ms, e0, e1 ) != hipSuccess )
            status = fail( HYG_ERR_HIP, "timing the stencil configurations failed" );
         else if ( cfg == 0 || ms < best )
         {
            best    = ms;
            bestCfg = cfg;
         }
      }
      if ( p2 )
         setP2Config( g, dtype, mode, bestCfg );
      else
         setCellConfig( g, dtype, mode, bestCfg );
      if ( chosen )
         chosen[mode] = bestCfg;
   }
   (void) hipEventDestroy( e0 );
   (void) hipEventDestroy( e1 );
   release();
   if ( status != HYG_OK )
      return status;
   return launched( "tune" );
}

int hyg_op_set_config( hyg_op op, int level, int dtype, int mode, int config )
{
   level = lat( op, level );
   HYG_TRY( checkOpLevel( op, level ) );
   if ( op->kind != 0 && op->kind != 2 )
      return fail( HYG_ERR_ARG, "configurations exist for the constant-coefficient operators only" );
   if ( ( dtype != HYG_F64 && dtype != HYG_F32 ) || mode < 0 || mode > 5 || config < 0 ||
        config >= ( op->kind == 2 ? kNumP2Cfgs : kNumCellCfgs ) )
      return fail( HYG_ERR_ARG, "bad dtype, mode or configuration" );
   if ( op->kind == 2 )
      setP2Config( op->mesh->ctx.geom( level ), dtype, mode, config );
   else
      setCellConfig( op->mesh->ctx.geom( level ), dtype, mode, config );
   return HYG_OK;
}

int hyg_op_get_config( hyg_op op, int level, int dtype, int mode, int* config, int* rows, int* points )
{
   level = lat( op, level );
   HYG_TRY( checkOpLevel( op, level ) );
   if ( op->kind != 0 && op->kind != 2 )
      return fail( HYG_ERR_ARG, "configurations exist for the constant-coefficient operators only" );
   if ( ( dtype != HYG_F64 && dtype != HYG_F32 ) || mode < 0 || mode > 5 )
      return fail( HYG_ERR_ARG, "bad dtype or mode" );
   if ( op->kind == 2 )
   {
      const int c = p2Config( op->mesh->ctx.geom( level ), dtype, mode );
      if ( config )
         *config = c;
      if ( rows )
         *rows = kP2Chunks[c / 2];
      if ( points )
         *points = c % 2 + 1;
      return HYG_OK;
   }
   const int c = cellConfig( op->mesh->ctx.geom( level ), dtype, mode );
   if ( config )
      *config = c;
   if ( rows )
      *rows = kCellRows[c];
   if ( points )
      *points = kCellPoints[c];
   return HYG_OK;
}

int hyg_apply( hyg_op op, hyg_func src, hyg_func dst, int level, int flag, int update, void* stream )
{
   level = lat( op, level );
   HYG_TRY( checkOpLevel( op, level ) );
   HYG_TRY( checkPair( src, dst, level ) );
   HYG_TRY( checkSpace( op, src ) );
   if ( src == dst )
      return fail( HYG_ERR_ARG, "apply requires src != dst" );
   if ( src->mesh != op->mesh )
      return fail( HYG_ERR_ARG, "function and operator live on different meshes" );
   if ( update != HYG_REPLACE && update != HYG_ADD )
      return fail( HYG_ERR_ARG, "update must be HYG_REPLACE or HYG_ADD" );
   StencilCall c;
   c.mode = update == HYG_ADD ? MODE_APPLY_ADD : MODE_APPLY;
   c.flag = flag;
   c.x    = src->data[level];
   c.out0 = dst->data[level];
   return stencilImpl( op, level, src->dtype, c, S( stream ) );
}

int hyg_apply_vector( hyg_op op, int n, const hyg_func* srcs, const hyg_func* dsts, int level, int flag, int update, void* stream )
{
   if ( n < 1 || n > 3 || !srcs || !dsts )
      return fail( HYG_ERR_ARG, "hyg_apply_vector takes 1 to 3 components" );
   const int L = lat( op, level );
   HYG_TRY( checkOpLevel( op, L ) );
   if ( update != HYG_REPLACE && update != HYG_ADD )
      return fail( HYG_ERR_ARG, "update must be HYG_REPLACE or HYG_ADD" );
   StencilCall cs[3];
   for ( int k = 0; k < n; ++k )
   {
      HYG_TRY( checkPair( srcs[k], dsts[k], L ) );
      HYG_TRY( checkSpace( op, srcs[k] ) );
      if ( srcs[k]->dtype != srcs[0]->dtype )
         return fail( HYG_ERR_ARG, "components of different value types" );
      if ( srcs[k]->mesh != op->mesh )
         return fail( HYG_ERR_ARG, "function and operator live on different meshes" );
      for ( int j = 0; j < n; ++j )
         if ( srcs[k] == dsts[j] || ( j != k && dsts[k] == dsts[j] ) )
            return fail( HYG_ERR_ARG, "apply requires distinct sources and destinations" );
      cs[k].mode = update == HYG_ADD ? MODE_APPLY_ADD : MODE_APPLY;
      cs[k].flag = flag;
      cs[k].x    = srcs[k]->data[L];
      cs[k].out0 = dsts[k]->data[L];
   }
   MeshCtx& mc = op->mesh->ctx;
   if ( op->kind == 2 || op->kind == 3 )
   {
      launchP2Vector( mc, mc.geom( L ), *op, L, srcs[0]->dtype, cs, n, S( stream ) );
      return launched( "vector apply" );
   }
   for ( int k = 0; k < n; ++k )
      if ( int rc = stencilImpl( op, L, srcs[0]->dtype, cs[k], S( stream ) ) )
         return rc;
   return HYG_OK;
}

int hyg_smooth_jac( hyg_op op, hyg_func dst, hyg_func rhs, hyg_func src, double relax, int level, int flag, void* stream )
{
   level = lat( op, level );
   HYG_TRY( checkOpLevel( op, level ) );
   HYG_TRY( checkPair( src, dst, level ) );
   HYG_TRY( checkOpFunc( op, src ) );
   HYG_TRY( checkPair( src, rhs, level ) );
   if ( src == dst || rhs == dst )
      return fail( HYG_ERR_ARG, "smooth_jac requires dst different from src and rhs" );
   StencilCall c;
   c.mode = MODE_JACOBI;
   c.flag = flag;
   c.x    = src->data[level];
   c.in1  = rhs->data[level];
   c.out0 = dst->data[level];
   c.a    = relax;
   return stencilImpl( op, level, src->dtype, c, S( stream ) );
}

int hyg_smooth_sor( hyg_op op, hyg_func dst, hyg_func rhs, double relax, int level, int flag, int backwards, void* stream )
{
   level = lat( op, level );
   HYG_TRY( checkOpLevel( op, level ) );
   HYG_TRY( checkPair( dst, rhs, level ) );
   HYG_TRY( checkOpFunc( op, dst ) );
   if ( op->kind != 0 )
      return fail( HYG_ERR_ARG, "smooth_sor is implemented for the constant P1 operator" );
   if ( dst == rhs )
      return fail( HYG_ERR_ARG, "smooth_sor requires dst different from rhs" );
   MeshCtx& mc = op->mesh->ctx;
   launchSor( mc, mc.geom( level ), op->levels[level], level, dst->dtype, dst->data[level], rhs->data[level], relax, flag,
              backwards != 0, S( stream ) );
   return launched( "sor" );
}

int hyg_smooth_sor_ordered( hyg_op op, hyg_func dst, hyg_func rhs, double relax, int level, int flag, int backwards, void* stream )
{
   level = lat( op, level );
   HYG_TRY( checkOpLevel( op, level ) );
   HYG_TRY( checkPair( dst, rhs, level ) );
   HYG_TRY( checkOpFunc( op, dst ) );
   if ( op->kind != 0 )
      return fail( HYG_ERR_ARG, "smooth_sor is implemented for the constant P1 operator" );
   if ( dst == rhs )
      return fail( HYG_ERR_ARG, "smooth_sor requires dst different from rhs" );
   MeshCtx& mc = op->mesh->ctx;
   try
   {
      launchSorOrdered( mc, mc.geom( level ), op->levels[level], level, dst->dtype, dst->data[level], rhs->data[level], relax, flag,
                        backwards != 0, S( stream ) );
   } catch ( const std::exception& e )
   {
      return fail( HYG_ERR_HIP, e.what() );
   }
   return launched( "ordered sor" );
}

int hyg_mesh_sor_colours( hyg_mesh mesh, int* ncolours, int* vertex_colours )
{
   HYG_TRY( checkInit() );
   if ( !mesh || !ncolours )
      return fail( HYG_ERR_ARG, "null argument" );
   const MeshCtx& mc = mesh->ctx;
   *ncolours         = mc.nColours;
   if ( vertex_colours )
      for ( int v = 0; v < mc.mesh->nV(); ++v )
         vertex_colours[v] = mc.vertexColours[v];
   return HYG_OK;
}

int hyg_inverse_diagonal( hyg_op op, hyg_func dst, int level, void* stream )
{
   level = lat( op, level );
   HYG_TRY( checkOpLevel( op, level ) );
   HYG_TRY( checkLevel( dst, level ) );
   HYG_TRY( checkOpFunc( op, dst ) );
   MeshCtx& mc = op->mesh->ctx;
   if ( op->kind >= 1 )
   {
      // the elementwise and P2 operators keep their inverse diagonal as a function array
      const auto& ol = op->levels[level];
      const i64   n  = levelValues( dst, level );
      if ( dst->dtype == HYG_F64 )
         launchCopy( n, 8, dst->data[level], ol.dinv, S( stream ) );
      else
         launchCopy( n, 4, dst->data[level], ol.dinvF, S( stream ) );
   }
   else
      launchInvDiag( mc, mc.geom( level ), op->levels[level], level, dst->dtype, dst->data[level], S( stream ) );
   return launched( "inverse diagonal" );
}

int hyg_residual( hyg_op op, hyg_func r, hyg_func b, hyg_func x, int level, int flag, void* stream )
{
   level = lat( op, level );
   HYG_TRY( checkOpLevel( op, level ) );
   HYG_TRY( checkOpFunc( op, x ) );
   HYG_TRY( checkPair( r, b, level ) );
   HYG_TRY( checkPair( r, x, level ) );
   if ( r == x )
      return fail( HYG_ERR_ARG, "residual requires r != x" );
   StencilCall c;
   c.mode = MODE_RESIDUAL;
   c.flag = flag;
   c.x    = x->data[level];
   c.in1  = b->data[level];
   c.out0 = r->data[level];
   return stencilImpl( op, level, x->dtype, c, S( stream ) );
}

int hyg_restrict( hyg_func f, int source_level, int flag, void* stream )
{
   source_level = lat( f, source_level );
   HYG_TRY( checkLevel( f, source_level ) );
   HYG_TRY( checkLevel( f, source_level - 1 ) );
   MeshCtx& mc = f->mesh->ctx;
   if ( f->lat )
   {
      launchP2Restrict( mc, mc.geom( source_level - 1 ), mc.p2Tables(), source_level - 1, f->dtype, f->data[source_level - 1],
                        f->data[source_level], flag, S( stream ) );
      return launched( "restrict" );
   }
   launchRestrict( mc, mc.geom( source_level - 1 ), source_level - 1, f->dtype, f->data[source_level - 1], f->data[source_level],
                   flag, S( stream ) );
   return launched( "restrict" );
}

int hyg_prolongate( hyg_func f, int source_level, int flag, int update, void* stream )
{
   source_level = lat( f, source_level );
   HYG_TRY( checkLevel( f, source_level ) );
   HYG_TRY( checkLevel( f, source_level + 1 ) );
   MeshCtx& mc = f->mesh->ctx;
   if ( f->lat )
   {
      launchP2Prolongate( mc, mc.geom( source_level + 1 ), mc.p2Tables(), source_level + 1, f->dtype, f->data[source_level + 1],
                          f->data[source_level], flag, update == HYG_ADD, S( stream ) );
      return launched( "prolongate" );
   }
   launchProlongate( mc, mc.geom( source_level + 1 ), source_level + 1, f->dtype, f->data[source_level + 1], f->data[source_level],
                     flag, update == HYG_ADD, S( stream ) );
   return launched( "prolongate" );
}

int hyg_cheb_coefficients( int order, double lower, double upper, double* c )
{
   if ( !c || order < 1 || order > 5 )
      return fail( HYG_ERR_ARG, "Chebyshev order must be 1..5" );
   // coefficients of q with x <- x + q(D^-1 A) D^-1 r, 1 - t q(t) = T_k((theta - t)/delta) / T_k(theta/delta)
   const double th = 0.5 * ( upper + lower ), de = 0.5 * ( upper - lower );
   const double th2 = th * th, de2 = de * de;
   switch ( order )
   {
   case 1:
      c[0] = 1.0 / th;
      break;
   case 2: {
      const double q = 1.0 / ( de2 - 2.0 * th2 );
      c[0]           = -4.0 * th * q;
      c[1]           = 2.0 * q;
      break;
   }
   case 3: {
      const double q = 1.0 / ( th * ( 3.0 * de2 - 4.0 * th2 ) );
      c[0]           = q * ( 3.0 * de2 - 12.0 * th2 );
      c[1]           = 12.0 / ( 3.0 * de2 - 4.0 * th2 );
      c[2]           = -4.0 * q;
      break;
   }
   case 4: {
      const double q = 1.0 / ( de2 * de2 + 8.0 * th2 * th2 - 8.0 * de2 * th2 );
      c[0]           = q * ( 32.0 * th2 * th - 16.0 * th * de2 );
      c[1]           = q * ( 8.0 * de2 - 48.0 * th2 );
      c[2]           = 32.0 * th * q;
      c[3]           = -8.0 * q;
      break;
   }
   default: {
      const double d4 = de2 * de2, t4 = th2 * th2;
      const double q5 = 1.0 / ( th * ( 16.0 * t4 - 20.0 * de2 * th2 + 5.0 * d4 ) );
      const double q4 = 1.0 / ( 5.0 * d4 + 16.0 * t4 - 20.0 * de2 * th2 );
      c[0]            = q5 * ( 5.0 * d4 + 80.0 * t4 - 60.0 * de2 * th2 );
      c[1]            = q4 * ( 60.0 * de2 - 160.0 * th2 );
      c[2]            = q5 * ( 160.0 * th2 - 20.0 * de2 );
      c[3]            = -80.0 * q4;
      c[4]            = 16.0 * q5;
      break;
   }
   }
   return HYG_OK;
}

int hyg_power_iteration( hyg_op op, hyg_func x, hyg_func tmp, int iterations, int level, double* radius, void* stream )
{
   // the nested C ABI calls take the level of the operator's space, the internal routines the lattice level
   const int user = level;
   level          = lat( op, level );
   HYG_TRY( checkOpLevel( op, level ) );
   HYG_TRY( checkPair( x, tmp, level ) );
   HYG_TRY( checkOpFunc( op, x ) );
   if ( x == tmp || !radius || iterations < 1 )
      return fail( HYG_ERR_ARG, "bad argument" );
   hipStream_t st = S( stream );
   FuncGuard   dinv;
   HYG_TRY( createFunc( x->mesh, level, level, x->dtype, &dinv.f, x->lat ) );
   HYG_TRY( hyg_inverse_diagonal( op, dinv.f, user, stream ) );
   auto invDiagApply = [&]( hyg_func_s* s, hyg_func_s* d ) -> int {
      HYG_TRY( hyg_apply( op, s, d, user, HYG_ALL, HYG_REPLACE, stream ) );
      return hyg_mult_elementwise( d, dinv.f, d, user, HYG_ALL, stream );
   };
   double nrm = 0.0;
   HYG_TRY( dotImpl( x, x, level, HYG_ALL, &nrm, st ) );
   double s = 1.0 / std::sqrt( nrm );
   HYG_TRY( lincombImpl( x, 1, &s, &x, false, 0.0, level, HYG_ALL, st ) );
   HYG_TRY( invDiagApply( x, tmp ) );
   double r = 0.0;
   for ( int it = 1; it <= iterations; ++it )
   {
      HYG_TRY( dotImpl( tmp, tmp, level, HYG_ALL, &nrm, st ) );
      s = 1.0 / std::sqrt( nrm );
      HYG_TRY( lincombImpl( x, 1, &s, &tmp, false, 0.0, level, HYG_ALL, st ) );
      HYG_TRY( invDiagApply( x, tmp ) );
      HYG_TRY( dotImpl( x, tmp, level, HYG_ALL, &r, st ) );
   }
   *radius = r;
   if ( !std::isfinite( r ) )
      return fail( HYG_ERR_NUMERIC, "power iteration produced a non-finite value" );
   return HYG_OK;
}

int hyg_gmg_create( hyg_op op, int dtype, const int* ip, const double* dp, hyg_gmg* out )
{
   HYG_TRY( checkInit() );
   if ( !op || !ip || !dp || !out )
      return fail( HYG_ERR_ARG, "null argument" );
   if ( dtype != HYG_F64 && dtype != HYG_F32 )
      return fail( HYG_ERR_ARG, "dtype must be HYG_F64 or HYG_F32" );
   if ( lat( op, ip[0] ) < op->minL || lat( op, ip[1] ) > op->maxL || ip[0] > ip[1] )
      return fail( HYG_ERR_ARG, "solver levels outside the operator's range" );
   if ( ip[5] < 0 || ip[5] > 5 )
      return fail( HYG_ERR_ARG, "smoother must be 0 (weighted Jacobi), 1 (Chebyshev), 2 (SOR), 3 (symmetric SOR), 4 (SOR in the "
                                "reference's order) or 5 (symmetric SOR in the reference's order)" );
   if ( ip[5] >= 2 && op->kind != 0 )
      return fail( HYG_ERR_ARG, "the SOR smoother is implemented for the constant P1 operator" );
   if ( ip[5] == 1 && ( ip[6] < 1 || ip[6] > 5 ) )
      return fail( HYG_ERR_ARG, "Chebyshev order must be 1..5" );
   auto* g      = new hyg_gmg_s;
   g->op        = op;
   g->dtype     = dtype;
   g->minL      = lat( op, ip[0] );
   g->maxL      = lat( op, ip[1] );
   g->pre       = ip[2];
   g->post      = ip[3];
   g->inc       = ip[4];
   g->smoother  = ip[5];
   g->chebOrder = ip[6];
   g->cgMaxIter = ip[7];
   g->relax     = dp[0];
   g->chebLo    = dp[1];
   g->chebHi    = dp[2];
   g->cgTol     = dp[3];
   if ( g->smoother == 1 )
      hyg_cheb_coefficients( g->chebOrder, g->chebLo, g->chebHi, g->cheb );
   if ( g->smoother >= 4 )
   {
      // tables and snapshots now, so that neither a cycle nor its capture allocates
      MeshCtx& mc = op->mesh->ctx;
      try
      {
         for ( int L = g->minL; L <= g->maxL; ++L )
            prepareSorOrdered( mc, mc.geom( L ), op->levels[L], L, dtype );
      } catch ( const std::exception& e )
      {
         delete g;
         return fail( HYG_ERR_HIP, e.what() );
      }
   }
   hyg_mesh_s* m  = op->mesh;
   int         rc = HYG_OK;
   for ( hyg_func_s** f : { &g->tmp, &g->jt, &g->t1, &g->t2, &g->cy, &g->cgp, &g->cgz, &g->cgap, &g->cgr } )
      if ( rc == HYG_OK )
         rc = createFunc( m, g->minL, g->maxL, dtype, f, op->lat );
   if ( rc != HYG_OK )
   {
      delete g;
      return rc;
   }
   *out = g;
   return HYG_OK;
}

int hyg_gmg_destroy( hyg_gmg g )
{
   delete g;
   return HYG_OK;
}

int hyg_gmg_vcycle( hyg_gmg g, hyg_func x, hyg_func b, int level, void* stream )
{
   level = lat( g ? g->op : nullptr, level );
   HYG_TRY( checkGmgArgs( g, x, b, level ) );
   HYG_TRY( vcycleImpl( g, x, b, level, level, S( stream ) ) );
   return HYG_OK;
}

int hyg_gmg_vcycle_graph( hyg_gmg g, hyg_func x, hyg_func b, int level, int cg_iterations, void* stream )
{
   if ( !g )
      return fail( HYG_ERR_ARG, "null solver handle" );
   level = lat( g->op, level );
   HYG_TRY( checkPair( x, b, level ) );
   HYG_TRY( checkOpFunc( g->op, x ) );
   if ( x == b )
      return fail( HYG_ERR_ARG, "x and b must differ" );
   if ( !stream )
      return fail( HYG_ERR_ARG, "capturing a V-cycle needs a stream other than the default stream" );
   if ( level < g->minL || level > g->maxL || x->dtype != g->dtype || b->dtype != g->dtype || x->minL > g->minL ||
        b->minL > g->minL )
      return fail( HYG_ERR_ARG, "functions must hold all solver levels in the solver's value type" );
   if ( cg_iterations < 1 )
      return fail( HYG_ERR_ARG, "cg_iterations must be positive" );
   hipStream_t st = S( stream );
   // the cache key uses the creation numbers of x and b, never their addresses: a destroyed function's handle
   // address can be reused by a new function whose device arrays differ
   if ( !g->graphExec || g->graphX != x->gen || g->graphB != b->gen || g->graphLevel != level || g->cgGraphIters != cg_iterations )
   {
      if ( g->graphExec )
         (void) hipGraphExecDestroy( g->graphExec );
      if ( g->graph )
         (void) hipGraphDestroy( g->graph );
      g->graphExec = nullptr;
      g->graph     = nullptr;
      MeshCtx& mc  = g->op->mesh->ctx;
      if ( !g->cgDev )
         HYG_HIP( hipMalloc( &g->cgDev, 4 * sizeof( double ) ) );
      if ( !g->cgPartials )
         HYG_HIP( hipMalloc( &g->cgPartials, size_t( dotPartialsNeeded( mc, mc.geom( g->minL ) ) ) * sizeof( double ) ) );
      for ( int L = g->minL; L <= level; ++L )
         mc.geom( L );
      g->cgGraphIters = cg_iterations;
      HYG_HIP( hipStreamBeginCapture( st, hipStreamCaptureModeThreadLocal ) );
      const int  rc = vcycleImpl( g, x, b, level, level, st, true );
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture( st, &graph );
      if ( rc != HYG_OK || e != hipSuccess )
      {
         if ( graph )
            (void) hipGraphDestroy( graph );
         return rc != HYG_OK ? rc : fail( HYG_ERR_HIP, std::string( "capturing the V-cycle: " ) + hipGetErrorString( e ) );
      }
      hipGraphExec_t exec = nullptr;
      if ( hipGraphInstantiate( &exec, graph, nullptr, nullptr, 0 ) != hipSuccess )
      {
         (void) hipGraphDestroy( graph );
         return fail( HYG_ERR_HIP, "instantiating the V-cycle graph failed" );
      }
      g->graph      = graph;
      g->graphExec  = exec;
      g->graphX     = x->gen;
      g->graphB     = b->gen;
      g->graphLevel = level;
   }
   HYG_HIP( hipGraphLaunch( g->graphExec, st ) );
   return HYG_OK;
}

int hyg_gmg_solve( hyg_gmg g, hyg_func x, hyg_func b, int level, int cycles, double* res, void* stream )
{
   level = lat( g ? g->op : nullptr, level );
   HYG_TRY( checkGmgArgs( g, x, b, level ) );
   if ( cycles < 0 || !res )
      return fail( HYG_ERR_ARG, "bad argument" );
   hipStream_t st = S( stream );
   HYG_TRY( residualNorm( g, x, b, g->tmp, level, &res[0], st ) );
   for ( int k = 0; k < cycles; ++k )
   {
      HYG_TRY( vcycleImpl( g, x, b, level, level, st ) );
      HYG_TRY( residualNorm( g, x, b, g->tmp, level, &res[k + 1], st ) );
   }
   return HYG_OK;
}

int hyg_gmg_solve_tol( hyg_gmg g, hyg_func x, hyg_func b, int level, int max_cycles, double rel_tol, double* res, int* cycles,
                       void* stream )
{
   level = lat( g ? g->op : nullptr, level );
   HYG_TRY( checkGmgArgs( g, x, b, level ) );
   if ( max_cycles < 0 || !res || !cycles )
      return fail( HYG_ERR_ARG, "bad argument" );
   hipStream_t st = S( stream );
   HYG_TRY( residualNorm( g, x, b, g->tmp, level, &res[0], st ) );
   *cycles = 0;
   for ( int k = 0; k < max_cycles && res[k] > rel_tol * res[0]; ++k )
   {
      HYG_TRY( vcycleImpl( g, x, b, level, level, st ) );
      HYG_TRY( residualNorm( g, x, b, g->tmp, level, &res[k + 1], st ) );
      *cycles = k + 1;
   }
   return HYG_OK;
}

int hyg_ir_solve( hyg_op op64, hyg_gmg g32, hyg_func x, hyg_func b, int level, int inner, int max_outer, double rel_tol,
                  double* res, int* steps, void* stream )
{
   level = lat( op64, level );
   HYG_TRY( checkOpLevel( op64, level ) );
   if ( !g32 || g32->dtype != HYG_F32 || g32->maxL != level )
      return fail( HYG_ERR_ARG, "defect correction needs an FP32 solver whose max level is the solve level" );
   HYG_TRY( checkPair( x, b, level ) );
   HYG_TRY( checkOpFunc( op64, x ) );
   if ( g32->op->lat != op64->lat )
      return fail( HYG_ERR_ARG, "P1 and P2 operators and solvers cannot be combined" );
   if ( g32->op->mesh != op64->mesh )
      return fail( HYG_ERR_ARG, "the FP64 operator and the FP32 solver live on different meshes" );
   if ( x->dtype != HYG_F64 )
      return fail( HYG_ERR_ARG, "defect correction works on FP64 x and b" );
   if ( !res || !steps || inner < 1 || max_outer < 0 )
      return fail( HYG_ERR_ARG, "bad argument" );
   hipStream_t st = S( stream );
   hyg_mesh_s* m  = op64->mesh;
   if ( !g32->x32 )
   {
      HYG_TRY( createFunc( m, g32->minL, g32->maxL, HYG_F32, &g32->x32, op64->lat ) );
      HYG_TRY( createFunc( m, g32->minL, g32->maxL, HYG_F32, &g32->b32, op64->lat ) );
      HYG_TRY( createFunc( m, level, level, HYG_F64, &g32->r64, op64->lat ) );
      HYG_TRY( createFunc( m, level, level, HYG_F64, &g32->t64, op64->lat ) );
   }
   // residual in FP64 (the FP64 operator shares the FP32 one's mesh)
   hyg_gmg_s dummy{};
   dummy.op    = op64;
   dummy.dtype = HYG_F64;
   double r0   = 0.0;
   HYG_TRY( residualNorm( &dummy, x, b, g32->r64, level, &r0, st ) );
   res[0]   = r0;
   *steps   = 0;
   for ( int it = 0; it < max_outer; ++it )
   {
      if ( r0 == 0.0 || res[it] / r0 <= rel_tol )
         break;
      launchConvert( levelValues( x, level ), HYG_F32, g32->b32->data[level], HYG_F64, g32->r64->data[level], st );
      HYG_TRY( launched( "convert" ) );
      HYG_TRY( lincombImpl( g32->x32, 0, nullptr, nullptr, false, 0.0, level, HYG_ALL, st ) );
      for ( int k = 0; k < inner; ++k )
         HYG_TRY( vcycleImpl( g32, g32->x32, g32->b32, level, level, st ) );
      launchConvert( levelValues( x, level ), HYG_F64, g32->t64->data[level], HYG_F32, g32->x32->data[level], st );
      HYG_TRY( launched( "convert" ) );
      const double one = 1.0;
      hyg_func_s*  t   = g32->t64;
      HYG_TRY( lincombImpl( x, 1, &one, &t, true, 0.0, level, kGmgFlag, st ) );
      HYG_TRY( residualNorm( &dummy, x, b, g32->r64, level, &res[it + 1], st ) );
      *steps = it + 1;
   }
   return HYG_OK;
}

} // extern "C"

// ================================================================== batched independent problems
// A batch runs the V-cycles (and the defect-correction steps around them) of many independent problems
// together: every grid operation is one persistent launch over all problems (BatchPlan). Problem i has its own
// solver (operator, levels, smoother scalars), mesh and functions; the cycle parameters (value type, smoother
// kind and order, pre/post/increment) are those of the first solver and must be shared by all.
struct hyg_batch_s
{
   std::vector< hyg_gmg_s* >  g;
   std::vector< hyg_func_s* > x, b;
   int                        dtype = 8;
   int                        cfg   = -1; // forced cell-interior configuration of the sweeps, -1: per step (tuned)
   int                        depth = 0; // largest number of levels below the top
   std::unique_ptr< BatchPlan > cycle;  // one V-cycle of every problem
   std::unique_ptr< BatchPlan > norms;  // residual b - A x at the top level into the solvers' tmp, its dot
   hipGraphExec_t             cycleExec = nullptr;
   double*                    partials  = nullptr; // dot partial sums, per problem at offset partOff[i]
   double*                    results   = nullptr; // one per problem
   double*                    host      = nullptr; // pinned copy of results
   // defect correction of the FP64 problems (x64, b64) around this FP32 batch: keyed by the creation numbers
   std::vector< uint64_t >    irKey;
   std::vector< hyg_func_s* > r64, t64;
   std::unique_ptr< BatchPlan > irPre, irPost; // residual + convert + zero; correction + residual + dot
   hipGraphExec_t             irExec  = nullptr;
   int                        irInner = 0;
   size_t                     irNormStart = 0; // first step of irPost's residual norm
   // batched apply plans keyed by the creation numbers of (src, dst) of every problem and the flag
   std::vector< uint64_t >    applyKey;
   std::unique_ptr< BatchPlan > applyPlan;
   ~hyg_batch_s()
   {
      if ( cycleExec )
         (void) hipGraphExecDestroy( cycleExec );
      if ( irExec )
         (void) hipGraphExecDestroy( irExec );
      for ( hyg_func_s* f : r64 )
         destroyFunc( f );
      for ( hyg_func_s* f : t64 )
         destroyFunc( f );
      (void) hipFree( partials );
      (void) hipFree( results );
      if ( host )
         (void) hipHostFree( host );
   }
};

namespace {

int levelOf( const hyg_batch_s* bt, size_t i, int d ) { return bt->g[i]->maxL - d; }

// cell-interior configuration of a batched sweep: the batch's forced one, else the one chosen for the largest
// problem of the step (hyg_op_tune, HYG_CELL_CFG or the default) at its level, value type and mode
int stepConfig( const hyg_batch_s* bt, const std::vector< BatchPlan::Item >& items, int dtype, int mode )
{
   if ( bt->cfg >= 0 )
      return bt->cfg;
   const BatchPlan::Item* big = nullptr;
   i64                    most = -1;
   for ( const auto& it : items )
   {
      const i64 n = i64( it.m->mesh->nC() ) * it.g->stride;
      if ( it.op->kind == 0 && n > most )
      {
         most = n;
         big  = &it;
      }
   }
   return big ? cellConfig( *big->g, dtype, mode ) : 1;
}

void addStencil( hyg_batch_s* bt, BatchPlan& plan, int dtype, int mode, const std::vector< BatchPlan::Item >& items )
{
   plan.stencil( dtype, mode, items, stepConfig( bt, items, dtype, mode ) );
}

// inner DoFs of level L (cell interiors and points of primitives flagged Inner)
i64 innerDofs( const MeshCtx& mc, int L )
{
   i64 n = 0;
   forEachCanonical( mc, L, [&]( i64, int p, int, int, int ) {
      if ( p < 0 || ( mc.primsHost[p].bc & HYG_INNER ) )
         ++n;
   } );
   return n;
}

BatchPlan::Item stencilItem( hyg_gmg_s* g, int L, int mode, const void* x, const void* in1, void* out0, void* out1, double a )
{
   MeshCtx&        mc = g->op->mesh->ctx;
   BatchPlan::Item it{ &mc, &mc.geom( L ), g->op, L, StencilCall{} };
   it.c.mode = mode;
   it.c.flag = kGmgFlag;
   it.c.x    = x;
   it.c.in1  = in1;
   it.c.out0 = out0;
   it.c.out1 = out1;
   it.c.a    = a;
   return it;
}

// one smoothing step (WeightedJacobiSmoother or ChebyshevSmoother::solve, as smoothImpl) of the problems in `act` at depth d
void batchSmooth( hyg_batch_s* bt, BatchPlan& plan, const std::vector< size_t >& act, int d )
{
   const int dt = bt->dtype;
   if ( bt->g[0]->smoother == 0 )
   {
      std::vector< BatchPlan::Vec >  cp;
      std::vector< BatchPlan::Item > it;
      for ( size_t i : act )
      {
         hyg_gmg_s* g = bt->g[i];
         const int  L = levelOf( bt, i, d );
         MeshCtx&   mc = g->op->mesh->ctx;
         BatchPlan::Vec v{ &mc, &mc.geom( L ), L, HYG_ALL, 1 };
         v.dst    = g->jt->data[L];
         v.src[0] = bt->x[i]->data[L];
         v.a[0]   = 1.0;
         cp.push_back( v );
         it.push_back( stencilItem( g, L, MODE_JACOBI, g->jt->data[L], bt->b[i]->data[L], bt->x[i]->data[L], nullptr, g->relax ) );
      }
      plan.vec( dt, cp );
      addStencil( bt, plan, dt, MODE_JACOBI, it );
      return;
   }
   const int K = bt->g[0]->chebOrder;
   {
      std::vector< BatchPlan::Item > it;
      for ( size_t i : act )
      {
         hyg_gmg_s* g = bt->g[i];
         const int  L = levelOf( bt, i, d );
         it.push_back( stencilItem( g, L, MODE_CHEB0, bt->x[i]->data[L], bt->b[i]->data[L], g->t1->data[L], g->cy->data[L], g->cheb[0] ) );
      }
      addStencil( bt, plan, dt, MODE_CHEB0, it );
   }
   bool swapped = false; // t1 -> t2 -> t1 ... as smoothImpl's tin / tout
   for ( int k = 1; k < K; ++k )
   {
      std::vector< BatchPlan::Item > it;
      for ( size_t i : act )
      {
         hyg_gmg_s*  g    = bt->g[i];
         const int   L    = levelOf( bt, i, d );
         hyg_func_s* tin  = swapped ? g->t2 : g->t1;
         hyg_func_s* tout = swapped ? g->t1 : g->t2;
         it.push_back( stencilItem( g, L, MODE_CHEBK, tin->data[L], g->cy->data[L], tout->data[L],
                                    k == K - 1 ? bt->x[i]->data[L] : g->cy->data[L], g->cheb[k] ) );
      }
      addStencil( bt, plan, dt, MODE_CHEBK, it );
      swapped = !swapped;
   }
   if ( K == 1 )
   {
      std::vector< BatchPlan::Vec > cp;
      for ( size_t i : act )
      {
         hyg_gmg_s* g  = bt->g[i];
         const int  L  = levelOf( bt, i, d );
         MeshCtx&   mc = g->op->mesh->ctx;
         BatchPlan::Vec v{ &mc, &mc.geom( L ), L, kGmgFlag, 1 };
         v.dst    = bt->x[i]->data[L];
         v.src[0] = g->cy->data[L];
         v.a[0]   = 1.0;
         cp.push_back( v );
      }
      plan.vec( dt, cp );
   }
}

// GeometricMultigridSolver::solveRecursively (as vcycleImpl) for every problem: depth d acts on level maxL_i - d of the
// problems that still have a coarser level; a problem's coarse level has no inner DoFs (checked at creation), so its
// coarse solve changes nothing and is left out
void buildCycle( hyg_batch_s* bt, BatchPlan& plan )
{
   const int dt   = bt->dtype;
   auto      act  = [&]( int d ) {
      std::vector< size_t > a;
      for ( size_t i = 0; i < bt->g.size(); ++i )
         if ( bt->g[i]->maxL - d > bt->g[i]->minL )
            a.push_back( i );
      return a;
   };
   const hyg_gmg_s* g0 = bt->g[0];
   for ( int d = 0; d < bt->depth; ++d )
   {
      const auto A = act( d );
      for ( int s = 0; s < g0->pre + g0->inc * d; ++s )
         batchSmooth( bt, plan, A, d );
      std::vector< BatchPlan::Item >     res;
      std::vector< BatchPlan::Transfer > rst;
      std::vector< BatchPlan::Vec >      vv;
      for ( size_t i : A )
      {
         hyg_gmg_s* g  = bt->g[i];
         const int  L  = levelOf( bt, i, d );
         MeshCtx&   mc = g->op->mesh->ctx;
         res.push_back( stencilItem( g, L, MODE_RESIDUAL, bt->x[i]->data[L], bt->b[i]->data[L], g->tmp->data[L], nullptr, 0.0 ) );
         rst.push_back( BatchPlan::Transfer{ &mc, &mc.geom( L - 1 ), L - 1, kGmgFlag, false, g->tmp->data[L - 1], g->tmp->data[L] } );
         BatchPlan::Vec cb{ &mc, &mc.geom( L - 1 ), L - 1, kGmgFlag, 1 };
         cb.dst    = bt->b[i]->data[L - 1];
         cb.src[0] = g->tmp->data[L - 1];
         cb.a[0]   = 1.0;
         vv.push_back( cb );
         BatchPlan::Vec zx{ &mc, &mc.geom( L - 1 ), L - 1, HYG_ALL, 0 };
         zx.dst = bt->x[i]->data[L - 1];
         vv.push_back( zx );
      }
      addStencil( bt, plan, dt, MODE_RESIDUAL, res );
      plan.restrictLevels( dt, rst );
      plan.vec( dt, vv );
   }
   for ( int d = bt->depth - 1; d >= 0; --d )
   {
      const auto                         A = act( d );
      std::vector< BatchPlan::Transfer > pro;
      for ( size_t i : A )
      {
         hyg_gmg_s* g  = bt->g[i];
         const int  L  = levelOf( bt, i, d );
         MeshCtx&   mc = g->op->mesh->ctx;
         pro.push_back( BatchPlan::Transfer{ &mc, &mc.geom( L ), L, kGmgFlag, true, bt->x[i]->data[L], bt->x[i]->data[L - 1] } );
      }
      plan.prolongate( dt, pro );
      for ( int s = 0; s < g0->post + g0->inc * d; ++s )
         batchSmooth( bt, plan, A, d );
   }
}

// r = b - A x at the top level of every problem (into r[i]), then the per-problem dot r . r into bt->results
void buildNorms( hyg_batch_s* bt, BatchPlan& plan, int dt, const std::vector< hyg_func_s* >& xs, const std::vector< hyg_func_s* >& bs,
                 const std::vector< hyg_func_s* >& rs )
{
   std::vector< BatchPlan::Item >     res;
   std::vector< BatchPlan::Vec >      dots;
   std::vector< BatchPlan::DotFinal > fin;
   double*                            part = bt->partials;
   for ( size_t i = 0; i < bt->g.size(); ++i )
   {
      hyg_gmg_s* g  = bt->g[i];
      const int  L  = g->maxL;
      MeshCtx&   mc = g->op->mesh->ctx;
      res.push_back( stencilItem( g, L, MODE_RESIDUAL, xs[i]->data[L], bs[i]->data[L], rs[i]->data[L], nullptr, 0.0 ) );
      BatchPlan::Vec v{ &mc, &mc.geom( L ), L, kGmgFlag, 0 };
      v.dot      = true;
      v.src[0]   = rs[i]->data[L];
      v.src[1]   = rs[i]->data[L];
      v.partials = part;
      dots.push_back( v );
      const int np = dotPartialsNeeded( mc, mc.geom( L ) );
      fin.push_back( BatchPlan::DotFinal{ part, np, bt->results + i } );
      part += np;
   }
   addStencil( bt, plan, dt, MODE_RESIDUAL, res );
   plan.vec( dt, dots );
   plan.dotFinal( fin );
}

int batchLaunchGraph( const BatchPlan& plan, hipGraphExec_t* exec, size_t from, size_t to, hipStream_t st )
{
   if ( !*exec )
   {
      hipGraph_t graph = nullptr;
      HYG_HIP( hipStreamBeginCapture( st, hipStreamCaptureModeThreadLocal ) );
      plan.launch( st, from, to );
      const hipError_t e = hipStreamEndCapture( st, &graph );
      if ( e != hipSuccess )
         return fail( HYG_ERR_HIP, std::string( "capturing the batched cycle: " ) + hipGetErrorString( e ) );
      const hipError_t e2 = hipGraphInstantiate( exec, graph, nullptr, nullptr, 0 );
      (void) hipGraphDestroy( graph );
      if ( e2 != hipSuccess )
      {
         *exec = nullptr;
         return fail( HYG_ERR_HIP, std::string( "instantiating the batched cycle: " ) + hipGetErrorString( e2 ) );
      }
   }
   HYG_HIP( hipGraphLaunch( *exec, st ) );
   return HYG_OK;
}

int batchPlanCycle( hyg_batch_s* bt )
{
   bt->cycle.reset( new BatchPlan );
   buildCycle( bt, *bt->cycle );
   HYG_HIP( bt->cycle->upload() );
   if ( bt->cycleExec )
   {
      (void) hipGraphExecDestroy( bt->cycleExec );
      bt->cycleExec = nullptr;
   }
   std::vector< hyg_func_s* > tmps;
   for ( hyg_gmg_s* g : bt->g )
      tmps.push_back( g->tmp );
   bt->norms.reset( new BatchPlan );
   buildNorms( bt, *bt->norms, bt->dtype, bt->x, bt->b, tmps );
   HYG_HIP( bt->norms->upload() );
   bt->irKey.clear(); // rebuilt with the new configuration on the next defect-correction call
   bt->applyKey.clear();
   return HYG_OK;
}

int batchReadNorms( hyg_batch_s* bt, double* out, hipStream_t st )
{
   const size_t n = bt->g.size();
   HYG_HIP( hipMemcpyAsync( bt->host, bt->results, n * sizeof( double ), hipMemcpyDeviceToHost, st ) );
   HYG_HIP( hipStreamSynchronize( st ) );
   for ( size_t i = 0; i < n; ++i )
   {
      out[i] = std::sqrt( bt->host[i] );
      if ( !std::isfinite( out[i] ) )
         return fail( HYG_ERR_NUMERIC, "non-finite residual norm in problem " + std::to_string( i ) );
   }
   return HYG_OK;
}

} // namespace

extern "C" {

int hyg_batch_create( int n, const hyg_gmg* solvers, const hyg_func* x, const hyg_func* b, hyg_batch* out )
{
   HYG_TRY( checkInit() );
   if ( n < 1 || !solvers || !x || !b || !out )
      return fail( HYG_ERR_ARG, "bad argument" );
   const hyg_gmg_s* g0 = solvers[0];
   if ( !g0 )
      return fail( HYG_ERR_ARG, "null solver handle" );
   for ( int i = 0; i < n; ++i )
   {
      hyg_gmg_s* g = solvers[i];
      const std::string who = "problem " + std::to_string( i ) + ": ";
      HYG_TRY( checkGmgArgs( g, x[i], b[i], g ? g->maxL : 0 ) );
      if ( g->dtype != g0->dtype || g->smoother != g0->smoother || g->pre != g0->pre || g->post != g0->post || g->inc != g0->inc ||
           ( g->smoother == 1 && g->chebOrder != g0->chebOrder ) )
         return fail( HYG_ERR_ARG, who + "value type, smoother and cycle shape must be those of problem 0" );
      if ( g->smoother > 1 )
         return fail( HYG_ERR_ARG, who + "batched cycles take the weighted Jacobi or the Chebyshev smoother" );
      if ( g->op->lat != 0 || g->op->kind > 1 )
         return fail( HYG_ERR_ARG, who + "batched cycles take P1 constant and P1 elementwise operators" );
      if ( g->op->kind == 1 && !g->op->stored )
         return fail( HYG_ERR_ARG, who + "an elementwise operator in a batch must keep its local matrices (hyg_op_store_local_matrices)" );
      if ( g->op->surDegree >= 0 || g->op->blend )
         return fail( HYG_ERR_ARG, who + "surrogate and blended operators are not batched" );
      if ( innerDofs( g->op->mesh->ctx, g->minL ) != 0 )
         return fail( HYG_ERR_ARG, who + "the coarse level has inner DoFs (batched cycles have no coarse solve)" );
   }
   auto* bt  = new hyg_batch_s;
   bt->dtype = g0->dtype;
   for ( int i = 0; i < n; ++i )
   {
      bt->g.push_back( solvers[i] );
      bt->x.push_back( x[i] );
      bt->b.push_back( b[i] );
      bt->depth = std::max( bt->depth, solvers[i]->maxL - solvers[i]->minL );
   }
   size_t np = 0;
   for ( hyg_gmg_s* g : bt->g )
   {
      MeshCtx& mc = g->op->mesh->ctx;
      np += size_t( dotPartialsNeeded( mc, mc.geom( g->maxL ) ) );
   }
   int rc = HYG_OK;
   if ( hipMalloc( &bt->partials, np * sizeof( double ) ) != hipSuccess || hipMalloc( &bt->results, n * sizeof( double ) ) != hipSuccess ||
        hipHostMalloc( (void**) &bt->host, n * sizeof( double ), hipHostMallocDefault ) != hipSuccess )
      rc = fail( HYG_ERR_HIP, "allocating the batch's reduction buffers" );
   if ( rc == HYG_OK )
      rc = batchPlanCycle( bt );
   if ( rc != HYG_OK )
   {
      delete bt;
      return rc;
   }
   *out = bt;
   return HYG_OK;
}

int hyg_batch_destroy( hyg_batch bt )
{
   delete bt;
   return HYG_OK;
}

int hyg_batch_set_config( hyg_batch bt, int config )
{
   if ( !bt || config < -1 || config >= kNumCellCfgs )
      return fail( HYG_ERR_ARG, "batch configuration must be -1 (per level, as tuned) or a cell-kernel configuration 0.." +
                                    std::to_string( kNumCellCfgs - 1 ) );
   if ( config == bt->cfg )
      return HYG_OK;
   bt->cfg = config;
   return batchPlanCycle( bt );
}

int hyg_batch_vcycle( hyg_batch bt, int use_graph, void* stream )
{
   if ( !bt )
      return fail( HYG_ERR_ARG, "null batch handle" );
   hipStream_t st = S( stream );
   if ( use_graph )
      return batchLaunchGraph( *bt->cycle, &bt->cycleExec, 0, size_t( -1 ), st );
   bt->cycle->launch( st );
   return launched( "batched V-cycle" );
}

int hyg_batch_residual_norms( hyg_batch bt, double* out, void* stream )
{
   if ( !bt || !out )
      return fail( HYG_ERR_ARG, "bad argument" );
   bt->norms->launch( S( stream ) );
   HYG_TRY( launched( "batched residual" ) );
   return batchReadNorms( bt, out, S( stream ) );
}

int hyg_batch_apply( hyg_batch bt, const hyg_func* src, const hyg_func* dst, int flag, void* stream )
{
   if ( !bt || !src || !dst )
      return fail( HYG_ERR_ARG, "bad argument" );
   const size_t            n = bt->g.size();
   std::vector< uint64_t > key{ uint64_t( flag ) };
   for ( size_t i = 0; i < n; ++i )
   {
      const int L = bt->g[i]->maxL;
      HYG_TRY( checkPair( src[i], dst[i], L ) );
      HYG_TRY( checkOpFunc( bt->g[i]->op, src[i] ) );
      if ( src[i]->dtype != bt->dtype )
         return fail( HYG_ERR_ARG, "functions must have the batch's value type" );
      key.push_back( src[i]->gen );
      key.push_back( dst[i]->gen );
   }
   if ( key != bt->applyKey )
   {
      std::vector< BatchPlan::Item > it;
      for ( size_t i = 0; i < n; ++i )
      {
         BatchPlan::Item item = stencilItem( bt->g[i], bt->g[i]->maxL, MODE_APPLY, src[i]->data[bt->g[i]->maxL], nullptr,
                                             dst[i]->data[bt->g[i]->maxL], nullptr, 0.0 );
         item.c.flag          = flag;
         it.push_back( item );
      }
      bt->applyPlan.reset( new BatchPlan );
      addStencil( bt, *bt->applyPlan, bt->dtype, MODE_APPLY, it );
      HYG_HIP( bt->applyPlan->upload() );
      bt->applyKey = key;
   }
   bt->applyPlan->launch( S( stream ) );
   return launched( "batched apply" );
}

int hyg_batch_ir_solve( hyg_batch bt, const hyg_func* x64, const hyg_func* b64, int inner, int max_outer, double rel_tol,
                        int use_graph, double* res, int* steps, void* stream )
{
   if ( !bt || !x64 || !b64 || !res || !steps || inner < 1 || max_outer < 0 )
      return fail( HYG_ERR_ARG, "bad argument" );
   if ( bt->dtype != HYG_F32 )
      return fail( HYG_ERR_ARG, "defect correction runs FP32 batched cycles" );
   const size_t            n = bt->g.size();
   std::vector< uint64_t > key{ uint64_t( inner ) };
   for ( size_t i = 0; i < n; ++i )
   {
      hyg_gmg_s* g = bt->g[i];
      HYG_TRY( checkPair( x64[i], b64[i], g->maxL ) );
      HYG_TRY( checkOpFunc( g->op, x64[i] ) );
      if ( x64[i]->dtype != HYG_F64 )
         return fail( HYG_ERR_ARG, "defect correction works on FP64 x and b" );
      key.push_back( x64[i]->gen );
      key.push_back( b64[i]->gen );
   }
   hipStream_t st = S( stream );
   if ( key != bt->irKey )
   {
      // FP64 residual and correction of every problem at its top level
      if ( bt->r64.empty() )
         for ( hyg_gmg_s* g : bt->g )
         {
            hyg_func_s *r = nullptr, *t = nullptr;
            HYG_TRY( createFunc( g->op->mesh, g->maxL, g->maxL, HYG_F64, &r ) );
            bt->r64.push_back( r );
            HYG_TRY( createFunc( g->op->mesh, g->maxL, g->maxL, HYG_F64, &t ) );
            bt->t64.push_back( t );
         }
      std::vector< hyg_func_s* > xs( x64, x64 + n ), bs( b64, b64 + n );
      // before the cycles: b32 = (float) r64, x32 = 0; after: t64 = (double) x32, x64 += t64, r64 = b64 - A x64, r64 . r64
      bt->irPre.reset( new BatchPlan );
      bt->irPost.reset( new BatchPlan );
      std::vector< BatchPlan::Convert > toF, toD;
      std::vector< BatchPlan::Vec >     zero, add;
      for ( size_t i = 0; i < n; ++i )
      {
         hyg_gmg_s* g  = bt->g[i];
         const int  L  = g->maxL;
         MeshCtx&   mc = g->op->mesh->ctx;
         const i64  nv = levelValues( bt->x[i], L );
         toF.push_back( BatchPlan::Convert{ nv, bt->b[i]->data[L], bt->r64[i]->data[L] } );
         BatchPlan::Vec z{ &mc, &mc.geom( L ), L, HYG_ALL, 0 };
         z.dst = bt->x[i]->data[L];
         zero.push_back( z );
         toD.push_back( BatchPlan::Convert{ nv, bt->t64[i]->data[L], bt->x[i]->data[L] } );
         BatchPlan::Vec a{ &mc, &mc.geom( L ), L, kGmgFlag, 1 };
         a.acc    = true;
         a.dst    = x64[i]->data[L];
         a.src[0] = bt->t64[i]->data[L];
         a.a[0]   = 1.0;
         add.push_back( a );
      }
      bt->irPre->convert( HYG_F32, HYG_F64, toF );
      bt->irPre->vec( HYG_F32, zero );
      bt->irPost->convert( HYG_F64, HYG_F32, toD );
      bt->irPost->vec( HYG_F64, add );
      bt->irNormStart = bt->irPost->steps.size();
      buildNorms( bt, *bt->irPost, HYG_F64, xs, bs, bt->r64 );
      HYG_HIP( bt->irPre->upload() );
      HYG_HIP( bt->irPost->upload() );
      if ( bt->irExec )
      {
         (void) hipGraphExecDestroy( bt->irExec );
         bt->irExec = nullptr;
      }
      bt->irKey   = key;
      bt->irInner = inner;
   }
   // initial residual: the norm steps at the end of irPost (residual, dot partial sums, final sums)
   bt->irPost->launch( st, bt->irNormStart );
   HYG_TRY( launched( "batched residual" ) );
   std::vector< double > r( n );
   HYG_TRY( batchReadNorms( bt, r.data(), st ) );
   const size_t stride = size_t( max_outer ) + 1;
   for ( size_t i = 0; i < n; ++i )
      res[i * stride] = r[i];
   std::vector< double > r0 = r;
   *steps = 0;
   for ( int it = 0; it < max_outer; ++it )
   {
      bool done = true;
      for ( size_t i = 0; i < n; ++i )
         done = done && ( r0[i] == 0.0 || r[i] / r0[i] <= rel_tol );
      if ( done )
         break;
      if ( use_graph )
      {
         if ( !bt->irExec )
         {
            hipGraph_t graph = nullptr;
            HYG_HIP( hipStreamBeginCapture( st, hipStreamCaptureModeThreadLocal ) );
            bt->irPre->launch( st );
            for ( int k = 0; k < inner; ++k )
               bt->cycle->launch( st );
            bt->irPost->launch( st );
            const hipError_t e = hipStreamEndCapture( st, &graph );
            if ( e != hipSuccess )
               return fail( HYG_ERR_HIP, std::string( "capturing the defect-correction step: " ) + hipGetErrorString( e ) );
            const hipError_t e2 = hipGraphInstantiate( &bt->irExec, graph, nullptr, nullptr, 0 );
            (void) hipGraphDestroy( graph );
            if ( e2 != hipSuccess )
            {
               bt->irExec = nullptr;
               return fail( HYG_ERR_HIP, std::string( "instantiating the defect-correction step: " ) + hipGetErrorString( e2 ) );
            }
         }
         HYG_HIP( hipGraphLaunch( bt->irExec, st ) );
      }
      else
      {
         bt->irPre->launch( st );
         for ( int k = 0; k < inner; ++k )
            bt->cycle->launch( st );
         bt->irPost->launch( st );
         HYG_TRY( launched( "batched defect correction" ) );
      }
      HYG_TRY( batchReadNorms( bt, r.data(), st ) );
      for ( size_t i = 0; i < n; ++i )
         res[i * stride + it + 1] = r[i];
      *steps = it + 1;
   }
   return HYG_OK;
}

} // extern "C"
