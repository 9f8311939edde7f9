// Launch wrappers of the gfx950 kernels (implemented in kernels.hip).
#pragma once

#include "hyg_internal.hpp"

namespace hyg {

// Stencil modes; x = stencil input, A x computed on the fly, d = inverse diagonal
enum StencilMode : int
{
   MODE_APPLY     = 0, // out0 = A x
   MODE_APPLY_ADD = 1, // out0 += A x
   MODE_RESIDUAL  = 2, // out0 = in1 - A x
   MODE_JACOBI    = 3, // out0 = x + a * d * (in1 - A x)
   MODE_CHEB0     = 4, // t = d * (in1 - A x); out0 = t; out1 = x + a * t
   MODE_CHEBK     = 5, // t = d * A x;         out0 = t; out1 = in1 + a * t
   MODE_SOR       = 6, // one colour of multi-colour SOR: out0 = x + a * d * (in1 - A x) at points of that colour (out0 = x)
};

struct StencilCall
{
   const void* x    = nullptr; // stencil input
   const void* in1  = nullptr; // rhs (residual / Jacobi / Cheb0) or accumulator (ChebK)
   void*       out0 = nullptr;
   void*       out1 = nullptr;
   double      a    = 0.0;
   int         flag = 0;
   int         mode = 0;
};

// Runs the cell-interior kernel and the macro-face/edge/vertex kernel of one stencil operation.
void launchStencil( const MeshCtx& m, LevelGeom& g, const hyg_op_s::Level& ol, int L, int dtype, const StencilCall& c,
                    hipStream_t s );
// One multi-colour SOR sweep of the constant P1 operator on x (in place) with right-hand side b.
void launchSor( const MeshCtx& m, LevelGeom& g, const hyg_op_s::Level& ol, int L, int dtype, void* x, const void* b, double relax,
                int flag, bool backwards, hipStream_t s );

// One SOR sweep of the constant P1 operator in the reference's order (primitives by dimension, each lexicographic;
// neighbours outside a primitive's closure and of no higher dimension read from a snapshot taken before the sweep).
// Builds the level's tables and snapshot buffer on first use.
void launchSorOrdered( const MeshCtx& m, const LevelGeom& g, hyg_op_s::Level& ol, int L, int dtype, void* x, const void* b, double relax,
                       int flag, bool backwards, hipStream_t s );
// builds the tables and the snapshot buffer of one level and value type (so that later sweeps allocate nothing)
void prepareSorOrdered( const MeshCtx& m, const LevelGeom& g, hyg_op_s::Level& ol, int L, int dtype );

// Variable-coefficient elementwise operator (op.kind == 1): one stencil operation, and the inverse
// diagonal of level L written into op.levels[L].dinv / dinvF (setup).
void launchElementwise( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, int dtype, const StencilCall& c,
                        hipStream_t s );
void launchElementwiseInverseDiagonal( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, hipStream_t s );
void launchElementwiseStore( const MeshCtx& m, hyg_op_s& op, hipStream_t s );
void launchKBarTables( const MeshCtx& m, const hyg_op_s& op, int L, hipStream_t s );
// Surrogate stencils of the elementwise operator (P1SurrogateOperator): fit the cell-interior stencils of every
// level by polynomials of degree deg from the samples of the interpolation level (needs the stored coefficient
// integrals; synchronous), whether level L uses them, the interior part of one stencil operation, and the interior
// inverse diagonal (written into op.levels[L].dinv)
void surrogateInterpolate( const MeshCtx& m, hyg_op_s& op, int deg, int interpolationLevel );
// blended Laplacian of a shell mesh (op.blend): the stencil weights of every level (Level::bw, bwF)
void assembleBlendWeights( const MeshCtx& m, hyg_op_s& op, hipStream_t s );
bool surrogateActive( const hyg_op_s& op, int L );
void launchSurrogateInterior( const MeshCtx& m, const LevelGeom& g, const hyg_op_s& op, int L, int dtype, const StencilCall& c, hipStream_t s );
void launchSurrogateInverseDiagonal( const MeshCtx& m, const LevelGeom& g, const hyg_op_s& op, int L, hipStream_t s );
// P2 operator (kind 2) at lattice level L: the stencil modes, the inverse diagonal, and the quadratic transfers
// (fine lattice Lf from Lf-1; coarse lattice Lc from Lc+1)
void launchP2( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, int dtype, const StencilCall& c, hipStream_t s );
void launchP2InverseDiagonal( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, hipStream_t s );
void launchP2Coefficients( const MeshCtx& m, const hyg_op_s& op, int L, hipStream_t s );
void launchP2StoreStencils( const MeshCtx& m, hyg_op_s& op, int L, hipStream_t s );
void launchP2Vector( const MeshCtx& m, LevelGeom& g, const hyg_op_s& op, int L, int dtype, const StencilCall* cs, int n, hipStream_t s );
void launchP2Prolongate( const MeshCtx& m, const LevelGeom& gf, const P2Tables* tb, int Lf, int dtype, void* fine,
                         const void* coarse, int flag, bool add, hipStream_t st );
void launchP2Restrict( const MeshCtx& m, const LevelGeom& gc, const P2Tables* tb, int Lc, int dtype, void* coarse, const void* fine,
                       int flag, hipStream_t st );

// cell-interior kernel configuration (index into kCellRows / kCellPoints) of a level, value type and mode
int  cellConfig( LevelGeom& g, int dtype, int mode );
void setCellConfig( LevelGeom& g, int dtype, int mode, int cfg );
// P2 row-kernel configuration (0 .. kNumP2Cfgs-1: segment length kP2Chunks[cfg / 2], cfg % 2 + 1 planes in flight) of
// a lattice level, value type and mode
int  p2Config( const LevelGeom& g, int dtype, int mode );
void setP2Config( LevelGeom& g, int dtype, int mode, int cfg );

// Pointwise linear combination over all cell slots selected by flag:
// dst = (accumulate ? dst : 0) + sum_k s_k src_k  (nsrc <= 3), or dst = value if nsrc == 0.
void launchLincomb( const MeshCtx& m, const LevelGeom& g, int L, int dtype, void* dst, int nsrc, const double* s,
                    const void* const* src, bool accumulate, double value, int flag, hipStream_t st );
// dst = a * b pointwise (flag)
void launchMult( const MeshCtx& m, const LevelGeom& g, int L, int dtype, void* dst, const void* a, const void* b, int flag,
                 hipStream_t st );
// Dot product over owned DoFs selected by flag; the FP64 result is left in result[0] (device).
// `partials` must hold dotPartialsNeeded() doubles.
void launchDot( const MeshCtx& m, const LevelGeom& g, int L, int dtype, const void* a, const void* b, int flag, double* partials,
                double* result, hipStream_t st );
int  dotPartialsNeeded( const MeshCtx& m, const LevelGeom& g );
// conjugate-gradient steps with device-resident scalars sc[]: x += a p, r -= a Ap with a = sc[rsOld] / sc[pAp];
// p = r + (sc[rsNew] / sc[rsOld]) p; no step once sc[rsOld] < tol2
void launchCgUpdate( const MeshCtx& m, const LevelGeom& g, int L, int dtype, void* x, void* r, const void* p, const void* ap,
                     const double* sc, int rsOld, int pAp, double tol2, int flag, hipStream_t st );
void launchCgDirection( const MeshCtx& m, const LevelGeom& g, int L, int dtype, void* p, const void* r, const double* sc, int rsOld,
                        int rsNew, double tol2, int flag, hipStream_t st );
// value-type conversion over n values
void launchConvert( i64 n, int dstType, void* dst, int srcType, const void* src, hipStream_t st );
// plain copy of n values
void launchCopy( i64 n, int dtype, void* dst, const void* src, hipStream_t st );
// inverse diagonal written into a function (all slots)
void launchInvDiag( const MeshCtx& m, const LevelGeom& g, const hyg_op_s::Level& ol, int L, int dtype, void* dst, hipStream_t st );

// grid transfer: restriction level Lc+1 -> Lc, prolongation Lf-1 -> Lf
void launchRestrict( const MeshCtx& m, const LevelGeom& gc, int Lc, int dtype, void* coarse, const void* fine, int flag,
                     hipStream_t st );
void launchProlongate( const MeshCtx& m, const LevelGeom& gf, int Lf, int dtype, void* fine, const void* coarse, int flag,
                       bool add, hipStream_t st );

// reference-layout halos (HaloMaps): out[k] = f[gather[k]]; every replica of own value j = in[ownSrc[j]]; every replica
// of own value j (DoF type in flag) = the sum of its replicas
void launchHaloGather( const HaloMaps& h, int dtype, const void* f, void* out, hipStream_t st );
void launchHaloScatter( const HaloMaps& h, int dtype, const void* in, void* f, hipStream_t st );
void launchHaloSum( const HaloMaps& h, int dtype, int flag, void* f, hipStream_t st );

// ------------------------------------------------------------------ batched problems
// A sequence of batched grid operations (steps). Each step applies one operation to every problem of a batch
// (each at its own level, mesh, operator and functions) in one persistent launch over a task table; the
// builders append steps on the host, upload() copies every task table to the device once, launch() issues
// steps [from, to) in order (capturable: no allocation, copy or memset).
struct BatchPlan
{
   enum Phase : int
   {
      STENCIL, // constant-operator sweeps
      ELEM,    // elementwise-operator applies
      COMBINE,
      TRANSFER,
      VEC,
      DOTFINAL,
      CONVERT
   };
   struct Step
   {
      int                 phase = 0, mode = 0, dtype = 8, nTasks = 0, total = 0;
      int                 cfg   = 1; // sweep steps: cell-interior configuration (index of kCellRows / kCellDepth)
      std::vector< char > host;
      void*               dev = nullptr;
   };
   // one problem of a stencil operation (constant operators: kind 0, elementwise with stored integrals: kind 1)
   struct Item
   {
      const MeshCtx*  m;
      LevelGeom*      g;
      const hyg_op_s* op;
      int             L;
      StencilCall     c;
   };
   // restriction to level L (from L+1) or prolongation to level L (from L-1)
   struct Transfer
   {
      const MeshCtx*   m;
      const LevelGeom* g; // geometry of level L
      int              L, flag;
      bool             add;
      void*            dst;
      const void*      src;
   };
   // linear combination (dot = false) or dot-product partial sums of src[0] . src[1] into `partials`
   struct Vec
   {
      const MeshCtx*   m;
      const LevelGeom* g;
      int              L, flag, nsrc = 0;
      bool             acc = false, dot = false;
      void*            dst = nullptr;
      const void*      src[3] = { nullptr, nullptr, nullptr };
      double           a[3]   = { 0, 0, 0 };
      double           value  = 0;
      double*          partials = nullptr;
   };
   struct DotFinal
   {
      const double* partials;
      int           n;
      double*       result;
   };
   struct Convert
   {
      i64         n;
      void*       dst;
      const void* src;
   };

   std::vector< Step > steps;
   void*               dev = nullptr;

   static bool modeSupported( int mode );
   // sweep step (constant operators, cell-interior configuration cfg), elementwise step, combine step
   void        stencil( int dtype, int mode, const std::vector< Item >& items, int cfg );
   void        restrictLevels( int dtype, const std::vector< Transfer >& items );
   void        prolongate( int dtype, const std::vector< Transfer >& items );
   void        vec( int dtype, const std::vector< Vec >& items );
   void        dotFinal( const std::vector< DotFinal >& items );
   void        convert( int dstType, int srcType, const std::vector< Convert >& items );
   hipError_t  upload();
   void        launch( hipStream_t st, size_t from = 0, size_t to = size_t( -1 ) ) const;
   BatchPlan() = default;
   BatchPlan( const BatchPlan& ) = delete;
   BatchPlan& operator=( const BatchPlan& ) = delete;
   ~BatchPlan();
};

} // namespace hyg
