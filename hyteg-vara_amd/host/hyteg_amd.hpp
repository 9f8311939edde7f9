// C++ host classes over the C ABI of include/hyteg_amd.h, shaped like the reference's interfaces so that
// code written against HyTeG's operator/solver API reads the same:
//   Operator::apply(src, dst, level, flag, updateType)          operators/Operator.hpp:73-80
//   WeightedJacobiSmoothable::smooth_jac(dst, rhs, tmp, relax, level, flag)   solvers/Smoothables.hpp:32-43
//   SORSmoothable::smooth_sor(dst, rhs, relax, level, flag)      solvers/Smoothables.hpp:72-98
//   Solver<Op>::solve(A, x, b, level)                            solvers/Solver.hpp:26-36
//   ProlongationOperator::prolongate / prolongateAndAdd          gridtransferoperators/ProlongationOperator.hpp:30-39
//   RestrictionOperator::restrict                                gridtransferoperators/RestrictionOperator.hpp:29-35
//   GeometricMultigridSolver(storage, smoother, coarse, restriction, prolongation, minL, maxL, pre, post, inc)
//                                                                solvers/GeometricMultigridSolver.hpp:42-121
// Functions are passed by const& and mutated, as in the reference. A failing call throws
// std::runtime_error carrying hyg_last_error() (the reference aborts or throws from its smoothers).
#pragma once

#include "../../include/hyteg_amd.h"

#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace hyteg_amd {

enum DoFType : int
{
   None              = HYG_NONE,
   Inner             = HYG_INNER,
   DirichletBoundary = HYG_DIRICHLET,
   NeumannBoundary   = HYG_NEUMANN,
   FreeslipBoundary  = HYG_FREESLIP,
   All               = HYG_ALL,
};
enum UpdateType : int
{
   Replace = HYG_REPLACE,
   Add     = HYG_ADD,
};

inline void check( int rc )
{
   if ( rc != HYG_OK )
      throw std::runtime_error( std::string( "hyteg_amd: " ) + hyg_last_error() );
}

// Device selection; call once per process before anything else.
inline void initialise( int device = 0 ) { check( hyg_init( device ) ); }

// Macro mesh + boundary flags (MeshInfo::fromGmshFile / meshSphericalShell / meshSymmetricCuboid +
// SetupPrimitiveStorage + PrimitiveStorage).
class PrimitiveStorage
{
 public:
   explicit PrimitiveStorage( const std::string& gmshFile )
   {
      hyg_mesh m = nullptr;
      check( hyg_mesh_create_from_msh( gmshFile.c_str(), &m ) );
      h_.reset( m, []( hyg_mesh p ) { hyg_mesh_destroy( p ); } );
   }
   // MeshInfo::meshSphericalShell( ntan, nrad, rmin, rmax )
   static std::shared_ptr< PrimitiveStorage > sphericalShell( int ntan, int nrad, double rmin, double rmax )
   {
      hyg_mesh m = nullptr;
      check( hyg_mesh_create_spherical_shell( ntan, nrad, rmin, rmax, &m ) );
      return std::shared_ptr< PrimitiveStorage >( new PrimitiveStorage( m ) );
   }
   // MeshInfo::meshSymmetricCuboid( lowerLeftFront, upperRightBack, nx, ny, nz )
   static std::shared_ptr< PrimitiveStorage > symmetricCuboid( const double lower[3], const double upper[3], int nx, int ny, int nz )
   {
      hyg_mesh m = nullptr;
      check( hyg_mesh_create_symmetric_cuboid( lower, upper, nx, ny, nz, &m ) );
      return std::shared_ptr< PrimitiveStorage >( new PrimitiveStorage( m ) );
   }
   hyg_mesh handle() const { return h_.get(); }
   // vertex DoFs of a P1 level in canonical order (a P2 level L has the DoFs of P1 level L + 1)
   int64_t numberOfDoFs( int level ) const
   {
      int64_t n = 0;
      check( hyg_mesh_num_dofs( handle(), level, &n ) );
      return n;
   }

 private:
   explicit PrimitiveStorage( hyg_mesh m ) { h_.reset( m, []( hyg_mesh p ) { hyg_mesh_destroy( p ); } ); }
   std::shared_ptr< hyg_mesh_s > h_;
};

// Device arrays of a P1Function<ValueType> or P2Function<ValueType> (double or float) for levels
// minLevel..maxLevel. A P2 function of level L exchanges its values in the canonical order of P1 level L + 1.
template < typename ValueType, bool IsP2 >
class LatticeFunction
{
   static_assert( std::is_same< ValueType, double >::value || std::is_same< ValueType, float >::value, "double or float" );

 public:
   using valueType = ValueType;
   LatticeFunction( const std::string& name, const std::shared_ptr< PrimitiveStorage >& storage, int minLevel, int maxLevel )
   : name_( name )
   , storage_( storage )
   {
      hyg_func  f  = nullptr;
      const int dt = sizeof( ValueType ) == 8 ? HYG_F64 : HYG_F32;
      check( IsP2 ? hyg_func_create_p2( storage->handle(), minLevel, maxLevel, dt, &f )
                  : hyg_func_create( storage->handle(), minLevel, maxLevel, dt, &f ) );
      h_.reset( f, []( hyg_func p ) { hyg_func_destroy( p ); } );
   }
   hyg_func           handle() const { return h_.get(); }
   const std::string& getFunctionName() const { return name_; }

   void interpolate( double constant, int level, DoFType flag = All ) const
   {
      check( hyg_interpolate_const( handle(), constant, level, flag, stream_ ) );
   }
   void assign( const std::vector< double >& scalars, const std::vector< LatticeFunction* >& functions, int level,
                DoFType flag = All ) const
   {
      std::vector< hyg_func > f;
      for ( auto* p : functions )
         f.push_back( p->handle() );
      check( hyg_assign( handle(), int( f.size() ), scalars.data(), f.data(), level, flag, stream_ ) );
   }
   void add( const std::vector< double >& scalars, const std::vector< LatticeFunction* >& functions, int level,
             DoFType flag = All ) const
   {
      std::vector< hyg_func > f;
      for ( auto* p : functions )
         f.push_back( p->handle() );
      check( hyg_add( handle(), int( f.size() ), scalars.data(), f.data(), level, flag, stream_ ) );
   }
   void multElementwise( const LatticeFunction& a, const LatticeFunction& b, int level, DoFType flag = All ) const
   {
      check( hyg_mult_elementwise( handle(), a.handle(), b.handle(), level, flag, stream_ ) );
   }
   double dotGlobal( const LatticeFunction& other, int level, DoFType flag = All ) const
   {
      double r = 0.0;
      check( hyg_dot( handle(), other.handle(), level, flag, &r, stream_ ) );
      return r;
   }
   template < typename Other >
   void copyFrom( const LatticeFunction< Other, IsP2 >& other, int level ) const
   {
      check( hyg_convert( handle(), other.handle(), level, stream_ ) );
   }
   // canonical-order host copies (double on the host side)
   void setValues( int level, const std::vector< double >& v ) const { check( hyg_func_upload( handle(), level, v.data(), stream_ ) ); }
   std::vector< double > getValues( int level ) const
   {
      std::vector< double > v( size_t( storage_->numberOfDoFs( level + ( IsP2 ? 1 : 0 ) ) ) );
      check( hyg_func_download( handle(), level, v.data(), stream_ ) );
      return v;
   }
   void setStream( void* s ) { stream_ = s; }

 private:
   std::string                         name_;
   std::shared_ptr< PrimitiveStorage > storage_;
   std::shared_ptr< hyg_func_s >       h_;
   void*                               stream_ = nullptr;
};
template < typename ValueType >
using P1Function = LatticeFunction< ValueType, false >;
template < typename ValueType >
using P2Function = LatticeFunction< ValueType, true >;

// The Operator<Src, Dst> interface shared by every operator handle: apply, smooth_jac, inverse diagonal.
template < typename Function >
class MatrixFreeOperator
{
 public:
   using srcType = Function;
   using dstType = Function;
   hyg_op handle() const { return h_.get(); }
   int    minLevel() const { return minLevel_; }
   int    maxLevel() const { return maxLevel_; }

   void apply( const srcType& src, const dstType& dst, int level, DoFType flag, UpdateType updateType = Replace ) const
   {
      check( hyg_apply( handle(), src.handle(), dst.handle(), level, flag, updateType, stream_ ) );
   }
   void smooth_jac( const dstType& dst, const dstType& rhs, const srcType& tmp, double relax, int level, DoFType flag ) const
   {
      check( hyg_smooth_jac( handle(), dst.handle(), rhs.handle(), tmp.handle(), relax, level, flag, stream_ ) );
   }
   // writes D^-1 into `dst` (getInverseDiagonalValues; the values are computed at construction)
   void getInverseDiagonalValues( const dstType& dst, int level ) const
   {
      check( hyg_inverse_diagonal( handle(), dst.handle(), level, stream_ ) );
   }
   void setStream( void* s ) { stream_ = s; }

 protected:
   MatrixFreeOperator( const std::shared_ptr< PrimitiveStorage >& storage, int minLevel, int maxLevel, hyg_op op )
   : storage_( storage )
   , minLevel_( minLevel )
   , maxLevel_( maxLevel )
   {
      h_.reset( op, []( hyg_op p ) { hyg_op_destroy( p ); } );
   }
   void* stream() const { return stream_; }

 private:
   std::shared_ptr< PrimitiveStorage > storage_;
   std::shared_ptr< hyg_op_s >         h_;
   int                                 minLevel_, maxLevel_;
   void*                               stream_ = nullptr;
};

// P1ConstantLaplaceOperator (P1ConstantOperator<P1FenicsForm<p1_diffusion, p1_tet_diffusion>>), one value type.
template < typename ValueType >
class P1ConstantLaplaceOperator : public MatrixFreeOperator< P1Function< ValueType > >
{
 public:
   P1ConstantLaplaceOperator( const std::shared_ptr< PrimitiveStorage >& storage, int minLevel, int maxLevel )
   : MatrixFreeOperator< P1Function< ValueType > >( storage, minLevel, maxLevel, make( storage, minLevel, maxLevel ) )
   {}
   // SORSmoothable::smooth_sor / smooth_sor_backwards, smooth_gs (P1Operator.hpp:325-423), multi-colour order
   void smooth_sor( const P1Function< ValueType >& dst, const P1Function< ValueType >& rhs, double relax, int level, DoFType flag,
                    bool backwards = false ) const
   {
      check( hyg_smooth_sor( this->handle(), dst.handle(), rhs.handle(), relax, level, flag, backwards ? 1 : 0, this->stream() ) );
   }
   void smooth_sor_backwards( const P1Function< ValueType >& dst, const P1Function< ValueType >& rhs, double relax, int level,
                              DoFType flag ) const
   {
      smooth_sor( dst, rhs, relax, level, flag, true );
   }
   void smooth_gs( const P1Function< ValueType >& dst, const P1Function< ValueType >& rhs, int level, DoFType flag ) const
   {
      smooth_sor( dst, rhs, 1.0, level, flag, false );
   }

 private:
   static hyg_op make( const std::shared_ptr< PrimitiveStorage >& storage, int minLevel, int maxLevel )
   {
      hyg_op op = nullptr;
      check( hyg_op_create_p1_const( storage->handle(), minLevel, maxLevel, &op ) );
      return op;
   }
};

// P1ElementwiseOperator<p1_div_k_grad_affine_q3> with k = k0 exp( sum_m a_m cos(2 pi kappa_m . x + phi_m) ),
// modes: M rows (kappa_x, kappa_y, kappa_z, phi, a).
template < typename ValueType >
class P1ElementwiseDivKGradOperator : public MatrixFreeOperator< P1Function< ValueType > >
{
 public:
   P1ElementwiseDivKGradOperator( const std::shared_ptr< PrimitiveStorage >& storage, int minLevel, int maxLevel, double k0,
                                  const std::vector< std::array< double, 5 > >& modes )
   : MatrixFreeOperator< P1Function< ValueType > >( storage, minLevel, maxLevel, make( storage, minLevel, maxLevel, k0, modes ) )
   {}
   // polynomial coefficient k = sum_i coeffs[i] m_i(x), monomials of degree <= degree in MonomialBasis3D order
   P1ElementwiseDivKGradOperator( const std::shared_ptr< PrimitiveStorage >& storage, int minLevel, int maxLevel, int degree,
                                  const std::vector< double >& coeffs )
   : MatrixFreeOperator< P1Function< ValueType > >( storage, minLevel, maxLevel, makePoly( storage, minLevel, maxLevel, degree, coeffs ) )
   {}
   void computeAndStoreLocalElementMatrices() const { check( hyg_op_store_local_matrices( this->handle() ) ); }

 private:
   static hyg_op make( const std::shared_ptr< PrimitiveStorage >& storage, int minLevel, int maxLevel, double k0,
                       const std::vector< std::array< double, 5 > >& modes )
   {
      hyg_op op = nullptr;
      check( hyg_op_create_p1_divkgrad( storage->handle(), minLevel, maxLevel, k0, modes.empty() ? nullptr : modes[0].data(),
                                        int( modes.size() ), &op ) );
      return op;
   }
   static hyg_op makePoly( const std::shared_ptr< PrimitiveStorage >& storage, int minLevel, int maxLevel, int degree,
                           const std::vector< double >& coeffs )
   {
      if ( int( coeffs.size() ) != ( degree + 1 ) * ( degree + 2 ) * ( degree + 3 ) / 6 )
         throw std::runtime_error( "polynomial coefficient count does not match the degree" );
      hyg_op op = nullptr;
      check( hyg_op_create_p1_divkgrad_poly( storage->handle(), minLevel, maxLevel, degree, coeffs.data(), &op ) );
      return op;
   }
};

// P1SurrogateOperator< p1_div_k_grad_affine_q3 > (p1functionspace/P1SurrogateOperator.hpp): constructed like the
// elementwise operator; interpolateStencils( polyDegree, interpolationLevel ) switches the cell interiors to the
// fitted polynomial stencils.
template < typename ValueType >
class P1SurrogateAffineDivkGradOperator : public P1ElementwiseDivKGradOperator< ValueType >
{
 public:
   using P1ElementwiseDivKGradOperator< ValueType >::P1ElementwiseDivKGradOperator;
   void interpolateStencils( int polyDegree, int interpolationLevel ) const
   {
      check( hyg_op_interpolate_stencils( this->handle(), polyDegree, interpolationLevel ) );
   }
};

// P2ConstantLaplaceOperator, one value type.
template < typename ValueType >
class P2ConstantLaplaceOperator : public MatrixFreeOperator< P2Function< ValueType > >
{
 public:
   P2ConstantLaplaceOperator( const std::shared_ptr< PrimitiveStorage >& storage, int minLevel, int maxLevel )
   : MatrixFreeOperator< P2Function< ValueType > >( storage, minLevel, maxLevel, make( storage, minLevel, maxLevel ) )
   {}

 private:
   static hyg_op make( const std::shared_ptr< PrimitiveStorage >& storage, int minLevel, int maxLevel )
   {
      hyg_op op = nullptr;
      check( hyg_op_create_p2_const( storage->handle(), minLevel, maxLevel, &op ) );
      return op;
   }
};

// P2ElementwiseOperator<p2_div_k_grad_affine_q4>; kind 0: k = k0 exp(alpha (1 - |x|)) (1 + beta xi_e) with xi_e
// seeded per micro-element, kind 1: k = 1 + x/2 + y^2/4.
template < typename ValueType >
class P2ElementwiseDivKGradOperator : public MatrixFreeOperator< P2Function< ValueType > >
{
 public:
   P2ElementwiseDivKGradOperator( const std::shared_ptr< PrimitiveStorage >& storage, int minLevel, int maxLevel, int kind, double k0,
                                  double alpha, double beta, uint64_t seed )
   : MatrixFreeOperator< P2Function< ValueType > >( storage, minLevel, maxLevel,
                                                    make( storage, minLevel, maxLevel, kind, k0, alpha, beta, seed ) )
   {}
   void computeAndStoreLocalElementMatrices() const { check( hyg_op_store_local_matrices( this->handle() ) ); }

 private:
   static hyg_op make( const std::shared_ptr< PrimitiveStorage >& storage, int minLevel, int maxLevel, int kind, double k0, double alpha,
                       double beta, uint64_t seed )
   {
      hyg_op op = nullptr;
      check( hyg_op_create_p2_divkgrad( storage->handle(), minLevel, maxLevel, kind, k0, alpha, beta, seed, &op ) );
      return op;
   }
};

class P1toP1LinearRestriction
{
 public:
   template < typename F >
   void restrict( const F& function, int sourceLevel, DoFType flag ) const
   {
      check( hyg_restrict( function.handle(), sourceLevel, flag, nullptr ) );
   }
};

class P1toP1LinearProlongation
{
 public:
   template < typename F >
   void prolongate( const F& function, int sourceLevel, DoFType flag ) const
   {
      check( hyg_prolongate( function.handle(), sourceLevel, flag, HYG_REPLACE, nullptr ) );
   }
   template < typename F >
   void prolongateAndAdd( const F& function, int sourceLevel, DoFType flag ) const
   {
      check( hyg_prolongate( function.handle(), sourceLevel, flag, HYG_ADD, nullptr ) );
   }
};

// P2toP2QuadraticRestriction / P2toP2QuadraticProlongation: the same calls on a P2 function (the library
// applies the quadratic transfers to P2 handles)
using P2toP2QuadraticRestriction  = P1toP1LinearRestriction;
using P2toP2QuadraticProlongation = P1toP1LinearProlongation;

// Smoother / coarse-solver descriptions consumed by GeometricMultigridSolver (the library runs them).
struct WeightedJacobiSmoother
{
   double relax = 2.0 / 3.0;
};
struct ChebyshevSmoother
{
   int    order = 2;
   double lower = 0.0, upper = 0.0;
   // setupCoefficients(order, spectralRadius): bounds [0.3, 1.2] x spectralRadius (ChebyshevSmoother.hpp:109-190)
   void setupCoefficients( int ord, double spectralRadius )
   {
      order = ord;
      lower = 0.3 * spectralRadius;
      upper = 1.2 * spectralRadius;
   }
};
struct CGSolver
{
   int    maxIter   = 10000;
   double tolerance = 1e-16;
};

// chebyshev::estimateRadius: power iteration on D^-1 A with `tmp` as workspace; x holds the start vector.
template < typename Op, typename F >
double estimateRadius( const Op& A, int level, int iterations, const F& x, const F& tmp )
{
   double rho = 0.0;
   check( hyg_power_iteration( A.handle(), x.handle(), tmp.handle(), iterations, level, &rho, nullptr ) );
   return rho;
}

template < typename Op >
class GeometricMultigridSolver
{
 public:
   using FunctionType = typename Op::srcType;
   template < typename Smoother >
   GeometricMultigridSolver( const std::shared_ptr< PrimitiveStorage >&, const Op& A, const Smoother& smoother, const CGSolver& coarse,
                             int minLevel, int maxLevel, int preSmoothSteps = 3, int postSmoothSteps = 3,
                             int smoothIncrementOnCoarserGrids = 0 )
   {
      int    ip[8] = { minLevel, maxLevel, preSmoothSteps, postSmoothSteps, smoothIncrementOnCoarserGrids, 0, 1, coarse.maxIter };
      double dp[4] = { 2.0 / 3.0, 0.0, 0.0, coarse.tolerance };
      setSmoother( smoother, ip, dp );
      hyg_gmg g = nullptr;
      check( hyg_gmg_create( A.handle(), sizeof( typename FunctionType::valueType ) == 8 ? HYG_F64 : HYG_F32, ip, dp, &g ) );
      h_.reset( g, []( hyg_gmg p ) { hyg_gmg_destroy( p ); } );
   }
   hyg_gmg handle() const { return h_.get(); }
   // one V-cycle (Solver::solve)
   void solve( const Op&, const FunctionType& x, const FunctionType& b, int level ) const
   {
      check( hyg_gmg_vcycle( handle(), x.handle(), b.handle(), level, nullptr ) );
   }
   // V-cycles until the residual dropped by relTol; returns the residual norms (FP64)
   std::vector< double > solveTo( const FunctionType& x, const FunctionType& b, int level, int maxCycles, double relTol ) const
   {
      std::vector< double > res( size_t( maxCycles ) + 1 );
      int                   n = 0;
      check( hyg_gmg_solve_tol( handle(), x.handle(), b.handle(), level, maxCycles, relTol, res.data(), &n, nullptr ) );
      res.resize( size_t( n ) + 1 );
      return res;
   }

 private:
   static void setSmoother( const WeightedJacobiSmoother& s, int* ip, double* dp )
   {
      ip[5] = 0;
      dp[0] = s.relax;
   }
   static void setSmoother( const ChebyshevSmoother& s, int* ip, double* dp )
   {
      ip[5] = 1;
      ip[6] = s.order;
      dp[1] = s.lower;
      dp[2] = s.upper;
   }
   std::shared_ptr< hyg_gmg_s > h_;
};

// Mixed-precision defect correction: FP64 residual and update, `inner` FP32 V-cycles per step (the FP32
// multigrid of the same operator handle, built on an FP32 function type).
template < typename Op64, typename Op32 >
std::vector< double > mixedPrecisionSolve( const Op64& A64, const GeometricMultigridSolver< Op32 >& gmg32, const typename Op64::srcType& x,
                                           const typename Op64::srcType& b, int level, int inner, int maxOuter, double relTol )
{
   std::vector< double > res( size_t( maxOuter ) + 1 );
   int                   n = 0;
   check( hyg_ir_solve( A64.handle(), gmg32.handle(), x.handle(), b.handle(), level, inner, maxOuter, relTol, res.data(), &n,
                        nullptr ) );
   res.resize( size_t( n ) + 1 );
   return res;
}

} // namespace hyteg_amd
