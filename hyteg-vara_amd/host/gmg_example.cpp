// Host program written against the C++ classes of hyteg_amd.hpp the way a HyTeG application is written:
// mesh -> functions -> operator -> apply -> geometric multigrid. Prints the apply result checksum and the
// residual history as one JSON line so that a test can compare it with the oracle.
//   usage: gmg_example <mesh.msh> <level> <cycles> <input.bin>
// input.bin holds the canonical-order vector (n doubles) used as apply source and as right-hand side.
#include "hyteg_amd.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>

using namespace hyteg_amd;

int main( int argc, char** argv )
{
   if ( argc < 5 )
   {
      std::fprintf( stderr, "usage: %s mesh level cycles input.bin\n", argv[0] );
      return 2;
   }
   const std::string mesh   = argv[1];
   const int         level  = std::atoi( argv[2] );
   const int         cycles = std::atoi( argv[3] );
   try
   {
      initialise( 0 );
      auto storage = std::make_shared< PrimitiveStorage >( mesh );
      const auto n = size_t( storage->numberOfDoFs( level ) );

      P1Function< double > u( "u", storage, 0, level ), f( "f", storage, 0, level ), Au( "Au", storage, 0, level );
      P1ConstantLaplaceOperator< double > A( storage, 0, level );

      std::vector< double > v( n );
      std::ifstream         in( argv[4], std::ios::binary );
      if ( !in.read( reinterpret_cast< char* >( v.data() ), std::streamsize( n * sizeof( double ) ) ) )
         throw std::runtime_error( std::string( "cannot read " ) + std::to_string( n ) + " doubles from " + argv[4] );
      u.setValues( level, v );
      A.apply( u, Au, level, Inner, Replace );
      const double applyNorm = std::sqrt( Au.dotGlobal( Au, level, Inner ) );

      // -Laplace x = f with f = v on inner DoFs, zero Dirichlet values
      f.setValues( level, v );
      f.interpolate( 0.0, level, DirichletBoundary );
      P1Function< double > x( "x", storage, 0, level );
      GeometricMultigridSolver< P1ConstantLaplaceOperator< double > > gmg( storage, A, WeightedJacobiSmoother{}, CGSolver{}, 0, level, 2,
                                                                             2 );
      std::printf( "{\"n\": %zu, \"apply_norm\": %.17g, \"residuals\": [", n, applyNorm );
      P1Function< double > r( "r", storage, 0, level );
      for ( int c = 0; c <= cycles; ++c )
      {
         if ( c > 0 )
            gmg.solve( A, x, f, level );
         A.apply( x, r, level, Inner, Replace );
         r.assign( { 1.0, -1.0 }, { &f, &r }, level, Inner );
         std::printf( "%s%.17g", c ? ", " : "", std::sqrt( r.dotGlobal( r, level, Inner ) ) );
      }
      std::printf( "]" );
      // the P2 operator on the same mesh: P2 level level - 1 holds the DoFs of P1 level `level`
      if ( level >= 1 )
      {
         P2Function< double >                p2u( "p2u", storage, level - 1, level - 1 ), p2Au( "p2Au", storage, level - 1, level - 1 );
         P2ConstantLaplaceOperator< double > A2( storage, level - 1, level - 1 );
         p2u.setValues( level - 1, v );
         A2.apply( p2u, p2Au, level - 1, Inner, Replace );
         std::printf( ", \"p2_apply_norm\": %.17g", std::sqrt( p2Au.dotGlobal( p2Au, level - 1, Inner ) ) );
      }
      std::printf( "}\n" );
   } catch ( const std::exception& e )
   {
      std::fprintf( stderr, "error: %s\n", e.what() );
      return 1;
   }
   return 0;
}
