// C-ABI solver drivers: instantiate the header-only HyTeG-shaped solver
// classes of include/hyteg_amd/hyteg_amd.hpp on existing handles.
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>

#include "hyteg_amd/hyteg_amd.hpp"

using namespace hyteg_amd;

namespace {

std::shared_ptr< PrimitiveStorage > storageOf( hg_ctx* ctx, hg_handle fn )
{
   hg_handle s;
   check( ctx, hg_function_storage( ctx, fn, &s ), "storage of function" );
   return PrimitiveStorage::borrow( ctx, s );
}

double residualNorm( const P1ConstantLaplaceOperator& A, const P1Function< double >& x, const P1Function< double >& b,
                     const P1Function< double >& r, uint_t level )
{
   A.apply( x, r, level, Inner, Replace );
   r.assign( { 1.0, -1.0 }, { b, r }, level, Inner );
   return std::sqrt( r.dotGlobal( r, level, Inner ) );
}

template < class F >
int guarded( F&& f )
{
   try
   {
      f();
      return HG_OK;
   } catch ( const std::exception& )
   {
      return HG_ERR_STATE;
   }
}

} // namespace

extern "C" {

int hg_p1_cg_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int level, int maxIter, double relTol, double absTol,
                    int* iterations )
{
   if ( !ctx || maxIter < 0 )
      return HG_ERR_ARG;
   return guarded( [&] {
      auto storage = storageOf( ctx, x );
      auto A       = P1ConstantLaplaceOperator::borrow( ctx, op );
      auto X       = P1Function< double >::borrow( ctx, x );
      auto B       = P1Function< double >::borrow( ctx, b );
      CGSolver< P1ConstantLaplaceOperator > cg( storage, uint_t( level ), uint_t( level ), uint_t( maxIter ), relTol, absTol );
      cg.solve( *A, *X, *B, uint_t( level ) );
      if ( iterations )
         *iterations = int( cg.getIterations() );
   } );
}

int hg_p1_gmg_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel, int pre, int post,
                     int cycles, double tol, int coarseMaxIter, double coarseTol, double* res_hist, int* cycles_done )
{
   if ( !ctx || minLevel < 0 || maxLevel < minLevel || pre < 0 || post < 0 || cycles < 0 || coarseMaxIter < 0 )
      return HG_ERR_ARG;
   return guarded( [&] {
      auto storage = storageOf( ctx, x );
      auto A       = P1ConstantLaplaceOperator::borrow( ctx, op );
      auto X       = P1Function< double >::borrow( ctx, x );
      auto B       = P1Function< double >::borrow( ctx, b );
      auto smoother = std::make_shared< GaussSeidelSmoother< P1ConstantLaplaceOperator > >();
      auto coarse   = std::make_shared< CGSolver< P1ConstantLaplaceOperator > >( storage, uint_t( minLevel ), uint_t( minLevel ),
                                                                               uint_t( coarseMaxIter ), coarseTol, coarseTol );
      auto R        = std::make_shared< P1toP1LinearRestriction<> >();
      auto P        = std::make_shared< P1toP1LinearProlongation<> >();
      GeometricMultigridSolver< P1ConstantLaplaceOperator > gmg( storage, smoother, coarse, R, P, uint_t( minLevel ),
                                                                uint_t( maxLevel ), uint_t( pre ), uint_t( post ) );
      P1Function< double > r( "residual", storage, uint_t( maxLevel ), uint_t( maxLevel ) );
      const double         r0 = residualNorm( *A, *X, *B, r, uint_t( maxLevel ) );
      if ( res_hist )
         res_hist[0] = r0;
      int k = 0;
      for ( ; k < cycles; ++k )
      {
         gmg.solve( *A, *X, *B, uint_t( maxLevel ) );
         const double rk = residualNorm( *A, *X, *B, r, uint_t( maxLevel ) );
         if ( res_hist )
            res_hist[k + 1] = rk;
         if ( tol > 0 && rk <= tol * r0 )
         {
            ++k;
            break;
         }
      }
      if ( cycles_done )
         *cycles_done = k;
   } );
}

} // extern "C"
