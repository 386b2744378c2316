// C-ABI solver drivers: instantiate the header-only HyTeG-shaped solver
// classes of include/hyteg_amd/hyteg_amd.hpp on existing handles.
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>

#include "hyteg_amd/hyteg_amd.hpp"

namespace hg {
// records `msg` as the context's last error (defined with the context in runtime.hip)
void setContextError( hg_ctx* ctx, const char* msg );
} // namespace hg

using namespace hyteg_amd;

namespace {

using StokesOp = P2P1ElementwiseBlendingFullViscousStokesOperator;

std::shared_ptr< PrimitiveStorage > storageOf( hg_ctx* ctx, hg_handle fn )
{
   hg_handle s;
   check( ctx, hg_function_storage( ctx, fn, &s ), "storage of function" );
   return PrimitiveStorage::borrow( ctx, s );
}

template < class Op, class F >
double residualNorm( const Op& A, const F& x, const F& b, const F& r, uint_t level )
{
   A.apply( x, r, level, Inner, Replace );
   r.assign( { 1.0, -1.0 }, { b, r }, level, Inner );
   return std::sqrt( r.dotGlobal( r, level, Inner ) );
}

// runs a solver body; a failure is reported as HG_ERR_STATE with its message
// in hg_last_error (the failing ABI call inside already recorded its own)
template < class F >
int guarded( hg_ctx* ctx, F&& f )
{
   try
   {
      f();
      return HG_OK;
   } catch ( const std::exception& e )
   {
      hg::setContextError( ctx, e.what() );
      return HG_ERR_STATE;
   }
}

// Replays `body` (a V-cycle: kernels only) from a HIP graph captured on its
// first use when the context allows graphs; runs it directly otherwise.
class GraphedCycle
{
 public:
   explicit GraphedCycle( hg_ctx* ctx )
   : ctx_( ctx )
   {
      int en = 0;
      check( ctx_, hg_ctx_graphs_enabled( ctx_, &en ), "graphs" );
      enabled_ = en != 0;
   }
   ~GraphedCycle()
   {
      if ( graph_ )
         (void) hg_graph_destroy( ctx_, graph_ );
   }
   bool enabled() const { return enabled_; }
   template < class F >
   void run( F&& body )
   {
      if ( !enabled_ )
      {
         body();
         return;
      }
      if ( !graph_ )
      {
         check( ctx_, hg_graph_begin_capture( ctx_ ), "begin capture" );
         try
         {
            body();
         } catch ( ... )
         {
            hg_handle dummy = 0;
            (void) hg_graph_end_capture( ctx_, &dummy );
            throw;
         }
         check( ctx_, hg_graph_end_capture( ctx_, &graph_ ), "end capture" );
      }
      check( ctx_, hg_graph_launch( ctx_, graph_ ), "graph launch" );
   }

 private:
   hg_ctx*   ctx_;
   hg_handle graph_   = 0;
   bool      enabled_ = true;
};

// V-cycles of `gmg` until the Inner residual drops by `tol` (0: run all cycles)
template < class Op, class F, class G >
int cycle( const Op& A, G& gmg, const F& x, const F& b, const F& r, uint_t level, int cycles, double tol, double* hist,
           GraphedCycle* graph = nullptr )
{
   const double r0 = residualNorm( A, x, b, r, level );
   if ( hist )
      hist[0] = r0;
   int k = 0;
   for ( ; k < cycles; ++k )
   {
      // with a graph: the first cycle runs eagerly (it builds the per-level data),
      // the later ones replay the cycle captured at the second
      if ( graph && k > 0 )
         graph->run( [&] { gmg.solve( A, x, b, level ); } );
      else
         gmg.solve( A, x, b, level );
      const double rk = residualNorm( A, x, b, r, level );
      if ( hist )
         hist[k + 1] = rk;
      if ( tol > 0 && rk <= tol * r0 )
         return k + 1;
   }
   return k;
}

// Mixed-precision iterative refinement (mixedPrecisionIterativeRefinement.cpp:163-172):
// FP64 residual r = b - A x, scaled by the RMS of D^-1 r and cast to the low
// precision; `innerCycles` V-cycles of the low-precision hierarchy (FP32
// arithmetic, storage type L, FP64 PCG on the coarsest level) solve A e = r;
// x += scale * e in FP64.  Stops when the FP64 Inner residual drops by `tol`.
template < typename L >
int iterativeRefinement( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel, int pre,
                         int post, int order, double rho, int innerCycles, int maxOuter, double tol, int coarseMaxIter,
                         double coarseTol, double* hist )
{
   using Op64   = P2ElementwiseBlendingDivKGradOperator;
   using OpLo   = P2ElementwiseBlendingDivKGradOperatorT< L >;
   auto storage = storageOf( ctx, x );
   auto A       = Op64::borrow( storage, op, uint_t( minLevel ), uint_t( maxLevel ) );
   auto Alo     = OpLo::borrow( storage, op, uint_t( minLevel ), uint_t( maxLevel ) );
   auto X       = P2Function< double >::borrow( ctx, x );
   auto B       = P2Function< double >::borrow( ctx, b );
   const uint_t top = uint_t( maxLevel ), bottom = uint_t( minLevel );

   auto smoother = std::make_shared< ChebyshevSmoother< OpLo > >( storage, bottom, top );
   smoother->setupCoefficients( uint_t( order ), rho );
   auto pcg    = std::make_shared< CGSolver< Op64 > >( storage, bottom, bottom, uint_t( coarseMaxIter ), coarseTol, 1e-300,
                                                    std::make_shared< InverseDiagonalPreconditioner< Op64 > >() );
   auto coarse = std::make_shared< FP64CoarseSolver< OpLo, Op64 > >( storage, bottom, *A, pcg );
   GeometricMultigridSolver< OpLo > gmg( storage, smoother, coarse, std::make_shared< P2toP2QuadraticRestriction< L > >(),
                                         std::make_shared< P2toP2QuadraticProlongation< L > >(), bottom, top,
                                         uint_t( pre ), uint_t( post ) );
   P2Function< L >      xlo( "ir_x", storage, bottom, top ), blo( "ir_b", storage, bottom, top );
   P2Function< double > r( "ir_r", storage, top, top ), e( "ir_e", storage, top, top );
   int64_t              n = 0;
   check( ctx, hg_p2_ndofs( ctx, storage->handle(), maxLevel, &n ), "ndofs" );
   const DoFType flag = Inner | NeumannBoundary | FreeslipBoundary;

   const auto invDiag = A->getInverseDiagonalValues();
   double     r0      = 0.0;
   int        k       = 0;
   for ( ;; ++k )
   {
      // r = b - A x on the solver's DoFs; its Inner norm is the convergence measure
      A->apply( *X, r, top, flag, Replace );
      r.assign( { 1.0, -1.0 }, { *B, r }, top, flag );
      const double rk = std::sqrt( r.dotGlobal( r, top, Inner ) );
      if ( k == 0 )
         r0 = rk;
      if ( hist )
         hist[k] = rk;
      if ( k == maxOuter || ( k > 0 && tol > 0 && rk <= tol * r0 ) )
         return k;
      // scaled so that its Jacobi correction D^-1 r (the size of the unknown
      // correction) has unit RMS: keeps the low-precision vectors of every level
      // inside the FP16 range
      e.multElementwise( { *invDiag, r }, top, flag );
      const double scale = std::sqrt( e.dotGlobal( e, top, flag ) / double( n ) );
      if ( !( scale > 0 ) )
         return k;
      r.assign( { 1.0 / scale }, { r }, top, flag );
      blo.copyFrom( r, top );
      xlo.setToZero( top );
      for ( int j = 0; j < innerCycles; ++j )
         gmg.solve( *Alo, xlo, blo, top );
      e.copyFrom( xlo, top );
      X->add( { scale }, { e }, top, flag );
   }
}


// ---- interleaved multi-vector P1 solves -------------------------------------
// per-vector Inner residual norms of b - A x (r is scratch)
template < class Op, class F >
std::vector< double > residualNorms( const Op& A, const F& x, const F& b, const F& r, uint_t level )
{
   A.apply( x, r, level, Inner, Replace );
   r.assign( { 1.0, -1.0 }, { b, r }, level, Inner );
   std::vector< double > n = r.dotGlobalBatched( r, level, Inner );
   for ( double& v : n )
      v = std::sqrt( v );
   return n;
}

bool allBelow( const std::vector< double >& r, const std::vector< double >& r0, double tol )
{
   for ( size_t k = 0; k < r.size(); ++k )
      if ( r[k] > tol * r0[k] )
         return false;
   return true;
}

void storeHist( double* hist, int k, const std::vector< double >& r )
{
   if ( hist )
      for ( size_t j = 0; j < r.size(); ++j )
         hist[size_t( k ) * r.size() + j] = r[j];
}

// GMG V(pre,post) with Gauss-Seidel smoothing on functions of one value type
// holding nvec interleaved vectors; the coarsest level gets coarseSweeps sweeps
template < typename VT >
std::unique_ptr< GeometricMultigridSolver< P1ConstantLaplaceOperatorT< VT > > >
    batchedGMG( const std::shared_ptr< PrimitiveStorage >& storage, int minLevel, int maxLevel, int pre, int post,
                int coarseSweeps, int nvec )
{
   using Op      = P1ConstantLaplaceOperatorT< VT >;
   auto smoother = std::make_shared< GaussSeidelSmoother< Op > >();
   auto coarse   = std::make_shared< SmootherSweeps< Op > >( smoother, uint_t( coarseSweeps ) );
   return std::make_unique< GeometricMultigridSolver< Op > >(
       storage, smoother, coarse, std::make_shared< P1toP1LinearRestriction< VT > >(),
       std::make_shared< P1toP1LinearProlongation< VT > >(), uint_t( minLevel ), uint_t( maxLevel ), uint_t( pre ),
       uint_t( post ), 0, CycleType::VCYCLE, nvec );
}

// FP64 iterative refinement around the low-precision batched V-cycle; every
// vector's residual is scaled by its own RMS / (stencil centre) so that the
// correction is O(1) in FP16
template < typename L >
int batchedIR( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel, int pre, int post,
               int coarseSweeps, int innerCycles, int maxOuter, double tol, double* hist )
{
   auto storage = storageOf( ctx, x );
   auto A       = P1ConstantLaplaceOperator::borrow( ctx, op );
   auto Alo     = P1ConstantLaplaceOperatorT< L >::borrow( ctx, op );
   auto X       = P1Function< double >::borrow( ctx, x );
   auto B       = P1Function< double >::borrow( ctx, b );
   const int    nv  = X->nvec();
   const uint_t top = uint_t( maxLevel );
   auto         gmg = batchedGMG< L >( storage, minLevel, maxLevel, pre, post, coarseSweeps, nv );
   P1Function< L >      xlo( "bir_x", storage, uint_t( minLevel ), top, nv ), blo( "bir_b", storage, uint_t( minLevel ), top, nv );
   P1Function< double > r( "bir_r", storage, top, top, nv ), e( "bir_e", storage, top, top, nv );
   int64_t              n = 0;
   check( ctx, hg_p1_ndofs( ctx, storage->handle(), maxLevel, &n ), "ndofs" );
   double centre[15];
   check( ctx, hg_p1_cell_stencil( ctx, op, maxLevel, 0, centre ), "stencil" );
   const DoFType         flag = Inner | NeumannBoundary | FreeslipBoundary;
   std::vector< double > r0;
   GraphedCycle          cycleGraph( ctx );
   for ( int k = 0;; ++k )
   {
      A->apply( *X, r, top, flag, Replace );
      r.assign( { 1.0, -1.0 }, { *B, r }, top, flag );
      std::vector< double > rk = r.dotGlobalBatched( r, top, Inner );
      for ( double& v : rk )
         v = std::sqrt( v );
      if ( k == 0 )
         r0 = rk;
      storeHist( hist, k, rk );
      if ( k == maxOuter || ( k > 0 && tol > 0 && allBelow( rk, r0, tol ) ) )
         return k;
      std::vector< double > scale( static_cast< size_t >( nv ) ), inv( static_cast< size_t >( nv ) );
      const std::vector< double > ms = r.dotGlobalBatched( r, top, flag );
      for ( int j = 0; j < nv; ++j )
      {
         scale[j] = std::sqrt( ms[j] / double( n ) ) / std::fabs( centre[0] );
         inv[j]   = scale[j] > 0 ? 1.0 / scale[j] : 0.0;
      }
      r.scaleBatched( inv, top, flag );
      blo.copyFrom( r, top );
      xlo.setToZero( top );
      if ( k == 0 )
         gmg->solve( *Alo, xlo, blo, top ); // first cycle eager: creates every lazily built level
      else
         cycleGraph.run( [&] { gmg->solve( *Alo, xlo, blo, top ); } );
      for ( int j = 1; j < innerCycles; ++j )
         cycleGraph.run( [&] { gmg->solve( *Alo, xlo, blo, top ); } );
      e.copyFrom( xlo, top );
      e.scaleBatched( scale, top, flag );
      X->add( { 1.0 }, { e }, top, flag );
   }
}

} // namespace

namespace {
using TaylorHood = P2P1TaylorHoodFunction< double >;

// the Uzawa smoother of the free-slip GMG template (StokesFSGMGUzawaSolverTemplate.hpp:113-170):
// Chebyshev smoothing of the velocity block with coefficients from the spectral
// radius, inside FullStokesVelocityBlockBlockDiagonalPreconditioner; with a
// projection, the Chebyshev corrections and the velocity residual are projected
template < class Op >
std::shared_ptr< UzawaSmoother< Op > > stokesUzawa( const std::shared_ptr< PrimitiveStorage >& storage, int minLevel,
                                                    int maxLevel, double omega, int velocityIterations, int chebyshevOrder,
                                                    double spectralRadius, bool project,
                                                    const std::shared_ptr< P2ProjectNormalOperator >& projection )
{
   using Cheby = ChebyshevSmoother< typename Op::VelocityOperator_T, P2ProjectNormalOperator >;
   const uint_t minL = uint_t( minLevel ), maxL = uint_t( maxLevel );
   auto         cheby = projection ? std::make_shared< Cheby >( storage, minL, maxL, false, projection, FreeslipBoundary ) :
                                     std::make_shared< Cheby >( storage, minL, maxL );
   cheby->setupCoefficients( uint_t( chebyshevOrder ), spectralRadius );
   auto          velocity = std::make_shared< FullStokesVelocityBlockBlockDiagonalPreconditioner< Op > >( cheby );
   const DoFType flag     = DoFType( Inner | NeumannBoundary | FreeslipBoundary );
   if ( projection )
      return std::make_shared< UzawaSmootherWithFreeSlipProjection< Op > >( storage, velocity, minL, maxL, omega, flag,
                                                                             projection, uint_t( velocityIterations ),
                                                                             project );
   return std::make_shared< UzawaSmoother< Op > >( storage, velocity, minL, maxL, omega, flag,
                                                   uint_t( velocityIterations ), project );
}

// runs body(A, projection) on the elementwise Stokes operator, or (freeslip) on
// its free-slip wrapper P2P1FullStokesFreeslipWrapper with the radial projection
template < class Body >
void withStokesOperator( const std::shared_ptr< PrimitiveStorage >& storage, hg_handle op, uint_t minLevel,
                         uint_t maxLevel, bool freeslip, Body&& body )
{
   std::shared_ptr< StokesOp > A( StokesOp::borrow( storage, op, minLevel, maxLevel ) );
   if ( !freeslip )
   {
      body( *A, std::shared_ptr< P2ProjectNormalOperator >() );
      return;
   }
   auto                          projection = std::make_shared< P2ProjectNormalOperator >();
   P2P1FullStokesFreeslipWrapper W( storage, minLevel, maxLevel, A, projection );
   body( W, projection );
}

// residual 2-norm of a Stokes system over Inner | Neumann | FreeSlip
template < class Op >
double stokesResidualNorm( const Op& A, const TaylorHood& x, const TaylorHood& b, const TaylorHood& r, uint_t level )
{
   const DoFType flag = DoFType( Inner | NeumannBoundary | FreeslipBoundary );
   A.apply( x, r, level, flag, Replace );
   r.assign( { 1.0, -1.0 }, { b, r }, level, flag );
   return std::sqrt( r.dotGlobal( r, level, flag ) );
}
} // namespace

extern "C" {

int hg_p1_batched_gmg_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel, int pre,
                             int post, int coarseSweeps, int cycles, double tol, double* res_hist, int* cycles_done )
{
   if ( !ctx || minLevel < 0 || maxLevel < minLevel || pre < 0 || post < 0 || cycles < 0 || coarseSweeps < 0 )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      auto storage = storageOf( ctx, x );
      auto A       = P1ConstantLaplaceOperator::borrow( ctx, op );
      auto X       = P1Function< double >::borrow( ctx, x );
      auto B       = P1Function< double >::borrow( ctx, b );
      auto gmg     = batchedGMG< double >( storage, minLevel, maxLevel, pre, post, coarseSweeps, X->nvec() );
      P1Function< double >        r( "residual", storage, uint_t( maxLevel ), uint_t( maxLevel ), X->nvec() );
      const std::vector< double > r0 = residualNorms( *A, *X, *B, r, uint_t( maxLevel ) );
      storeHist( res_hist, 0, r0 );
      int          k = 0;
      GraphedCycle cycleGraph( ctx );
      while ( k < cycles )
      {
         if ( k == 0 )
            gmg->solve( *A, *X, *B, uint_t( maxLevel ) ); // eager: creates every lazily built level
         else
            cycleGraph.run( [&] { gmg->solve( *A, *X, *B, uint_t( maxLevel ) ); } );
         ++k;
         const std::vector< double > rk = residualNorms( *A, *X, *B, r, uint_t( maxLevel ) );
         storeHist( res_hist, k, rk );
         if ( tol > 0 && allBelow( rk, r0, tol ) )
            break;
      }
      if ( cycles_done )
         *cycles_done = k;
   } );
}

int hg_p1_batched_ir_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel,
                            int lowDtype, int pre, int post, int coarseSweeps, int innerCycles, int maxOuter, double tol,
                            double* res_hist, int* outer_done )
{
   if ( !ctx || minLevel < 0 || maxLevel < minLevel || pre < 0 || post < 0 || coarseSweeps < 0 || innerCycles < 1 ||
        maxOuter < 0 || ( lowDtype != HG_F32 && lowDtype != HG_F16 ) || !outer_done )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      *outer_done = lowDtype == HG_F32 ?
                        batchedIR< float >( ctx, op, x, b, minLevel, maxLevel, pre, post, coarseSweeps, innerCycles,
                                            maxOuter, tol, res_hist ) :
                        batchedIR< f16 >( ctx, op, x, b, minLevel, maxLevel, pre, post, coarseSweeps, innerCycles,
                                          maxOuter, tol, res_hist );
   } );
}

int hg_p1_cg_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int level, int maxIter, double relTol, double absTol,
                    int* iterations )
{
   if ( !ctx || maxIter < 0 )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
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
   return guarded( ctx, [&] {
      auto storage = storageOf( ctx, x );
      auto A       = P1ConstantLaplaceOperator::borrow( ctx, op );
      auto X       = P1Function< double >::borrow( ctx, x );
      auto B       = P1Function< double >::borrow( ctx, b );
      auto smoother = std::make_shared< GaussSeidelSmoother< P1ConstantLaplaceOperator > >();
      auto coarse   = std::make_shared< CGSolver< P1ConstantLaplaceOperator > >( storage, uint_t( minLevel ), uint_t( minLevel ),
                                                                               uint_t( coarseMaxIter ), coarseTol, coarseTol );
      GraphedCycle graph( ctx );
      if ( graph.enabled() )
      {
         // a V-cycle without host synchronisation, replayed from a HIP graph: the
         // coarse CG runs a fixed count, the number of free coarse DoFs (the
         // bound of CG in exact arithmetic) capped by coarseMaxIter, with its
         // stopping test on the device (the iterations after it holds leave the
         // iterate unchanged)
         const DoFType        freeDoFs = Inner | NeumannBoundary | FreeslipBoundary;
         P1Function< double > ones( "free", storage, uint_t( minLevel ), uint_t( minLevel ) );
         ones.interpolate( 1.0, uint_t( minLevel ), freeDoFs );
         const double nfree = ones.dotGlobal( ones, uint_t( minLevel ), freeDoFs );
         coarse             = std::make_shared< CGSolver< P1ConstantLaplaceOperator > >(
             storage, uint_t( minLevel ), uint_t( minLevel ), uint_t( std::min( double( coarseMaxIter ), nfree ) ),
             coarseTol, coarseTol );
         coarse->setFixedIterations( true );
      }
      auto R = std::make_shared< P1toP1LinearRestriction<> >();
      auto P = std::make_shared< P1toP1LinearProlongation<> >();
      GeometricMultigridSolver< P1ConstantLaplaceOperator > gmg( storage, smoother, coarse, R, P, uint_t( minLevel ),
                                                                uint_t( maxLevel ), uint_t( pre ), uint_t( post ) );
      P1Function< double > r( "residual", storage, uint_t( maxLevel ), uint_t( maxLevel ) );
      const int k = cycle( *A, gmg, *X, *B, r, uint_t( maxLevel ), cycles, tol, res_hist,
                           graph.enabled() ? &graph : nullptr );
      if ( cycles_done )
         *cycles_done = k;
   } );
}

int hg_p2_cg_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int level, int maxIter, double relTol,
                    double absTol, int jacobi, int* iterations )
{
   if ( !ctx || maxIter < 0 || level < 0 )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      using Op     = P2ElementwiseBlendingDivKGradOperator;
      auto storage = storageOf( ctx, x );
      auto A       = Op::borrow( storage, op, uint_t( level ), uint_t( level ) );
      auto X       = P2Function< double >::borrow( ctx, x );
      auto B       = P2Function< double >::borrow( ctx, b );
      std::shared_ptr< Solver< Op > > prec;
      if ( jacobi )
         prec = std::make_shared< InverseDiagonalPreconditioner< Op > >();
      CGSolver< Op > cg( storage, uint_t( level ), uint_t( level ), uint_t( maxIter ), relTol, absTol, prec );
      cg.solve( *A, *X, *B, uint_t( level ) );
      if ( iterations )
         *iterations = int( cg.getIterations() );
   } );
}

int hg_p2_chebyshev_smooth( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int level, int order,
                            double spectralRadius, int steps )
{
   if ( !ctx || level < 0 || order < 1 || order > 16 || steps < 0 || !( spectralRadius > 0 ) )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      using Op     = P2ElementwiseBlendingDivKGradOperator;
      auto storage = storageOf( ctx, x );
      auto A       = Op::borrow( storage, op, uint_t( level ), uint_t( level ) );
      auto X       = P2Function< double >::borrow( ctx, x );
      auto B       = P2Function< double >::borrow( ctx, b );
      ChebyshevSmoother< Op > smoother( storage, uint_t( level ), uint_t( level ) );
      smoother.setupCoefficients( uint_t( order ), spectralRadius );
      for ( int i = 0; i < steps; ++i )
         smoother.solve( *A, *X, *B, uint_t( level ) );
   } );
}

int hg_p2_estimate_radius( hg_ctx* ctx, hg_handle op, hg_handle x, int level, int iterations, double* radius )
{
   if ( !ctx || !radius || iterations < 1 || level < 0 )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      using Op     = P2ElementwiseBlendingDivKGradOperator;
      auto storage = storageOf( ctx, x );
      auto A       = Op::borrow( storage, op, uint_t( level ), uint_t( level ) );
      auto X       = P2Function< double >::borrow( ctx, x );
      P2Function< double > tmp( "radius_tmp", storage, uint_t( level ), uint_t( level ) );
      *radius = chebyshev::estimateRadius( *A, uint_t( level ), uint_t( iterations ), *X, tmp );
   } );
}

int hg_p2_gmg_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel, int pre, int post,
                     int chebyshevOrder, double spectralRadius, int cycles, double tol, int coarseMaxIter,
                     double coarseTol, int coarseJacobi, double* res_hist, int* cycles_done )
{
   if ( !ctx || minLevel < 0 || maxLevel < minLevel || pre < 0 || post < 0 || cycles < 0 || coarseMaxIter < 0 ||
        chebyshevOrder < 1 || chebyshevOrder > 16 || !( spectralRadius > 0 ) )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      using Op      = P2ElementwiseBlendingDivKGradOperator;
      auto storage  = storageOf( ctx, x );
      auto A        = Op::borrow( storage, op, uint_t( minLevel ), uint_t( maxLevel ) );
      auto X        = P2Function< double >::borrow( ctx, x );
      auto B        = P2Function< double >::borrow( ctx, b );
      auto smoother = std::make_shared< ChebyshevSmoother< Op > >( storage, uint_t( minLevel ), uint_t( maxLevel ) );
      smoother->setupCoefficients( uint_t( chebyshevOrder ), spectralRadius );
      std::shared_ptr< Solver< Op > > prec;
      if ( coarseJacobi )
         prec = std::make_shared< InverseDiagonalPreconditioner< Op > >();
      auto coarse = std::make_shared< CGSolver< Op > >( storage, uint_t( minLevel ), uint_t( minLevel ),
                                                        uint_t( coarseMaxIter ), coarseTol, 1e-300, prec );
      auto R      = std::make_shared< P2toP2QuadraticRestriction<> >();
      auto P      = std::make_shared< P2toP2QuadraticProlongation<> >();
      GeometricMultigridSolver< Op > gmg( storage, smoother, coarse, R, P, uint_t( minLevel ), uint_t( maxLevel ),
                                          uint_t( pre ), uint_t( post ) );
      P2Function< double > r( "residual", storage, uint_t( maxLevel ), uint_t( maxLevel ) );
      const int            k = cycle( *A, gmg, *X, *B, r, uint_t( maxLevel ), cycles, tol, res_hist );
      if ( cycles_done )
         *cycles_done = k;
   } );
}

int hg_p2_ir_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel, int lowDtype,
                    int pre, int post, int chebyshevOrder, double spectralRadius, int innerCycles, int maxOuter,
                    double tol, int coarseMaxIter, double coarseTol, double* res_hist, int* outer_done )
{
   if ( !ctx || minLevel < 0 || maxLevel < minLevel || pre < 0 || post < 0 || innerCycles < 1 || maxOuter < 0 ||
        coarseMaxIter < 0 || chebyshevOrder < 1 || chebyshevOrder > 16 || !( spectralRadius > 0 ) ||
        ( lowDtype != HG_F32 && lowDtype != HG_F16 ) )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      if ( lowDtype == HG_F32 )
         *outer_done = iterativeRefinement< float >( ctx, op, x, b, minLevel, maxLevel, pre, post, chebyshevOrder,
                                                     spectralRadius, innerCycles, maxOuter, tol, coarseMaxIter,
                                                     coarseTol, res_hist );
      else
         *outer_done = iterativeRefinement< f16 >( ctx, op, x, b, minLevel, maxLevel, pre, post, chebyshevOrder,
                                                   spectralRadius, innerCycles, maxOuter, tol, coarseMaxIter, coarseTol,
                                                   res_hist );
   } );
}

int hg_stokes_freeslip_apply( hg_ctx* ctx, hg_handle op, const hg_handle* src, const hg_handle* dst, int level,
                              int flag, int mask, int update )
{
   if ( !ctx || !src || !dst || level < 0 || ( mask != 1 && mask != 2 && mask != 4 && mask != 7 ) ||
        ( update != HG_REPLACE && update != HG_ADD ) )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      const hg_handle any     = ( mask == 2 ) ? src[3] : src[0];
      auto            storage = storageOf( ctx, any );
      withStokesOperator( storage, op, uint_t( level ), uint_t( level ), true, [&]( const auto& A, const auto& ) {
         const uint_t     L = uint_t( level );
         const DoFType    f = DoFType( flag );
         const UpdateType u = UpdateType( update );
         if ( mask == 7 )
            A.apply( TaylorHood( ctx, src ), TaylorHood( ctx, dst ), L, f, u );
         else if ( mask == 1 )
            A.getA().apply( P2VectorFunction< double >( ctx, src ), P2VectorFunction< double >( ctx, dst ), L, f, u );
         else if ( mask == 2 )
            A.applyDivT( *P1Function< double >::borrow( ctx, src[3] ), P2VectorFunction< double >( ctx, dst ), L, f, u );
         else
            A.applyDiv( P2VectorFunction< double >( ctx, src ), *P1Function< double >::borrow( ctx, dst[3] ), L, f, u );
      } );
   } );
}

int hg_stokes_estimate_radius( hg_ctx* ctx, hg_handle op, const hg_handle* x, const hg_handle* tmp, int level,
                               int iterations, int freeslip, double* radius )
{
   if ( !ctx || !x || !tmp || !radius || level < 0 || iterations < 1 )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      auto storage = storageOf( ctx, x[0] );
      withStokesOperator( storage, op, uint_t( level ), uint_t( level ), freeslip != 0, [&]( const auto& A, const auto& ) {
         P2VectorFunction< double > X( ctx, x ), T( ctx, tmp );
         *radius = chebyshev::estimateRadius( A.getA(), uint_t( level ), uint_t( iterations ), X, T );
      } );
   } );
}

int hg_stokes_uzawa_smooth( hg_ctx* ctx, hg_handle op, const hg_handle* x, const hg_handle* b, int level, double omega,
                            int velocityIterations, int chebyshevOrder, double spectralRadius, int projectPressureMean,
                            int freeslip )
{
   if ( !ctx || !x || !b || level < 0 || velocityIterations < 0 || chebyshevOrder < 1 || chebyshevOrder > 16 ||
        !( spectralRadius > 0 ) )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      auto storage = storageOf( ctx, x[0] );
      withStokesOperator( storage, op, uint_t( level ), uint_t( level ), freeslip != 0,
                          [&]( const auto& A, const auto& projection ) {
                             using Op = std::decay_t< decltype( A ) >;
                             TaylorHood X( ctx, x ), B( ctx, b );
                             auto       uzawa = stokesUzawa< Op >( storage, level, level, omega, velocityIterations,
                                                             chebyshevOrder, spectralRadius, projectPressureMean != 0,
                                                             projection );
                             uzawa->solve( A, X, B, uint_t( level ) );
                          } );
   } );
}

int hg_stokes_minres_solve( hg_ctx* ctx, hg_handle op, const hg_handle* x, const hg_handle* b, int level, int maxIter,
                            double relTol, double absTol, int freeslip, int* iterations )
{
   if ( !ctx || !x || !b || level < 0 || maxIter < 0 )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      auto storage = storageOf( ctx, x[0] );
      withStokesOperator( storage, op, uint_t( level ), uint_t( level ), freeslip != 0, [&]( const auto& A, const auto& ) {
         using Op = std::decay_t< decltype( A ) >;
         TaylorHood          X( ctx, x ), B( ctx, b );
         MinResSolver< Op > minres( storage, uint_t( level ), uint_t( level ), uint_t( maxIter ), relTol, absTol,
                                    std::make_shared< StokesPressureBlockPreconditioner< Op > >() );
         minres.solve( A, X, B, uint_t( level ) );
         if ( iterations )
            *iterations = int( minres.getIterations() );
      } );
   } );
}

int hg_stokes_gmg_solve( hg_ctx* ctx, hg_handle op, const hg_handle* x, const hg_handle* b, int minLevel, int maxLevel,
                         int pre, int post, int smoothIncrement, double omega, int velocityIterations, int chebyshevOrder,
                         double spectralRadius, int coarseMaxIter, double coarseTol, int projectPressureMean, int freeslip,
                         int cycles, double tol, double* res_hist, int* cycles_done )
{
   if ( !ctx || !x || !b || minLevel < 0 || maxLevel < minLevel || pre < 0 || post < 0 || smoothIncrement < 0 ||
        velocityIterations < 0 || chebyshevOrder < 1 || chebyshevOrder > 16 || !( spectralRadius > 0 ) ||
        coarseMaxIter < 0 || cycles < 0 )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      auto         storage = storageOf( ctx, x[0] );
      const uint_t minL = uint_t( minLevel ), L = uint_t( maxLevel );
      withStokesOperator( storage, op, minL, L, freeslip != 0, [&]( const auto& A, const auto& projection ) {
         using Op = std::decay_t< decltype( A ) >;
         TaylorHood X( ctx, x ), B( ctx, b );
         auto       uzawa  = stokesUzawa< Op >( storage, minLevel, maxLevel, omega, velocityIterations, chebyshevOrder,
                                         spectralRadius, projectPressureMean != 0, projection );
         auto       coarse = std::make_shared< MinResSolver< Op > >(
             storage, minL, minL, uint_t( coarseMaxIter ), coarseTol, 0.0,
             std::make_shared< StokesPressureBlockPreconditioner< Op > >() );
         std::shared_ptr< RestrictionOperator< TaylorHood > >  restriction;
         std::shared_ptr< ProlongationOperator< TaylorHood > > prolongation;
         if ( projection )
         {
            restriction  = std::make_shared< P2P1StokesToP2P1StokesRestrictionWithFreeSlipProjection >( projection );
            prolongation = std::make_shared< P2P1StokesToP2P1StokesProlongationWithFreeSlipProjection >(
                std::make_shared< TaylorHood >( "tmpProlongation", storage, minL, L ), projection );
         }
         else
         {
            restriction  = std::make_shared< P2P1StokesToP2P1StokesRestriction >();
            prolongation = std::make_shared< P2P1StokesToP2P1StokesProlongation >();
         }
         GeometricMultigridSolver< Op > gmg( storage, uzawa, coarse, restriction, prolongation, minL, L, uint_t( pre ),
                                             uint_t( post ), uint_t( smoothIncrement ) );
         TaylorHood   r( "residual", storage, L, L );
         const double r0 = stokesResidualNorm( A, X, B, r, L );
         if ( res_hist )
            res_hist[0] = r0;
         int k = 0;
         for ( double rk = r0; k < cycles && !( tol > 0 && rk <= tol * r0 ); )
         {
            gmg.solve( A, X, B, L );
            if ( projectPressureMean )
               vertexdof::projectMean( X.p(), L );
            rk = stokesResidualNorm( A, X, B, r, L );
            ++k;
            if ( res_hist )
               res_hist[k] = rk;
         }
         if ( cycles_done )
            *cycles_done = k;
      } );
   } );
}

} // extern "C"

namespace {
// the velocity-block solver of the block-factorisation preconditioner: one GMG
// V-cycle (Chebyshev on [0.1, 1.2] x radius, the TerraNeo interval factors of
// terraneo/helpers/TerraNeoParameters.hpp:82-83; MINRES with identity
// preconditioner on minLevel, in FP64).  With free-slip the transfers project
// their result: the coarse P A P is singular in the normal directions, and an
// unprojected coarse right-hand side makes the V-cycle diverge.
template < class VT, class VOp64, class VOpT >
std::shared_ptr< Solver< VOpT > > velocityVCycle( const std::shared_ptr< PrimitiveStorage >& storage, uint_t minL,
                                                 uint_t L, const VOp64& A64, const VOpT& A, int aPre, int aPost,
                                                 int order, double rho, int coarseIt, double coarseTol,
                                                 const std::shared_ptr< P2ProjectNormalOperator >& projection,
                                                 std::vector< std::shared_ptr< void > >& keep )
{
   auto cheby = std::make_shared< ChebyshevSmoother< VOpT > >( storage, minL, L );
   cheby->setupCoefficients( uint_t( order ), rho, 1.2, 0.1 );
   auto minres = std::make_shared< MinResSolver< VOp64 > >( storage, minL, minL, uint_t( coarseIt ), coarseTol, 0.0,
                                                            std::make_shared< IdentityPreconditioner< VOp64 > >() );
   std::shared_ptr< Solver< VOpT > > coarse;
   if constexpr ( std::is_same< VOpT, VOp64 >::value )
      coarse = minres;
   else
      coarse = std::make_shared< FP64CoarseSolver< VOpT, VOp64 > >( storage, minL, A64, minres );
   keep.push_back( minres );
   std::shared_ptr< RestrictionOperator< P2VectorFunction< VT > > >  restriction;
   std::shared_ptr< ProlongationOperator< P2VectorFunction< VT > > > prolongation;
   if ( projection )
   {
      restriction  = std::make_shared< P2toP2QuadraticVectorRestrictionWithProjectionT< VT > >( projection );
      prolongation = std::make_shared< P2toP2QuadraticVectorProlongationWithProjectionT< VT > >( projection );
   }
   else
   {
      restriction  = std::make_shared< P2toP2QuadraticVectorRestrictionT< VT > >();
      prolongation = std::make_shared< P2toP2QuadraticVectorProlongationT< VT > >();
   }
   (void) A;
   return std::make_shared< GeometricMultigridSolver< VOpT > >( storage, cheby, coarse, restriction, prolongation, minL, L,
                                                                uint_t( aPre ), uint_t( aPost ) );
}
} // namespace

extern "C" {

int hg_p2_const_gmg_solve( hg_ctx* ctx, hg_handle op, hg_handle x, hg_handle b, int minLevel, int maxLevel, int pre,
                           int post, int coarseMaxIter, double coarseTol, int cycles, double tol, double* res_hist,
                           int* cycles_done )
{
   if ( !ctx || minLevel < 0 || maxLevel < minLevel || pre < 0 || post < 0 || coarseMaxIter < 0 || cycles < 0 )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      using Op     = P2ConstantLaplaceOperator;
      auto storage = storageOf( ctx, x );
      auto A       = Op::borrow( storage, op, uint_t( minLevel ), uint_t( maxLevel ) );
      auto X       = P2Function< double >::borrow( ctx, x );
      auto B       = P2Function< double >::borrow( ctx, b );
      auto coarse  = std::make_shared< CGSolver< Op > >( storage, uint_t( minLevel ), uint_t( minLevel ),
                                                        uint_t( coarseMaxIter ), coarseTol, 1e-16 );
      GeometricMultigridSolver< Op > gmg( storage, std::make_shared< GaussSeidelSmoother< Op > >(), coarse,
                                          std::make_shared< P2toP2QuadraticRestriction<> >(),
                                          std::make_shared< P2toP2QuadraticProlongation<> >(), uint_t( minLevel ),
                                          uint_t( maxLevel ), uint_t( pre ), uint_t( post ) );
      P2Function< double > r( "residual", storage, uint_t( maxLevel ), uint_t( maxLevel ) );
      const int done = cycle( *A, gmg, *X, *B, r, uint_t( maxLevel ), cycles, tol, res_hist );
      if ( cycles_done )
         *cycles_done = done;
   } );
}

int hg_stokes_fgmres_solve( hg_ctx* ctx, hg_handle op, const hg_handle* x, const hg_handle* b, int minLevel,
                            int maxLevel, int aPre, int aPost, int chebyshevOrder, double spectralRadius,
                            int aCoarseMaxIter, double aCoarseTol, int schurMaxIter, double schurTol, int freeslip,
                            int velocityDtype, int maxIter, int restart, double relTol, double* res_hist,
                            int* iterations )
{
   if ( !ctx || !x || !b || minLevel < 0 || maxLevel < minLevel || aPre < 0 || aPost < 0 || chebyshevOrder < 1 ||
        chebyshevOrder > 16 || !( spectralRadius > 0 ) || aCoarseMaxIter < 0 || schurMaxIter < 0 || maxIter < 1 ||
        restart < 2 || velocityDtype < HG_F64 || velocityDtype > HG_F16 )
      return HG_ERR_ARG;
   return guarded( ctx, [&] {
      auto         storage = storageOf( ctx, x[0] );
      const uint_t minL = uint_t( minLevel ), L = uint_t( maxLevel );
      withStokesOperator( storage, op, minL, L, freeslip != 0, [&]( const auto& A, const auto& projection ) {
         using Op  = std::decay_t< decltype( A ) >;
         using VOp = typename Op::VelocityOperator_T;
         using SOp = P1ElementwiseKMassOperator;
         TaylorHood                               X( ctx, x ), B( ctx, b );
         std::vector< std::shared_ptr< void > > keep;
         // velocity block: FP64, or FP32 arithmetic on FP32 / FP16 vectors inside the FP64 iteration
         std::shared_ptr< Solver< VOp > > ablock;
         auto lowPrecision = [&]( auto tag ) {
            using VT   = decltype( tag );
            using VOpT = P2ElementwiseBlendingFullViscousOperatorT< VT >;
            auto Alow  = std::make_shared< VOpT >( storage, op, minL, L, projection );
            keep.push_back( Alow );
            auto gmg = velocityVCycle< VT, VOp, VOpT >( storage, minL, L, A.getA(), *Alow, aPre, aPost, chebyshevOrder,
                                                        spectralRadius, aCoarseMaxIter, aCoarseTol, projection, keep );
            ablock = std::make_shared< MixedPrecisionSolver< VOp, VOpT > >( storage, minL, L, *Alow, gmg, projection );
         };
         if ( velocityDtype == HG_F64 )
            ablock = velocityVCycle< double, VOp, VOp >( storage, minL, L, A.getA(), A.getA(), aPre, aPost,
                                                         chebyshevOrder, spectralRadius, aCoarseMaxIter, aCoarseTol,
                                                         projection, keep );
         else if ( velocityDtype == HG_F32 )
            lowPrecision( float() );
         else
            lowPrecision( f16() );
         // Schur block: CG on the 1/eta pressure mass, preconditioned with its inverse lumped diagonal
         SOp  S( ctx, op );
         auto cgS = std::make_shared< CGSolver< SOp > >(
             storage, L, L, uint_t( schurMaxIter ), schurTol, 1e-16,
             std::make_shared< FunctionMultiplicationPreconditioner< SOp > >( S.inverseLumpedDiagonal( storage, L, L ),
                                                                              All ) );
         // the Krylov vectors carry the whole hierarchy: the velocity V-cycle runs on them
         auto prec = std::make_shared< BlockFactorisationPreconditioner< Op, VOp, SOp > >( storage, minL, L, S, ablock, cgS,
                                                                                           1, projection );
         FGMRESSolver< Op > fgmres( storage, minL, L, uint_t( maxIter ) + 1, relTol, 0.0, prec, uint_t( restart ) );
         fgmres.setIterationHook( [&]( uint_t j, double r ) {
            if ( res_hist )
               res_hist[j] = r;
            return false;
         } );
         fgmres.solve( A, X, B, L );
         if ( res_hist )
            res_hist[0] = fgmres.getInitialResidual();
         if ( iterations )
            *iterations = int( fgmres.getIterations() );
      } );
   } );
}

} // extern "C"
