/* CPU oracle, test infrastructure only: multigrid components of the P2 path.
 *
 *   quadratic prolongation    P2toP2QuadraticProlongation.cpp:37-63,217-425
 *   quadratic restriction     P2toP2QuadraticRestriction.cpp:131-290 (transpose)
 *   Chebyshev smoother        ChebyshevSmoother.hpp:112-232 (solve), :246-330
 *                             (setupCoefficients / setupCoefficientsInternal)
 *   power iteration           SpectrumEstimation.hpp:56-90, ChebyshevSmoother.hpp:560-600,656-665
 *   (preconditioned) CG       CGSolver.hpp:94-203
 *   V-cycle                   GeometricMultigridSolver.hpp:180-308
 *
 * Prolongation is the interpolation of the coarse P2 function at the fine P2
 * nodes.  It is evaluated element by element: every coarse micro-cell maps
 * its 35 level+2 lattice points through the 10 quadratic basis functions; a
 * fine node shared by several micro-cells takes its value from the first one
 * (all give the same value).  Restriction is the exact transpose of that
 * matrix.  The reference reaches the same operators with per-primitive push
 * kernels scaled by 1/(number of neighbour cells) and additive communication.
 *
 * The Chebyshev coefficients are derived from the defining property of the
 * smoother polynomial (Adams, Brezina, Hu, Tuminaro 2003):
 *   1 - lambda p(lambda) = T_k((theta - lambda)/delta) / T_k(theta/delta),
 * theta = (upper + lower)/2, delta = (upper - lower)/2, expanded in monomials. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hg_oracle.h"
#include "hg_oracle_internal.h"

static int64_t tetidx( int64_t w, int64_t x, int64_t y, int64_t z )
{
   const int64_t tw = w * ( w + 1 ) * ( w + 2 ) / 6, tz = ( w - z ) * ( w - z + 1 ) * ( w - z + 2 ) / 6;
   return tw - tz + y * ( w - z ) - y * ( y - 1 ) / 2 + x;
}

static const int P2EDGE[6][2] = { { 0, 1 }, { 0, 2 }, { 0, 3 }, { 1, 2 }, { 1, 3 }, { 2, 3 } };

typedef struct
{
   int64_t fine;    /* canonical fine id */
   int64_t c[10];   /* canonical coarse ids of the element's 10 nodes */
   double  phi[10]; /* basis values at the fine node */
} prow_t;

/* one row of P per fine P2 node of level coarseLevel+1 (rows in canonical fine order) */
static prow_t* build_prolongation( const or_mesh* m, int coarseLevel, int64_t* nrows )
{
   const int     Lc1 = coarseLevel + 1, Lf1 = coarseLevel + 2; /* storage lattices */
   lvl_t         lc = or_lvl_info( m, Lc1 ), lf = or_lvl_info( m, Lf1 );
   int64_t       scc, scf;
   int64_t*      cmc = or_build_cmap( m, Lc1, &scc );
   int64_t*      cmf = or_build_cmap( m, Lf1, &scf );
   const int64_t wc = lc.mm + 1, wf = lf.mm + 1, mm = (int64_t) 1 << coarseLevel;
   prow_t*       rows = calloc( lf.n, sizeof( prow_t ) );
   char*         seen = calloc( lf.n, 1 );
   for ( int c = 0; c < m->nc; ++c )
      for ( int t = 0; t < 6; ++t )
      {
         const int64_t wt = mm - WSHIFT[t];
         for ( int64_t z = 0; z < wt; ++z )
            for ( int64_t y = 0; y + z < wt; ++y )
               for ( int64_t x = 0; x + y + z < wt; ++x )
               {
                  int64_t v[4][3], ids[10];
                  for ( int k = 0; k < 4; ++k )
                  {
                     v[k][0] = x + VOF[t][k][0];
                     v[k][1] = y + VOF[t][k][1];
                     v[k][2] = z + VOF[t][k][2];
                     ids[k]  = cmc[c * scc + tetidx( wc, 2 * v[k][0], 2 * v[k][1], 2 * v[k][2] )];
                  }
                  for ( int e = 0; e < 6; ++e )
                  {
                     const int64_t* a = v[P2EDGE[e][0]];
                     const int64_t* b = v[P2EDGE[e][1]];
                     ids[4 + e]       = cmc[c * scc + tetidx( wc, a[0] + b[0], a[1] + b[1], a[2] + b[2] )];
                  }
                  /* the 35 points sum_k n_k v_k / 4, n_k >= 0, sum n_k = 4, in level+2 lattice units */
                  for ( int n1 = 0; n1 <= 4; ++n1 )
                     for ( int n2 = 0; n1 + n2 <= 4; ++n2 )
                        for ( int n3 = 0; n1 + n2 + n3 <= 4; ++n3 )
                        {
                           const int n0 = 4 - n1 - n2 - n3;
                           const int n[4] = { n0, n1, n2, n3 };
                           int64_t   X[3] = { 0, 0, 0 };
                           for ( int k = 0; k < 4; ++k )
                              for ( int d = 0; d < 3; ++d )
                                 X[d] += n[k] * v[k][d];
                           const int64_t f = cmf[c * scf + tetidx( wf, X[0], X[1], X[2] )];
                           if ( seen[f] )
                              continue;
                           seen[f]     = 1;
                           prow_t* r   = rows + f;
                           double  lam[4];
                           for ( int k = 0; k < 4; ++k )
                              lam[k] = 0.25 * n[k];
                           r->fine = f;
                           for ( int k = 0; k < 4; ++k )
                           {
                              r->c[k]   = ids[k];
                              r->phi[k] = lam[k] * ( 2.0 * lam[k] - 1.0 );
                           }
                           for ( int e = 0; e < 6; ++e )
                           {
                              r->c[4 + e]   = ids[4 + e];
                              r->phi[4 + e] = 4.0 * lam[P2EDGE[e][0]] * lam[P2EDGE[e][1]];
                           }
                        }
               }
      }
   free( seen );
   free( cmc );
   free( cmf );
   *nrows = lf.n;
   return rows;
}

void or_p2_prolongate( const or_mesh* m, int coarseLevel, const double* coarse, double* fine, int flag, int add )
{
   int64_t n;
   prow_t* P  = build_prolongation( m, coarseLevel, &n );
   lvl_t   lf = or_lvl_info( m, coarseLevel + 2 );
   for ( int64_t f = 0; f < n; ++f )
   {
      if ( f < lf.coff && !( or_owner_flag( m, &lf, f ) & flag ) )
         continue;
      double v = 0.0;
      for ( int i = 0; i < 10; ++i )
         if ( P[f].phi[i] != 0.0 )
            v += P[f].phi[i] * coarse[P[f].c[i]];
      fine[f] = add ? fine[f] + v : v;
   }
   free( P );
}

void or_p2_restrict( const or_mesh* m, int fineLevel, const double* fine, double* coarse, int flag )
{
   int64_t n;
   prow_t* P  = build_prolongation( m, fineLevel - 1, &n );
   lvl_t   lc = or_lvl_info( m, fineLevel );
   double* acc = calloc( lc.n, sizeof( double ) );
   for ( int64_t f = 0; f < n; ++f )
      for ( int i = 0; i < 10; ++i )
         if ( P[f].phi[i] != 0.0 )
            acc[P[f].c[i]] += P[f].phi[i] * fine[f];
   for ( int64_t p = 0; p < lc.n; ++p )
      if ( p >= lc.coff || ( or_owner_flag( m, &lc, p ) & flag ) )
         coarse[p] = acc[p];
   free( acc );
   free( P );
}

/* ---- Chebyshev ------------------------------------------------------------ */
int or_chebyshev_coefficients( int order, double lower, double upper, double* coeffs )
{
   if ( order < 1 || order > 16 )
      return -1;
   const double theta = 0.5 * ( upper + lower ), delta = 0.5 * ( upper - lower );
   /* T_j(s) with s = (theta - lambda) / delta = a + b lambda, as polynomials in lambda */
   const double a = theta / delta, b = -1.0 / delta;
   double       Tm[18] = { 0 }, T0[18] = { 0 }, T1[18] = { 0 };
   T0[0] = 1.0;
   T1[0] = a;
   T1[1] = b;
   double* prev = T0;
   double* cur  = T1;
   for ( int j = 1; j < order; ++j )
   {
      /* T_{j+1} = 2 s T_j - T_{j-1} */
      for ( int i = 0; i <= j + 1; ++i )
      {
         const double si = 2.0 * ( a * cur[i] + ( i > 0 ? b * cur[i - 1] : 0.0 ) ) - prev[i];
         Tm[i]           = si;
      }
      for ( int i = 0; i <= j + 1; ++i )
      {
         prev[i] = cur[i];
         cur[i]  = Tm[i];
      }
   }
   /* T_k(theta/delta) = T_k(s at lambda = 0) = cur[0]; p_i = -cur[i+1] / cur[0] */
   for ( int i = 0; i < order; ++i )
      coeffs[i] = -cur[i + 1] / cur[0];
   return 0;
}

static void mult_f( int64_t n, const double* a, const double* x, double* y, const int32_t* fl, int flag )
{
   for ( int64_t i = 0; i < n; ++i )
      if ( fl[i] & flag )
         y[i] = a[i] * x[i];
}

void or_p2_chebyshev_smooth( const or_p2op* op, const double* invDiag, int order, const double* coeffs, double* x,
                             const double* b, int flag )
{
   const int64_t  n  = op->l1.n;
   const int32_t* fl = op->flags;
   double*        t1 = calloc( n, sizeof( double ) );
   double*        t2 = calloc( n, sizeof( double ) );
   or_p2_apply( op, x, t2, flag, 0 );
   or_axpby_f( n, 1.0, b, -1.0, t2, fl, flag );
   mult_f( n, invDiag, t2, t1, fl, flag );
   or_axpby_f( n, coeffs[0], t1, 1.0, x, fl, flag );
   for ( int k = 1; k < order; ++k )
   {
      or_p2_apply( op, t1, t2, flag, 0 );
      memset( t1, 0, sizeof( double ) * n );
      mult_f( n, invDiag, t2, t1, fl, flag );
      or_axpby_f( n, coeffs[k], t1, 1.0, x, fl, flag );
   }
   free( t1 );
   free( t2 );
}

double or_p2_power_iteration( const or_p2op* op, const double* invDiag, double* x, int its )
{
   const int64_t  n   = op->l1.n;
   const int32_t* fl  = op->flags;
   double*        aux = calloc( n, sizeof( double ) );
   double         nrm = sqrt( or_dot( n, x, x, fl, OR_ALL ) );
   for ( int64_t i = 0; i < n; ++i )
      x[i] /= nrm;
   or_p2_apply( op, x, aux, OR_ALL, 0 );
   mult_f( n, invDiag, aux, aux, fl, OR_ALL );
   double radius = 0.0;
   for ( int it = 1; it <= its; ++it )
   {
      nrm = sqrt( or_dot( n, aux, aux, fl, OR_ALL ) );
      for ( int64_t i = 0; i < n; ++i )
         x[i] = aux[i] / nrm;
      or_p2_apply( op, x, aux, OR_ALL, 0 );
      mult_f( n, invDiag, aux, aux, fl, OR_ALL );
      radius = or_dot( n, x, aux, fl, OR_ALL );
   }
   free( aux );
   return radius;
}

/* CGSolver::solve with the identity (jacobi = 0) or the inverse-diagonal
 * (jacobi = 1) preconditioner; returns the iteration index at exit */
int or_p2_cg( const or_p2op* op, double* x, const double* b, int flag, int maxIter, double relTol, double absTol,
              int jacobi )
{
   const int64_t  n  = op->l1.n;
   const int32_t* fl = op->flags;
   double*        p  = calloc( n, sizeof( double ) );
   double*        z  = calloc( n, sizeof( double ) );
   double*        ap = calloc( n, sizeof( double ) );
   double*        r  = calloc( n, sizeof( double ) );
   double*        id = NULL;
   if ( jacobi )
   {
      id = malloc( sizeof( double ) * n );
      or_p2_inv_diag( op, id );
   }
   or_p2_apply( op, x, ap, flag, 0 );
   or_axpby_f( n, 1.0, b, 0.0, r, fl, flag );
   or_axpby_f( n, -1.0, ap, 1.0, r, fl, flag );
   if ( jacobi )
      mult_f( n, id, r, z, fl, flag );
   else
      or_axpby_f( n, 1.0, r, 0.0, z, fl, flag );
   or_axpby_f( n, 1.0, z, 0.0, p, fl, flag );
   double       prsold    = or_dot( n, r, z, fl, flag );
   const double res_start = sqrt( or_dot( n, r, r, fl, flag ) );
   int          it        = 0;
   if ( res_start >= absTol )
      for ( int i = 0; i < maxIter; ++i )
      {
         or_p2_apply( op, p, ap, flag, 0 );
         const double pAp   = or_dot( n, p, ap, fl, flag );
         const double alpha = prsold / pAp;
         or_axpby_f( n, alpha, p, 1.0, x, fl, flag );
         or_axpby_f( n, -alpha, ap, 1.0, r, fl, flag );
         const double rs = sqrt( or_dot( n, r, r, fl, flag ) );
         it              = i;
         if ( rs / res_start < relTol || rs < absTol )
            break;
         memset( z, 0, sizeof( double ) * n );
         if ( jacobi )
            mult_f( n, id, r, z, fl, flag );
         else
            or_axpby_f( n, 1.0, r, 0.0, z, fl, flag );
         const double prsnew = or_dot( n, r, z, fl, flag );
         or_axpby_f( n, 1.0, z, prsnew / prsold, p, fl, flag );
         prsold = prsnew;
      }
   free( p );
   free( z );
   free( ap );
   free( r );
   free( id );
   return it;
}

/* ---- V-cycle ---------------------------------------------------------------- */
typedef struct
{
   or_p2op* op;
   int64_t  n;
   double * x, *b, *tmp, *invd;
} p2lev_t;

typedef struct
{
   const or_mesh* m;
   p2lev_t*       lv;
   int            minL, pre, post, order, coarseMaxIter, coarseJacobi;
   double         coeffs[16], coarseTol;
} p2gmg_t;

static void p2_vcycle( p2gmg_t* g, int L )
{
   const int flag = OR_INNER | OR_NEUMANN | OR_FREESLIP;
   p2lev_t*  c    = &g->lv[L];
   if ( L == g->minL )
   {
      or_p2_cg( c->op, c->x, c->b, flag, g->coarseMaxIter, g->coarseTol, 1e-300, g->coarseJacobi );
      return;
   }
   for ( int i = 0; i < g->pre; ++i )
      or_p2_chebyshev_smooth( c->op, c->invd, g->order, g->coeffs, c->x, c->b, flag );
   or_p2_apply( c->op, c->x, c->tmp, flag, 0 );
   or_axpby_f( c->n, 1.0, c->b, -1.0, c->tmp, c->op->flags, flag );
   p2lev_t* d = &g->lv[L - 1];
   or_p2_restrict( g->m, L, c->tmp, d->tmp, flag );
   or_axpby_f( d->n, 1.0, d->tmp, 0.0, d->b, d->op->flags, flag );
   memset( d->x, 0, sizeof( double ) * d->n );
   p2_vcycle( g, L - 1 );
   or_p2_prolongate( g->m, L - 1, d->x, c->x, flag, 1 );
   for ( int i = 0; i < g->post; ++i )
      or_p2_chebyshev_smooth( c->op, c->invd, g->order, g->coeffs, c->x, c->b, flag );
}

int or_p2_gmg( const or_mesh* m, int minLevel, int maxLevel, int nm, const double* ka, const double* kom,
               const double* kph, double* x, const double* b, int pre, int post, int order, double rho, int cycles,
               double tol, double* res_hist, int coarseMaxIter, double coarseTol, int coarseJacobi )
{
   p2gmg_t g = { 0 };
   g.m       = m;
   g.minL    = minLevel;
   g.pre     = pre;
   g.post    = post;
   g.order   = order;
   g.coarseMaxIter = coarseMaxIter;
   g.coarseTol     = coarseTol;
   g.coarseJacobi  = coarseJacobi;
   if ( or_chebyshev_coefficients( order, 0.3 * rho, 1.2 * rho, g.coeffs ) )
      return -1;
   g.lv = calloc( maxLevel + 1, sizeof( p2lev_t ) );
   for ( int L = minLevel; L <= maxLevel; ++L )
   {
      p2lev_t* l = &g.lv[L];
      l->op      = or_p2op_create( m, L, nm, ka, kom, kph );
      l->n       = l->op->l1.n;
      l->x       = L == maxLevel ? x : calloc( l->n, sizeof( double ) );
      l->b       = calloc( l->n, sizeof( double ) );
      l->tmp     = calloc( l->n, sizeof( double ) );
      l->invd    = malloc( sizeof( double ) * l->n );
      or_p2_inv_diag( l->op, l->invd );
   }
   p2lev_t* f = &g.lv[maxLevel];
   memcpy( f->b, b, sizeof( double ) * f->n );
   double* r = calloc( f->n, sizeof( double ) );
#define P2RESID()                                                                   \
   ( or_p2_apply( f->op, f->x, r, OR_INNER, 0 ),                                    \
     or_axpby_f( f->n, 1.0, f->b, -1.0, r, f->op->flags, OR_INNER ),                 \
     sqrt( or_dot( f->n, r, r, f->op->flags, OR_INNER ) ) )
   const double r0 = P2RESID();
   if ( res_hist )
      res_hist[0] = r0;
   int k = 0;
   while ( k < cycles )
   {
      p2_vcycle( &g, maxLevel );
      ++k;
      const double rk = P2RESID();
      if ( res_hist )
         res_hist[k] = rk;
      if ( rk <= tol * r0 )
         break;
   }
#undef P2RESID
   free( r );
   for ( int L = minLevel; L <= maxLevel; ++L )
   {
      p2lev_t* l = &g.lv[L];
      or_p2op_free( l->op );
      if ( L != maxLevel )
         free( l->x );
      free( l->b );
      free( l->tmp );
      free( l->invd );
   }
   free( g.lv );
   return k;
}
