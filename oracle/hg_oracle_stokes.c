/*
 * hg_oracle_stokes.c -- CPU restatement of the P2-P1 Taylor-Hood Stokes
 * operator of HyTeG with blending and variable viscosity, and of the P1
 * diagonal operators the Uzawa smoother and the MINRES coarse solver use.
 *
 * TEST INFRASTRUCTURE ONLY (see hg_oracle.h).
 *
 * Reference files followed:
 *   elementwiseoperators/P2P1ElementwiseBlendingStokesOperator.hpp:125-216
 *       (P2P1ElementwiseBlendingFullViscousStokesOperator: apply = A u + B^T p, B u)
 *   elementwiseoperators/P2ElementwiseBlendingFullViscousOperator.hpp:37-88 (9 blocks)
 *   forms/form_hyteg_generated/p2/p2_full_stokesvar_blending_q3.cpp:723-1580
 *       (2 eta eps(u):eps(v) - 2/3 eta div u div v, 5-point degree-3 rule with
 *        weights -2/15 and 3/40, eta at the blended points)
 *   forms/form_hyteg_generated/p2_to_p1/p2_to_p1_div_blending_q2.cpp:245-1060,
 *   forms/form_hyteg_generated/p1_to_p2/p1_to_p2_divt_blending_q2.cpp
 *       (B = -(psi, div u), 4-point degree-2 rule, weights 1/24)
 *   elementwiseoperators/P2ToP1ElementwiseOperator.cpp:108-290 (gemv data flow)
 *   forms/form_fenics_generated/p1_tet_pspg_tet.h:4113-4230
 *       (PSPG: -(|T|^(2/3) / 12) (grad p, grad q), affine element geometry)
 *   constant_stencil_operator/P1ConstantOperator.hpp:195-215
 *       (P1PSPGInvDiagOperator, P1LumpedInvMassOperator)
 *
 * Data layout: a velocity vector is [u_x | u_y | u_z], each block the canonical
 * P1 numbering of level L+1 (P2 level L); the pressure is the canonical P1
 * numbering of level L.  The pressure has no boundary DoFs in HyTeG's sense
 * (P2P1TaylorHoodFunction gives it createAllInnerBC, P2P1TaylorHoodFunction.hpp:61):
 * every pressure DoF counts as Inner.
 */
#include "hg_oracle_internal.h"

#include <stdio.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* 5-point degree-3 rule (negative centroid weight): centroid, then the points
 * with barycentric coordinate 1/2 at vertex 3, 2, 1, 0 */
static const double Q3P[5][4] = { { 0.25, 0.25, 0.25, 0.25 },
                                  { 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 0.5 },
                                  { 1.0 / 6.0, 1.0 / 6.0, 0.5, 1.0 / 6.0 },
                                  { 1.0 / 6.0, 0.5, 1.0 / 6.0, 1.0 / 6.0 },
                                  { 0.5, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0 } };
static const double Q3W[5]    = { -2.0 / 15.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0 };
/* 4-point degree-2 rule: (a, a, a, b) permutations, b at vertex 3, 2, 1, 0 */
#define QA 0.1381966011250105
#define QB 0.5854101966249685
static const double Q2P[4][4] = { { QA, QA, QA, QB }, { QA, QA, QB, QA }, { QA, QB, QA, QA }, { QB, QA, QA, QA } };
static const double Q2W       = 1.0 / 24.0;

static const int SE[6][2] = { { 0, 1 }, { 0, 2 }, { 0, 3 }, { 1, 2 }, { 1, 3 }, { 2, 3 } };

static void ref_grads( const double* lam, double g[10][3] )
{
   static const double dl[4][3] = { { -1, -1, -1 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
   for ( int i = 0; i < 4; ++i )
      for ( int d = 0; d < 3; ++d )
         g[i][d] = ( 4.0 * lam[i] - 1.0 ) * dl[i][d];
   for ( int e = 0; e < 6; ++e )
   {
      const int a = SE[e][0], b = SE[e][1];
      for ( int d = 0; d < 3; ++d )
         g[4 + e][d] = 4.0 * ( lam[a] * dl[b][d] + lam[b] * dl[a][d] );
   }
}

static void inverse3( const double* A, double* Ai, double* det )
{
   const double d = A[0] * ( A[4] * A[8] - A[5] * A[7] ) - A[1] * ( A[3] * A[8] - A[5] * A[6] ) +
                    A[2] * ( A[3] * A[7] - A[4] * A[6] );
   Ai[0] = ( A[4] * A[8] - A[5] * A[7] ) / d;
   Ai[1] = ( A[2] * A[7] - A[1] * A[8] ) / d;
   Ai[2] = ( A[1] * A[5] - A[2] * A[4] ) / d;
   Ai[3] = ( A[5] * A[6] - A[3] * A[8] ) / d;
   Ai[4] = ( A[0] * A[8] - A[2] * A[6] ) / d;
   Ai[5] = ( A[2] * A[3] - A[0] * A[5] ) / d;
   Ai[6] = ( A[3] * A[7] - A[4] * A[6] ) / d;
   Ai[7] = ( A[1] * A[6] - A[0] * A[7] ) / d;
   Ai[8] = ( A[0] * A[4] - A[1] * A[3] ) / d;
   *det  = d;
}

/* physical gradients of the 10 P2 basis functions, |det (DF J)| and the blended
 * point at barycentric point lam of the element with computational vertices p */
static void point_geometry( const double* p, const double* blend, const double* lam, double pg[10][3], double* adet,
                            double F[3] )
{
   double J[9], x[3], DF[9] = { 1, 0, 0, 0, 1, 0, 0, 0, 1 };
   for ( int r = 0; r < 3; ++r )
      for ( int k = 0; k < 3; ++k )
         J[3 * r + k] = p[3 * ( k + 1 ) + r] - p[r];
   for ( int d = 0; d < 3; ++d )
      x[d] = lam[0] * p[d] + lam[1] * p[3 + d] + lam[2] * p[6 + d] + lam[3] * p[9 + d];
   if ( blend )
      or_shell_map( blend, x, F, DF );
   else
      memcpy( F, x, 3 * sizeof( double ) );
   double M[9], Mi[9], detM;
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j )
         M[3 * i + j] = DF[3 * i] * J[j] + DF[3 * i + 1] * J[3 + j] + DF[3 * i + 2] * J[6 + j];
   inverse3( M, Mi, &detM );
   double g[10][3];
   ref_grads( lam, g );
   for ( int i = 0; i < 10; ++i )
      for ( int d = 0; d < 3; ++d )
         pg[i][d] = Mi[d] * g[i][0] + Mi[3 + d] * g[i][1] + Mi[6 + d] * g[i][2];
   *adet = fabs( detM );
}

/* eta(x) = exp(c0 + c1 |F(x)|) */
static double viscosity( double c0, double c1, const double F[3] )
{
   return exp( c0 + c1 * sqrt( F[0] * F[0] + F[1] * F[1] + F[2] * F[2] ) );
}

/* A (30x30, row (c,i) = 10 c + i: test component c, test function i) of
 * int 2 eta eps(u):eps(v) - 2/3 eta div u div v, and B (4x30) of -(psi_i, div u) */
void or_stokes_element( const double* p, const double* blend, double c0, double c1, double* A, double* B )
{
   if ( A )
   {
      memset( A, 0, 900 * sizeof( double ) );
      for ( int q = 0; q < 5; ++q )
      {
         double pg[10][3], adet, F[3];
         point_geometry( p, blend, Q3P[q], pg, &adet, F );
         const double wq = Q3W[q] * adet * viscosity( c0, c1, F );
         for ( int ci = 0; ci < 3; ++ci )
            for ( int i = 0; i < 10; ++i )
               for ( int cj = 0; cj < 3; ++cj )
                  for ( int j = 0; j < 10; ++j )
                  {
                     double v = pg[j][ci] * pg[i][cj] - ( 2.0 / 3.0 ) * pg[j][cj] * pg[i][ci];
                     if ( ci == cj )
                        v += pg[j][0] * pg[i][0] + pg[j][1] * pg[i][1] + pg[j][2] * pg[i][2];
                     A[( 10 * ci + i ) * 30 + 10 * cj + j] += wq * v;
                  }
      }
   }
   if ( B )
   {
      memset( B, 0, 120 * sizeof( double ) );
      for ( int q = 0; q < 4; ++q )
      {
         double pg[10][3], adet, F[3];
         point_geometry( p, blend, Q2P[q], pg, &adet, F );
         const double wq = Q2W * adet;
         for ( int i = 0; i < 4; ++i )
            for ( int cj = 0; cj < 3; ++cj )
               for ( int j = 0; j < 10; ++j )
                  B[i * 30 + 10 * cj + j] -= wq * Q2P[q][i] * pg[j][cj];
      }
   }
}

/* M (4x4) of int k psi_i psi_j with k = 1/eta = exp(-(c0 + c1 |F(x)|)), the
 * 4-point degree-2 rule on the blended element (the P1 "K-mass" Schur operator
 * of the free-slip Stokes wrapper, terraneo/operators/P2P1StokesOperatorWithWrapper.hpp:388-395) */
void or_stokes_kmass_element( const double* p, const double* blend, double c0, double c1, double* M )
{
   memset( M, 0, 16 * sizeof( double ) );
   for ( int q = 0; q < 4; ++q )
   {
      double pg[10][3], adet, F[3];
      point_geometry( p, blend, Q2P[q], pg, &adet, F );
      const double wq = Q2W * adet / viscosity( c0, c1, F );
      for ( int i = 0; i < 4; ++i )
         for ( int j = 0; j < 4; ++j )
            M[4 * i + j] += wq * Q2P[q][i] * Q2P[q][j];
   }
}

struct or_stokesop
{
   const or_mesh* m;
   int            L;
   lvl_t          l1, l0; /* velocity (level L+1) and pressure (level L) numbering */
   int64_t        sc1, w1, sc0, w0;
   int64_t*       cmap1;
   int64_t*       cmap0;
   int32_t*       flags1;
   double         c0, c1;
};

or_stokesop* or_stokesop_create( const or_mesh* m, int level, double c0, double c1 )
{
   or_stokesop* op = calloc( 1, sizeof( or_stokesop ) );
   op->m           = m;
   op->L           = level;
   op->l1          = or_lvl_info( m, level + 1 );
   op->l0          = or_lvl_info( m, level );
   op->cmap1       = or_build_cmap( m, level + 1, &op->sc1 );
   op->cmap0       = or_build_cmap( m, level, &op->sc0 );
   op->w1          = op->l1.mm + 1;
   op->w0          = op->l0.mm + 1;
   op->flags1      = malloc( sizeof( int32_t ) * op->l1.n );
   or_p1_flags( m, level + 1, op->flags1 );
   op->c0 = c0;
   op->c1 = c1;
   return op;
}

void or_stokesop_free( or_stokesop* op )
{
   if ( !op )
      return;
   free( op->cmap1 );
   free( op->cmap0 );
   free( op->flags1 );
   free( op );
}

static void cell_point( const or_mesh* m, int c, int64_t mm, const int64_t v[3], double* out )
{
   const double* x0 = m->X + 3 * m->cv[4 * c];
   const double* x1 = m->X + 3 * m->cv[4 * c + 1];
   const double* x2 = m->X + 3 * m->cv[4 * c + 2];
   const double* x3 = m->X + 3 * m->cv[4 * c + 3];
   const double  h  = 1.0 / (double) mm;
   for ( int d = 0; d < 3; ++d )
      out[d] = x0[d] + ( x1[d] - x0[d] ) * h * v[0] + ( x2[d] - x0[d] ) * h * v[1] + ( x3[d] - x0[d] ) * h * v[2];
}

/* every micro-cell of level L of cell c: computational vertices, the 10 velocity
 * ids (level L+1) and the 4 pressure ids (level L) */
typedef void ( *stokes_cb )( void* ctx, const double* p, const double* blend, const int64_t* uid, const int64_t* pid );

static void for_each_stokes_element( const or_stokesop* op, int c, stokes_cb f, void* ctx )
{
   const or_mesh* m  = op->m;
   const int64_t  mm = (int64_t) 1 << op->L;
   const double*  B  = m->blend ? m->blend + 14 * c : NULL;
   for ( int t = 0; t < 6; ++t )
   {
      const int64_t wt = mm - WSHIFT[t];
      for ( int64_t z = 0; z < wt; ++z )
         for ( int64_t y = 0; y + z < wt; ++y )
            for ( int64_t x = 0; x + y + z < wt; ++x )
            {
               int64_t v[4][3], uid[10], pid[4];
               double  p[12];
               for ( int k = 0; k < 4; ++k )
               {
                  v[k][0] = x + VOF[t][k][0];
                  v[k][1] = y + VOF[t][k][1];
                  v[k][2] = z + VOF[t][k][2];
                  cell_point( m, c, mm, v[k], p + 3 * k );
                  uid[k] = op->cmap1[c * op->sc1 + tet_index( op->w1, 2 * v[k][0], 2 * v[k][1], 2 * v[k][2] )];
                  pid[k] = op->cmap0[c * op->sc0 + tet_index( op->w0, v[k][0], v[k][1], v[k][2] )];
               }
               for ( int e = 0; e < 6; ++e )
               {
                  const int64_t* a = v[SE[e][0]];
                  const int64_t* b = v[SE[e][1]];
                  uid[4 + e]       = op->cmap1[c * op->sc1 + tet_index( op->w1, a[0] + b[0], a[1] + b[1], a[2] + b[2] )];
               }
               f( ctx, p, B, uid, pid );
            }
   }
}

typedef struct
{
   const or_stokesop* op;
   const double *     u, *p;
   double *           au, *ap;
   int                mask;
} sapply_ctx;

static void sapply_cb( void* vctx, const double* p, const double* blend, const int64_t* uid, const int64_t* pid )
{
   sapply_ctx*   c  = vctx;
   const int64_t n2 = c->op->l1.n;
   double        A[900], B[120], ue[30], pe[4], yu[30] = { 0 }, yp[4] = { 0 };
   or_stokes_element( p, blend, c->op->c0, c->op->c1, ( c->mask & 1 ) ? A : NULL, ( c->mask & 6 ) ? B : NULL );
   for ( int cc = 0; cc < 3; ++cc )
      for ( int i = 0; i < 10; ++i )
         ue[10 * cc + i] = c->u ? c->u[cc * n2 + uid[i]] : 0.0;
   for ( int i = 0; i < 4; ++i )
      pe[i] = c->p ? c->p[pid[i]] : 0.0;
   if ( c->mask & 1 )
      for ( int r = 0; r < 30; ++r )
         for ( int s = 0; s < 30; ++s )
            yu[r] += A[30 * r + s] * ue[s];
   if ( c->mask & 2 )
      for ( int r = 0; r < 30; ++r )
         for ( int i = 0; i < 4; ++i )
            yu[r] += B[30 * i + r] * pe[i];
   if ( c->mask & 4 )
      for ( int i = 0; i < 4; ++i )
         for ( int s = 0; s < 30; ++s )
            yp[i] += B[30 * i + s] * ue[s];
   if ( c->mask & 8 )
   {
      double M[16];
      or_stokes_kmass_element( p, blend, c->op->c0, c->op->c1, M );
      for ( int i = 0; i < 4; ++i )
         for ( int j = 0; j < 4; ++j )
            yp[i] += M[4 * i + j] * pe[j];
   }
   if ( c->mask & 3 )
      for ( int cc = 0; cc < 3; ++cc )
         for ( int i = 0; i < 10; ++i )
         {
#pragma omp atomic
            c->au[cc * n2 + uid[i]] += yu[10 * cc + i];
         }
   if ( c->mask & 12 )
      for ( int i = 0; i < 4; ++i )
      {
#pragma omp atomic
         c->ap[pid[i]] += yp[i];
      }
}

/* P2P1 Stokes apply of selected blocks (mask bit 1: A u, 2: B^T p into the
 * velocity; 4: B u into the pressure; mask 8: the 1/eta pressure mass), with the gemv data flow of the
 * elementwise operators (or_p2_apply): velocity DoFs of vertices, edges and
 * faces are written where their flag matches, cell-interior DoFs always
 * (zeroed first on Replace with Inner in flag); pressure DoFs are all Inner,
 * so pressure DoFs of vertices, edges and faces are written when flag
 * contains Inner. */
void or_stokes_apply( const or_stokesop* op, const double* u, const double* p, double* yu, double* yp, int flag, int mask,
                      int add )
{
   const int64_t n2 = op->l1.n, n1 = op->l0.n;
   double*       au = ( mask & 3 ) ? calloc( 3 * n2, sizeof( double ) ) : NULL;
   double*       ap = ( mask & 12 ) ? calloc( n1, sizeof( double ) ) : NULL;
   sapply_ctx    c  = { op, u, p, au, ap, mask };
#pragma omp parallel for schedule( dynamic, 1 )
   for ( int cell = 0; cell < op->m->nc; ++cell )
      for_each_stokes_element( op, cell, sapply_cb, &c );
   if ( mask & 3 )
      for ( int cc = 0; cc < 3; ++cc )
         for ( int64_t i = 0; i < n2; ++i )
         {
            double* d = yu + cc * n2 + i;
            if ( i >= op->l1.coff )
               *d = ( ( !add && ( flag & OR_INNER ) ) ? 0.0 : *d ) + au[cc * n2 + i];
            else if ( op->flags1[i] & flag )
               *d = add ? *d + au[cc * n2 + i] : au[cc * n2 + i];
         }
   /* pressure: every DoF has type Inner; cell-interior DoFs accumulate as in the
    * velocity (zeroed first on Replace with Inner in flag) */
   if ( mask & 12 )
      for ( int64_t i = 0; i < n1; ++i )
      {
         if ( i >= op->l0.coff )
            yp[i] = ( ( !add && ( flag & OR_INNER ) ) ? 0.0 : yp[i] ) + ap[i];
         else if ( flag & OR_INNER )
            yp[i] = add ? yp[i] + ap[i] : ap[i];
      }
   free( au );
   free( ap );
}

static void sdiag_cb( void* vctx, const double* p, const double* blend, const int64_t* uid, const int64_t* pid )
{
   sapply_ctx*   c  = vctx;
   const int64_t n2 = c->op->l1.n;
   double        A[900];
   (void) pid;
   or_stokes_element( p, blend, c->op->c0, c->op->c1, A, NULL );
   for ( int cc = 0; cc < 3; ++cc )
      for ( int i = 0; i < 10; ++i )
      {
#pragma omp atomic
         c->au[cc * n2 + uid[i]] += A[( 10 * cc + i ) * 31];
      }
}

/* inverse of the diagonal of the three diagonal blocks of A
 * (VectorToVectorOperator::computeInverseDiagonalOperatorValues) */
void or_stokes_inv_diag( const or_stokesop* op, double* out )
{
   const int64_t n2 = op->l1.n;
   memset( out, 0, 3 * n2 * sizeof( double ) );
   sapply_ctx c = { op, NULL, NULL, out, NULL, 1 };
#pragma omp parallel for schedule( dynamic, 1 )
   for ( int cell = 0; cell < op->m->nc; ++cell )
      for_each_stokes_element( op, cell, sdiag_cb, &c );
   for ( int64_t i = 0; i < 3 * n2; ++i )
      out[i] = 1.0 / out[i];
}

/* P1 diagonal operators of level L on the affine micro-cells: kind 0 the
 * inverse diagonal of the PSPG operator -(|T|^(2/3)/12) (grad p, grad q),
 * kind 1 the inverse lumped mass |T|/4 per vertex */
static void p1_affine_inv_diag( const or_mesh* m, int level, int kind, double* out )
{
   const lvl_t   l  = or_lvl_info( m, level );
   int64_t       sc;
   int64_t*      cmap = or_build_cmap( m, level, &sc );
   const int64_t w = l.mm + 1, mm = l.mm;
   memset( out, 0, l.n * sizeof( double ) );
#pragma omp parallel for schedule( dynamic, 1 )
   for ( int c = 0; c < m->nc; ++c )
      for ( int t = 0; t < 6; ++t )
      {
         const int64_t wt = mm - WSHIFT[t];
         for ( int64_t z = 0; z < wt; ++z )
            for ( int64_t y = 0; y + z < wt; ++y )
               for ( int64_t x = 0; x + y + z < wt; ++x )
               {
                  double  p[4][3], J[9], Ji[9], det;
                  int64_t ids[4];
                  for ( int k = 0; k < 4; ++k )
                  {
                     const int64_t v[3] = { x + VOF[t][k][0], y + VOF[t][k][1], z + VOF[t][k][2] };
                     cell_point( m, c, mm, v, p[k] );
                     ids[k] = cmap[c * sc + tet_index( w, v[0], v[1], v[2] )];
                  }
                  for ( int r = 0; r < 3; ++r )
                     for ( int k = 0; k < 3; ++k )
                        J[3 * r + k] = p[k + 1][r] - p[0][r];
                  inverse3( J, Ji, &det );
                  const double vol = fabs( det ) / 6.0;
                  for ( int k = 0; k < 4; ++k )
                  {
                     double v;
                     if ( kind == 0 )
                     {
                        /* grad lam_k: row k-1 of J^-1, lam_0 = 1 - sum */
                        double g[3];
                        for ( int d = 0; d < 3; ++d )
                           g[d] = k == 0 ? -( Ji[d] + Ji[3 + d] + Ji[6 + d] ) : Ji[3 * ( k - 1 ) + d];
                        v = -( pow( vol, 2.0 / 3.0 ) / 12.0 ) * vol * ( g[0] * g[0] + g[1] * g[1] + g[2] * g[2] );
                     }
                     else
                        v = vol / 4.0;
#pragma omp atomic
                     out[ids[k]] += v;
                  }
               }
      }
   for ( int64_t i = 0; i < l.n; ++i )
      out[i] = 1.0 / out[i];
   free( cmap );
}

void or_p1_pspg_inv_diag( const or_mesh* m, int level, double* out ) { p1_affine_inv_diag( m, level, 0, out ); }
void or_p1_lumped_inv_mass( const or_mesh* m, int level, double* out ) { p1_affine_inv_diag( m, level, 1, out ); }

/* boundary flags by surface (setMeshBoundaryFlagsInner/Outer of the shell
 * applications): primitives of boundary faces whose mean vertex radius is
 * below rmid get innerFlag, the others outerFlag */
void or_mesh_set_shell_flags( or_mesh* m, double rmid, int innerFlag, int outerFlag )
{
   for ( int f = 0; f < m->nf; ++f )
   {
      if ( m->fnc[f] != 1 )
         continue;
      const int* v = m->fv + 3 * f;
      double     r = 0.0;
      for ( int j = 0; j < 3; ++j )
      {
         const double* x = m->X + 3 * v[j];
         r += sqrt( x[0] * x[0] + x[1] * x[1] + x[2] * x[2] ) / 3.0;
      }
      const int flag = r < rmid ? innerFlag : outerFlag;
      m->fflag[f]    = flag;
      for ( int j = 0; j < 3; ++j )
         m->vflag[v[j]] = flag;
      for ( int e = 0; e < m->ne; ++e )
      {
         const int a = m->ev[2 * e], b = m->ev[2 * e + 1];
         int       na = 0;
         for ( int j = 0; j < 3; ++j )
            na += ( v[j] == a ) + ( v[j] == b );
         if ( na == 2 )
            m->eflag[e] = flag;
      }
   }
}

/* P2ProjectNormalOperator::project with the radial normal x/|x|
 * (p1functionspace/freeslip/VertexDoFProjectNormal.hpp:43-120): u -= (u.n) n
 * on the P2 DoFs of level `level` whose type is in flag; u = [ux|uy|uz] */
void or_p2_project_normal( const or_mesh* m, int level, double* u, int flag )
{
   if ( flag & OR_INNER )
   {
      fprintf( stderr, "or_p2_project_normal: the flag must select boundary types only\n" );
      abort();
   }
   const lvl_t   l  = or_lvl_info( m, level + 1 );
   const int64_t n2 = l.n;
   double*       X  = malloc( sizeof( double ) * 3 * n2 );
   int32_t*      fl = malloc( sizeof( int32_t ) * n2 );
   or_p1_coords( m, level + 1, X );
   or_p1_flags( m, level + 1, fl );
   for ( int64_t i = 0; i < l.coff; ++i )
   {
      if ( !( fl[i] & flag ) )
         continue;
      const double* x = X + 3 * i;
      const double  r = sqrt( x[0] * x[0] + x[1] * x[1] + x[2] * x[2] );
      const double  n[3] = { x[0] / r, x[1] / r, x[2] / r };
      const double  d    = n[0] * u[i] + n[1] * u[n2 + i] + n[2] * u[2 * n2 + i];
      for ( int c = 0; c < 3; ++c )
         u[c * n2 + i] -= d * n[c];
   }
   free( X );
   free( fl );
}
