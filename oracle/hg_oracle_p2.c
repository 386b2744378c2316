/*
 * hg_oracle_p2.c -- CPU restatement of the P2 part of HyTeG's hot path:
 * the icosahedral spherical-shell mesh, the IcosahedralShellMap blending,
 * the P2 div(k grad) element matrix with blending and 11-point quadrature,
 * and the elementwise P2 operator (apply, diagonal).
 *
 * TEST INFRASTRUCTURE ONLY (see hg_oracle.h).
 *
 * Reference files followed:
 *   mesh/MeshGenSphericalShell.cpp:38-700  (SHELLMESH_CLASSIC mesh, vertex/cell order)
 *   geometry/IcosahedralShellMap.hpp:109-236, 529-780 (evalF, evalDF, vertex classification)
 *   forms/form_hyteg_generated/p2/p2_div_k_grad_blending_q4.cpp:848-4481 (element matrix:
 *       quadrature points, k at blended points, blended Jacobians)
 *   elementwiseoperators/P2ElementwiseOperator.cpp:67-328, 418-568 (gemv, diagonal)
 *
 * P2 DoFs of level L are stored as the P1 lattice of level L+1: a micro-cell of
 * level L with vertices a_k (level-L lattice coordinates) has its 10 P2 nodes at
 * 2 a_k and a_i + a_j of the level-(L+1) lattice, so the canonical P2 numbering
 * is the canonical P1 numbering of level L+1.
 */
#include "hg_oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* spherical shell mesh (SHELLMESH_CLASSIC)                                  */
/* ------------------------------------------------------------------------- */
typedef struct
{
   int     ntan;
   double  iNode[12][3];
   int     dNode[10][4];
   double* nc; /* nodeCoords[is1][is2][id][3] */
} shell_t;

static double* NC( shell_t* s, int i1, int i2, int id ) { return s->nc + ( ( (size_t) i1 * s->ntan + i2 ) * 10 + id ) * 3; }

static void shell_setup( shell_t* s, int ntan )
{
   s->ntan              = ntan;
   const double fifthpi = 0.4 * asin( 1.0 );
   const double w       = 2.0 * acos( 1.0 / ( 2.0 * sin( fifthpi ) ) );
   const double cosw = cos( w ), sinw = sin( w );
   s->iNode[0][0] = 0.0, s->iNode[0][1] = 0.0, s->iNode[0][2] = 1.0;
   s->iNode[11][0] = 0.0, s->iNode[11][1] = 0.0, s->iNode[11][2] = -1.0;
   for ( int k = 1; k <= 5; ++k )
   {
      double phi     = 2.0 * ( k - 0.5 ) * fifthpi;
      s->iNode[k][0] = sinw * cos( phi );
      s->iNode[k][1] = sinw * sin( phi );
      s->iNode[k][2] = cosw;
      phi            = 2.0 * ( k - 1 ) * fifthpi;
      s->iNode[k + 5][0] = sinw * cos( phi );
      s->iNode[k + 5][1] = sinw * sin( phi );
      s->iNode[k + 5][2] = -cosw;
   }
   /* the ten diamonds: pole first, then counter-clockwise */
   static const int dn[10][4] = { { 0, 5, 6, 1 },  { 0, 1, 7, 2 },  { 0, 2, 8, 3 },  { 0, 3, 9, 4 },  { 0, 4, 10, 5 },
                                  { 11, 7, 1, 6 }, { 11, 8, 2, 7 }, { 11, 9, 3, 8 }, { 11, 10, 4, 9 }, { 11, 6, 5, 10 } };
   memcpy( s->dNode, dn, sizeof( dn ) );
   s->nc = calloc( (size_t) ntan * ntan * 10 * 3, sizeof( double ) );
   for ( int id = 0; id < 10; ++id )
   {
      const int L = id < 5 ? 1 : 3, R = id < 5 ? 3 : 1;
      memcpy( NC( s, 0, 0, id ), s->iNode[s->dNode[id][0]], 3 * sizeof( double ) );
      memcpy( NC( s, ntan - 1, 0, id ), s->iNode[s->dNode[id][L]], 3 * sizeof( double ) );
      memcpy( NC( s, ntan - 1, ntan - 1, id ), s->iNode[s->dNode[id][2]], 3 * sizeof( double ) );
      memcpy( NC( s, 0, ntan - 1, id ), s->iNode[s->dNode[id][R]], 3 * sizeof( double ) );
      /* midpoint refinement of the diamond: rows, columns, diagonals */
      for ( int m = 1; m < ntan - 1; m *= 2 )
      {
         const int l = ( ntan - 1 ) / m, l2 = l / 2;
         for ( int pass = 0; pass < 3; ++pass )
            for ( int j1 = 0; j1 <= m; ++j1 )
               for ( int j2 = 0; j2 < m; ++j2 )
               {
                  int i1, i2, a1, a2, b1, b2;
                  if ( pass == 0 )
                  {
                     i1 = j1 * l, i2 = j2 * l + l2;
                     a1 = i1, a2 = i2 - l2, b1 = i1, b2 = i2 + l2;
                  }
                  else if ( pass == 1 )
                  {
                     i1 = j2 * l + l2, i2 = j1 * l;
                     a1 = i1 - l2, a2 = i2, b1 = i1 + l2, b2 = i2;
                  }
                  else
                  {
                     if ( j1 == m )
                        continue;
                     i1 = j1 * l + l2, i2 = j2 * l + l2;
                     a1 = i1 - l2, a2 = i2 + l2, b1 = i1 + l2, b2 = i2 - l2;
                  }
                  const double* A = NC( s, a1, a2, id );
                  const double* B = NC( s, b1, b2, id );
                  double        x = A[0] + B[0], y = A[1] + B[1], z = A[2] + B[2];
                  double        n = sqrt( x * x + y * y + z * z );
                  double*       P = NC( s, i1, i2, id );
                  P[0] = x / n, P[1] = y / n, P[2] = z / n;
               }
      }
   }
}

static void vert_tuple( int ntan, int nrad, int j, int* is1, int* is2, int* id, int* ir )
{
   if ( j >= 2 * nrad )
   {
      int idx = j - 2 * nrad, aux = 10 * ( ntan - 1 ) * ( ntan - 1 );
      *ir = idx / aux;
      idx -= *ir * aux;
      aux = ( ntan - 1 ) * ( ntan - 1 );
      *id = idx / aux;
      idx -= *id * aux;
      *is2 = idx / ( ntan - 1 ) + 1;
      *is1 = idx % ( ntan - 1 );
   }
   else if ( j < nrad )
      *is1 = 0, *is2 = 0, *id = 0, *ir = j;
   else
      *is1 = 0, *is2 = 0, *id = 5, *ir = j - nrad;
}

static int tuple_vert( int ntan, int nrad, int is1, int is2, int id, int ir )
{
   if ( is1 == 0 && is2 == 0 )
      return id < 5 ? ir : nrad + ir;
   if ( is2 == 0 )
   {
      is2 = is1;
      is1 = 0;
      id  = id < 5 ? ( id + 4 ) % 5 : ( id + 4 ) % 5 + 5;
   }
   else if ( is1 == ntan - 1 )
   {
      is1 = ( ntan - 1 ) - is2;
      is2 = ntan - 1;
      id  = id < 5 ? ( id + 4 ) % 5 + 5 : id % 5;
   }
   return ir * ( ntan - 1 ) * ( ntan - 1 ) * 10 + id * ( ntan - 1 ) * ( ntan - 1 ) + ( is2 - 1 ) * ( ntan - 1 ) + is1 +
          2 * nrad;
}

/* map of the macro-cell onto its radial prism, see shell_classify */
static void blend_of_cell( const or_mesh* m, int c, int depth, double* out );

or_mesh* or_mesh_shell( int ntan, int nrad, double rmin, double rmax )
{
   if ( ntan < 2 || nrad < 2 || !( rmin > 0 && rmin < rmax ) )
      return NULL;
   shell_t s;
   shell_setup( &s, ntan );
   const int nv = ( 10 * ( ntan - 1 ) * ( ntan - 1 ) + 2 ) * nrad;
   double*   X  = malloc( sizeof( double ) * 3 * nv );
   for ( int j = 0; j < nv; ++j )
   {
      int is1, is2, id, ir;
      vert_tuple( ntan, nrad, j, &is1, &is2, &id, &ir );
      const double  layer = rmin + ( ( rmax - rmin ) / ( nrad - 1 ) ) * ir;
      const double* p     = NC( &s, is1, is2, id );
      for ( int d = 0; d < 3; ++d )
         X[3 * j + d] = p[d] * layer;
   }
   static const int off[8][3]  = { { 0, 0, 1 }, { 1, 0, 1 }, { 1, 1, 1 }, { 0, 1, 1 },
                                  { 0, 0, 0 }, { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 } };
   static const int tN[6][4]   = { { 0, 1, 3, 7 }, { 4, 7, 5, 0 }, { 0, 5, 1, 7 }, { 1, 2, 3, 6 }, { 5, 7, 6, 1 }, { 3, 6, 7, 1 } };
   static const int sN[6][4]   = { { 4, 7, 5, 1 }, { 0, 1, 3, 4 }, { 4, 3, 7, 1 }, { 7, 6, 5, 2 }, { 3, 1, 2, 7 }, { 5, 2, 1, 7 } };
   const int        nc         = ( ntan - 1 ) * ( ntan - 1 ) * ( nrad - 1 ) * 60;
   int64_t*         cells      = malloc( sizeof( int64_t ) * 4 * nc );
   for ( int e = 0; e < nc; ++e )
   {
      int idx = e, aux = 60 * ( ntan - 1 ) * ( ntan - 1 );
      int ir = idx / aux;
      idx -= ir * aux;
      aux    = 6 * ( ntan - 1 ) * ( ntan - 1 );
      int id = idx / aux;
      idx -= id * aux;
      aux     = 6 * ( ntan - 1 );
      int is2 = idx / aux;
      idx -= is2 * aux;
      int is1 = idx / 6;
      int it  = idx - 6 * is1;
      int v8[8];
      for ( int k = 0; k < 8; ++k )
      {
         const int ks1 = is1 + ( id <= 4 ? off[k][0] : off[k][1] );
         const int ks2 = is2 + ( id <= 4 ? off[k][1] : off[k][0] );
         v8[k]         = tuple_vert( ntan, nrad, ks1, ks2, id, ir + off[k][2] );
      }
      for ( int k = 0; k < 4; ++k )
         cells[4 * e + k] = v8[id <= 4 ? tN[it][k] : sN[it][k]];
   }
   or_mesh* m = or_mesh_create( nv, X, nc, cells );
   free( X );
   free( cells );
   free( s.nc );
   m->blend = malloc( sizeof( double ) * 14 * m->nc );
   for ( int c = 0; c < m->nc; ++c )
      blend_of_cell( m, c, 0, m->blend + 14 * c );
   return m;
}

static double norm3( const double* a ) { return sqrt( a[0] * a[0] + a[1] * a[1] + a[2] * a[2] ); }

/* IcosahedralShellMap::classifyVertices (IcosahedralShellMap.hpp:608-731) */
static void blend_of_cell( const or_mesh* m, int c, int depth, double* out )
{
   const int* cv = m->cv + 4 * c;
   double     P[4][3], r[4], rin = 1e300, rout = 0.0;
   for ( int k = 0; k < 4; ++k )
   {
      memcpy( P[k], m->X + 3 * cv[k], 3 * sizeof( double ) );
      r[k] = norm3( P[k] );
      rin  = r[k] < rin ? r[k] : rin;
      rout = r[k] > rout ? r[k] : rout;
   }
   int          nOuter = 0;
   const double eps    = ( rout - rin ) * 0.001;
   for ( int k = 0; k < 4; ++k )
      nOuter += ( rout - r[k] ) < eps ? 1 : 0;
   int iref = -1, iray = -1;
   for ( int k = 0; k < 4 && iref < 0; ++k )
      for ( int j = k + 1; j < 4 && iref < 0; ++j )
      {
         const double cx = P[k][1] * P[j][2] - P[k][2] * P[j][1];
         const double cy = P[k][2] * P[j][0] - P[k][0] * P[j][2];
         const double cz = P[k][0] * P[j][1] - P[k][1] * P[j][0];
         const double sc = r[k] * r[j];
         if ( fabs( cx ) / sc > 1e-14 || fabs( cy ) / sc > 1e-14 || fabs( cz ) / sc > 1e-14 )
            continue;
         iref = k;
         iray = j;
      }
   if ( nOuter == 2 && depth == 0 )
   {
      /* skew tet: take the map of the non-skew neighbour across the first
       * face that does not contain both radial vertices */
      const int gref = cv[iref], gray = cv[iray];
      for ( int f = 0; f < 4; ++f )
      {
         const int  fid = m->cf[4 * c + f];
         const int* fv  = m->fv + 3 * fid;
         int        hasRef = 0, hasRay = 0;
         for ( int j = 0; j < 3; ++j )
         {
            hasRef |= fv[j] == gref;
            hasRay |= fv[j] == gray;
         }
         if ( hasRef && hasRay )
            continue;
         for ( int o = 0; o < m->nc; ++o )
         {
            if ( o == c )
               continue;
            for ( int g = 0; g < 4; ++g )
               if ( m->cf[4 * o + g] == fid )
               {
                  blend_of_cell( m, o, 1, out );
                  return;
               }
         }
      }
      return;
   }
   int ithr = -1, ifor = -1;
   for ( int k = 0; k < 4; ++k )
      if ( k != iref && k != iray )
         ithr = k;
   for ( int k = 0; k < 4; ++k )
      if ( k != iref && k != iray && k != ithr )
         ifor = k;
   if ( ( nOuter == 1 && r[iref] < r[iray] ) || ( nOuter == 3 && r[iref] > r[iray] ) )
   {
      int t = iray;
      iray  = iref;
      iref  = t;
   }
   memcpy( out + 0, P[iray], 3 * sizeof( double ) );
   memcpy( out + 3, P[iref], 3 * sizeof( double ) );
   memcpy( out + 6, P[ithr], 3 * sizeof( double ) );
   memcpy( out + 9, P[ifor], 3 * sizeof( double ) );
   out[12] = norm3( P[iray] );
   out[13] = norm3( P[iref] );
}

void or_mesh_blend( const or_mesh* m, double* out14 )
{
   if ( m->blend )
      memcpy( out14, m->blend, sizeof( double ) * 14 * m->nc );
}

static double det3( const double a[3], const double b[3], const double c[3] )
{
   return a[0] * ( b[1] * c[2] - b[2] * c[1] ) - a[1] * ( b[0] * c[2] - b[2] * c[0] ) + a[2] * ( b[0] * c[1] - b[1] * c[0] );
}

/* IcosahedralShellMap::evalF and evalDF: radial projection, the radius
 * interpolated linearly in the barycentric coordinate of refVertex */
void or_shell_map( const double* B, const double* x, double* Fx, double* DF )
{
   const double *ray = B, *ref = B + 3, *thr = B + 6, *fr = B + 9;
   const double  radRay = B[12], radRef = B[13];
   double        a[3], t[3], f[3], d[3];
   for ( int k = 0; k < 3; ++k )
   {
      a[k] = ref[k] - ray[k];
      t[k] = thr[k] - ray[k];
      f[k] = fr[k] - ray[k];
      d[k] = x[k] - ray[k];
   }
   const double volT = det3( a, t, f );
   const double volX = det3( d, t, f );
   const double lam  = volX / volT;
   const double bary = fabs( lam );
   const double rho  = norm3( x );
   const double rad  = radRay + bary * ( radRef - radRay );
   for ( int k = 0; k < 3; ++k )
      Fx[k] = x[k] * rad / rho;
   if ( !DF )
      return;
   /* d(bary)/dx = sign(lam) * (t x f) / volT */
   double n[3] = { t[1] * f[2] - t[2] * f[1], t[2] * f[0] - t[0] * f[2], t[0] * f[1] - t[1] * f[0] };
   const double s = ( lam < 0 ? -1.0 : 1.0 ) * ( radRef - radRay ) / volT;
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j )
         DF[3 * i + j] = rad / rho * ( ( i == j ? 1.0 : 0.0 ) - x[i] * x[j] / ( rho * rho ) ) + x[i] * s * n[j] / rho;
}

/* ------------------------------------------------------------------------- */
/* P2 div(k grad) element matrix                                             */
/* ------------------------------------------------------------------------- */
/* Xiao-Gimbutas degree-4 rule with 11 points on the tetrahedron (barycentric
 * coordinates w.r.t. the element vertices 0..3, in the order the generated
 * form p2_div_k_grad_blending_q4 uses); weights are fixed by exactness on all
 * polynomials of degree <= 4 (computed at first use). */
static const double QP[11][4] = {
    { 0.17469405869723054, 0.040490506727590428, 0.01356070187980288, 0.77125473269537614 },
    { 0.081404918402859225, 0.75250850700965499, 0.068099370938206658, 0.097987203649279112 },
    { 0.30444840243449711, 0.24027666492807259, 0.12680172591539199, 0.32847320672203839 },
    { 0.74122888209362259, 0.067223294893383398, 0.035183929773598722, 0.15636389323939531 },
    { 0.053341239535745266, 0.41926631387951302, 0.047781435559086663, 0.47961101102565512 },
    { 0.43295349048135578, 0.45076587609127677, 0.059456616299433829, 0.056824017127933668 },
    { 0.53800720391618573, 0.12941137378891041, 0.33019041483746447, 0.0023910074574393651 },
    { 0.008991260093335729, 0.1215419913339278, 0.30649398842969028, 0.56297276014304609 },
    { 0.10660417256199362, 0.097204644587583267, 0.68439041545304002, 0.11180076739738309 },
    { 0.32923295974264682, 0.029569495206479609, 0.31790356021339461, 0.32329398483747901 },
    { 0.10384411641099309, 0.43271023904776862, 0.35382323920929709, 0.1096224053319412 },
};
static double QW[11];
static int    qw_ready = 0;

static double fact( int n ) { return n <= 1 ? 1.0 : n * fact( n - 1 ); }

/* least-squares solution of the 35 moment equations by Householder QR */
static void qw_init( void )
{
   if ( qw_ready )
      return;
   double M[35][12];
   int    row = 0;
   for ( int a = 0; a <= 4; ++a )
      for ( int b = 0; a + b <= 4; ++b )
         for ( int c = 0; a + b + c <= 4; ++c )
         {
            for ( int q = 0; q < 11; ++q )
               M[row][q] = pow( QP[q][1], a ) * pow( QP[q][2], b ) * pow( QP[q][3], c );
            M[row][11] = fact( a ) * fact( b ) * fact( c ) / fact( a + b + c + 3 );
            ++row;
         }
   for ( int k = 0; k < 11; ++k )
   {
      double nrm = 0.0;
      for ( int r = k; r < 35; ++r )
         nrm += M[r][k] * M[r][k];
      nrm           = sqrt( nrm );
      const double alpha = M[k][k] > 0 ? -nrm : nrm;
      double       v[35];
      double       vn = 0.0;
      for ( int r = 0; r < 35; ++r )
         v[r] = r < k ? 0.0 : M[r][k];
      v[k] -= alpha;
      for ( int r = k; r < 35; ++r )
         vn += v[r] * v[r];
      for ( int j = k; j < 12; ++j )
      {
         double d = 0.0;
         for ( int r = k; r < 35; ++r )
            d += v[r] * M[r][j];
         d *= 2.0 / vn;
         for ( int r = k; r < 35; ++r )
            M[r][j] -= d * v[r];
      }
   }
   for ( int i = 10; i >= 0; --i )
   {
      double s = M[i][11];
      for ( int j = i + 1; j < 11; ++j )
         s -= M[i][j] * QW[j];
      QW[i] = s / M[i][i];
   }
   qw_ready = 1;
}

void or_quadrature( double* pts44, double* w11 )
{
   qw_init();
   memcpy( pts44, QP, sizeof( QP ) );
   memcpy( w11, QW, sizeof( QW ) );
}


static double kfun( const kfun_t* K, const double* x )
{
   if ( K->nm == 0 )
      return 1.0;
   double s = 0.0;
   for ( int m = 0; m < K->nm; ++m )
      s += K->a[m] * sin( K->omega[3 * m] * x[0] + K->omega[3 * m + 1] * x[1] + K->omega[3 * m + 2] * x[2] + K->phi[m] );
   return exp( s );
}

/* local P2 nodes: vertices 0..3, then edges (0,1),(0,2),(0,3),(1,2),(1,3),(2,3) */
static const int P2E[6][2] = { { 0, 1 }, { 0, 2 }, { 0, 3 }, { 1, 2 }, { 1, 3 }, { 2, 3 } };

/* reference gradients (w.r.t. xi1..3, lambda_k = xi_k, lambda_0 = 1 - sum) of the 10 P2 basis functions */
static void p2_ref_grads( const double* lam, double g[10][3] )
{
   static const double dl[4][3] = { { -1, -1, -1 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
   for ( int i = 0; i < 4; ++i )
      for ( int d = 0; d < 3; ++d )
         g[i][d] = ( 4.0 * lam[i] - 1.0 ) * dl[i][d];
   for ( int e = 0; e < 6; ++e )
   {
      const int a = P2E[e][0], b = P2E[e][1];
      for ( int d = 0; d < 3; ++d )
         g[4 + e][d] = 4.0 * ( lam[a] * dl[b][d] + lam[b] * dl[a][d] );
   }
}

static void inv3( const double* A, double* Ai, double* det )
{
   const double d = A[0] * ( A[4] * A[8] - A[5] * A[7] ) - A[1] * ( A[3] * A[8] - A[5] * A[6] ) + A[2] * ( A[3] * A[7] - A[4] * A[6] );
   Ai[0]          = ( A[4] * A[8] - A[5] * A[7] ) / d;
   Ai[1]          = ( A[2] * A[7] - A[1] * A[8] ) / d;
   Ai[2]          = ( A[1] * A[5] - A[2] * A[4] ) / d;
   Ai[3]          = ( A[5] * A[6] - A[3] * A[8] ) / d;
   Ai[4]          = ( A[0] * A[8] - A[2] * A[6] ) / d;
   Ai[5]          = ( A[2] * A[3] - A[0] * A[5] ) / d;
   Ai[6]          = ( A[3] * A[7] - A[4] * A[6] ) / d;
   Ai[7]          = ( A[1] * A[6] - A[0] * A[7] ) / d;
   Ai[8]          = ( A[0] * A[4] - A[1] * A[3] ) / d;
   *det           = d;
}

/* A_ij = sum_q w_q k(F(x_q)) |det J| |det DF(x_q)| (G grad_i).(G grad_j),
 * G = (DF(x_q) J)^-T, x_q = sum_k lambda_qk p_k (p in computational space) */
void or_p2_element_matrix( const double* p /*4x3*/, const double* blend /*14 or NULL*/, int nm, const double* ka,
                           const double* kom, const double* kph, double* A /*10x10*/ )
{
   qw_init();
   kfun_t K = { nm, (double*) ka, (double*) kom, (double*) kph };
   double J[9];
   for ( int r = 0; r < 3; ++r )
      for ( int k = 0; k < 3; ++k )
         J[3 * r + k] = p[3 * ( k + 1 ) + r] - p[r];
   for ( int i = 0; i < 100; ++i )
      A[i] = 0.0;
   for ( int q = 0; q < 11; ++q )
   {
      double x[3], Fx[3], DF[9] = { 1, 0, 0, 0, 1, 0, 0, 0, 1 };
      for ( int d = 0; d < 3; ++d )
         x[d] = QP[q][0] * p[d] + QP[q][1] * p[3 + d] + QP[q][2] * p[6 + d] + QP[q][3] * p[9 + d];
      if ( blend )
         or_shell_map( blend, x, Fx, DF );
      else
         memcpy( Fx, x, sizeof( Fx ) );
      double M[9], Mi[9], detM;
      for ( int i = 0; i < 3; ++i )
         for ( int j = 0; j < 3; ++j )
            M[3 * i + j] = DF[3 * i] * J[j] + DF[3 * i + 1] * J[3 + j] + DF[3 * i + 2] * J[6 + j];
      inv3( M, Mi, &detM );
      double g[10][3], pg[10][3];
      p2_ref_grads( QP[q], g );
      for ( int i = 0; i < 10; ++i )
         for ( int d = 0; d < 3; ++d ) /* (M^-T g)_d = sum_e Mi[e][d] g_e */
            pg[i][d] = Mi[d] * g[i][0] + Mi[3 + d] * g[i][1] + Mi[6 + d] * g[i][2];
      const double wq = QW[q] * kfun( &K, Fx ) * fabs( detM );
      for ( int i = 0; i < 10; ++i )
         for ( int j = 0; j < 10; ++j )
            A[10 * i + j] += wq * ( pg[i][0] * pg[j][0] + pg[i][1] * pg[j][1] + pg[i][2] * pg[j][2] );
   }
}

/* binary16 rounding (round to nearest, ties to even) of a positive value in
 * the normal binary16 range: 11 significant bits */
static double round_half( double v )
{
   int          e;
   const double m = frexp( v, &e ); /* v = m 2^e, m in [0.5, 1) */
   return ldexp( nearbyint( ldexp( m, 11 ) ), e - 11 );
}

/* The reduced element form of the low-precision applies (the preconditioner
 * inside mixed-precision refinement; no counterpart in the reference): the
 * micro-cell is the affine tetrahedron of its four blended vertices F(p_k), k
 * is constant on it, k(F(x_c)) at the blended centroid rounded to binary16,
 * and the P2 stiffness of that tetrahedron is integrated by the 4-point
 * degree-2 rule (exact for it). */
void or_p2_reduced_element_matrix( const double* p, const double* blend, int nm, const double* ka, const double* kom,
                                   const double* kph, double* A )
{
   kfun_t K = { nm, (double*) ka, (double*) kom, (double*) kph };
   double P[4][3], xc[3] = { 0, 0, 0 }, Fc[3];
   for ( int k = 0; k < 4; ++k )
   {
      double DF[9];
      if ( blend )
         or_shell_map( blend, p + 3 * k, P[k], DF );
      else
         memcpy( P[k], p + 3 * k, sizeof( P[k] ) );
      for ( int d = 0; d < 3; ++d )
         xc[d] += 0.25 * p[3 * k + d];
   }
   if ( blend )
   {
      double DF[9];
      or_shell_map( blend, xc, Fc, DF );
   }
   else
      memcpy( Fc, xc, sizeof( Fc ) );
   /* rounded to binary32, then to binary16: the conversion the GPU table uses */
   const double kc = round_half( (double) (float) kfun( &K, Fc ) );
   double       J[9], Ji[9], detJ;
   for ( int r = 0; r < 3; ++r )
      for ( int k = 0; k < 3; ++k )
         J[3 * r + k] = P[k + 1][r] - P[0][r];
   inv3( J, Ji, &detJ );
   for ( int i = 0; i < 100; ++i )
      A[i] = 0.0;
   const double a = 0.1381966011250105, b = 0.5854101966249685;
   for ( int q = 0; q < 4; ++q )
   {
      double lam[4] = { a, a, a, a };
      lam[q]        = b;
      double g[10][3], pg[10][3];
      p2_ref_grads( lam, g );
      for ( int i = 0; i < 10; ++i )
         for ( int d = 0; d < 3; ++d )
            pg[i][d] = Ji[d] * g[i][0] + Ji[3 + d] * g[i][1] + Ji[6 + d] * g[i][2];
      const double wq = ( 1.0 / 24.0 ) * kc * fabs( detJ );
      for ( int i = 0; i < 10; ++i )
         for ( int j = 0; j < 10; ++j )
            A[10 * i + j] += wq * ( pg[i][0] * pg[j][0] + pg[i][1] * pg[j][1] + pg[i][2] * pg[j][2] );
   }
}

void or_p2op_set_reduced( or_p2op* op, int on ) { op->reduced = on; }

/* ------------------------------------------------------------------------- */
/* P2 elementwise operator                                                    */
/* ------------------------------------------------------------------------- */

or_p2op* or_p2op_create( const or_mesh* m, int level, int nm, const double* ka, const double* kom, const double* kph )
{
   or_p2op* op = calloc( 1, sizeof( or_p2op ) );
   op->m       = m;
   op->L       = level;
   op->l1      = or_lvl_info( m, level + 1 );
   op->cmap    = or_build_cmap( m, level + 1, &op->sc1 );
   op->w1      = op->l1.mm + 1;
   op->K.nm    = nm;
   if ( nm > 0 )
   {
      op->K.a     = malloc( sizeof( double ) * nm );
      op->K.omega = malloc( sizeof( double ) * 3 * nm );
      op->K.phi   = malloc( sizeof( double ) * nm );
      memcpy( op->K.a, ka, sizeof( double ) * nm );
      memcpy( op->K.omega, kom, sizeof( double ) * 3 * nm );
      memcpy( op->K.phi, kph, sizeof( double ) * nm );
   }
   op->flags = malloc( sizeof( int32_t ) * op->l1.n );
   or_p1_flags( m, level + 1, op->flags );
   return op;
}

void or_p2op_free( or_p2op* op )
{
   if ( !op )
      return;
   free( op->cmap );
   free( op->K.a );
   free( op->K.omega );
   free( op->K.phi );
   free( op->flags );
   free( op );
}

/* computational coordinates of level-L lattice point q of cell c (coordinateFromIndex) */
static void comp_coord( const or_mesh* m, int c, int64_t mm, const int64_t q[3], double* out )
{
   const double* v0 = m->X + 3 * m->cv[4 * c];
   const double* v1 = m->X + 3 * m->cv[4 * c + 1];
   const double* v2 = m->X + 3 * m->cv[4 * c + 2];
   const double* v3 = m->X + 3 * m->cv[4 * c + 3];
   const double  h  = 1.0 / (double) mm;
   for ( int d = 0; d < 3; ++d )
      out[d] = v0[d] + ( v1[d] - v0[d] ) * h * q[0] + ( v2[d] - v0[d] ) * h * q[1] + ( v3[d] - v0[d] ) * h * q[2];
}

/* loop over all micro-cells of level L; calls f with the 10 canonical ids and the element matrix */
typedef void ( *elem_cb )( void* ctx, const int64_t* ids, const double* A );

static void for_each_element_in( const or_p2op* op, int c0, int c1, elem_cb f, void* ctx )
{
   const or_mesh* m  = op->m;
   const int64_t  mm = (int64_t) 1 << op->L;
   for ( int c = c0; c < c1; ++c )
   {
      const double* B = m->blend ? m->blend + 14 * c : NULL;
      for ( int t = 0; t < 6; ++t )
      {
         const int64_t wt = mm - WSHIFT[t];
         for ( int64_t z = 0; z < wt; ++z )
            for ( int64_t y = 0; y + z < wt; ++y )
               for ( int64_t x = 0; x + y + z < wt; ++x )
               {
                  int64_t v[4][3];
                  double  p[12], A[100];
                  int64_t ids[10];
                  for ( int k = 0; k < 4; ++k )
                  {
                     v[k][0] = x + VOF[t][k][0];
                     v[k][1] = y + VOF[t][k][1];
                     v[k][2] = z + VOF[t][k][2];
                     comp_coord( m, c, mm, v[k], p + 3 * k );
                     ids[k] = op->cmap[c * op->sc1 + tet_index( op->w1, 2 * v[k][0], 2 * v[k][1], 2 * v[k][2] )];
                  }
                  for ( int e = 0; e < 6; ++e )
                  {
                     const int64_t* a = v[P2E[e][0]];
                     const int64_t* b = v[P2E[e][1]];
                     ids[4 + e]       = op->cmap[c * op->sc1 + tet_index( op->w1, a[0] + b[0], a[1] + b[1], a[2] + b[2] )];
                  }
                  if ( op->reduced )
                     or_p2_reduced_element_matrix( p, B, op->K.nm, op->K.a, op->K.omega, op->K.phi, A );
                  else
                     or_p2_element_matrix( p, B, op->K.nm, op->K.a, op->K.omega, op->K.phi, A );
                  f( ctx, ids, A );
               }
      }
   }
}

static void for_each_element( const or_p2op* op, elem_cb f, void* ctx ) { for_each_element_in( op, 0, op->m->nc, f, ctx ); }

typedef struct
{
   const double* src;
   double*       acc;
} apply_ctx;

static void apply_cb( void* vctx, const int64_t* ids, const double* A )
{
   apply_ctx* c = vctx;
   double     xe[10];
   for ( int i = 0; i < 10; ++i )
      xe[i] = c->src[ids[i]];
   for ( int i = 0; i < 10; ++i )
   {
      double s = 0.0;
      for ( int j = 0; j < 10; ++j )
         s += A[10 * i + j] * xe[j];
#pragma omp atomic
      c->acc[ids[i]] += s;
   }
}

/* P2ElementwiseOperator::apply = gemv(1, src, 0|1, dst) (P2ElementwiseOperator.cpp:110-328):
 * cell-interior DoFs accumulate (zeroed first on Replace when Inner is in
 * flag), DoFs of vertices/edges/faces whose type matches flag get the sum
 * of all cells (replacing or adding), the others are untouched. */
void or_p2_apply( const or_p2op* op, const double* src, double* dst, int flag, int add )
{
   const int64_t n   = op->l1.n;
   double*       acc = calloc( n, sizeof( double ) );
   apply_ctx     c   = { src, acc };
   /* macro-cells in parallel; contributions to shared DoFs are added atomically */
#pragma omp parallel for schedule( dynamic, 1 )
   for ( int cell = 0; cell < op->m->nc; ++cell )
      for_each_element_in( op, cell, cell + 1, apply_cb, &c );
   for ( int64_t i = 0; i < n; ++i )
   {
      if ( i >= op->l1.coff )
         dst[i] = ( ( !add && ( flag & OR_INNER ) ) ? 0.0 : dst[i] ) + acc[i];
      else if ( op->flags[i] & flag )
         dst[i] = add ? dst[i] + acc[i] : acc[i];
   }
   free( acc );
}

static void diag_cb( void* vctx, const int64_t* ids, const double* A )
{
   double* d = vctx;
   for ( int i = 0; i < 10; ++i )
      d[ids[i]] += A[11 * i];
}

/* computeInverseDiagonalOperatorValues (P2ElementwiseOperator.cpp:418-568) */
void or_p2_inv_diag( const or_p2op* op, double* out )
{
   const int64_t n = op->l1.n;
   memset( out, 0, sizeof( double ) * n );
   for_each_element( op, diag_cb, out );
   for ( int64_t i = 0; i < n; ++i )
      out[i] = 1.0 / out[i];
}

typedef struct
{
   int64_t* rows;
   int64_t* cols;
   double*  vals;
   int64_t  k;
} coo_ctx;

static void coo_cb( void* vctx, const int64_t* ids, const double* A )
{
   coo_ctx* c = vctx;
   for ( int i = 0; i < 10; ++i )
      for ( int j = 0; j < 10; ++j )
      {
         c->rows[c->k] = ids[i];
         c->cols[c->k] = ids[j];
         c->vals[c->k] = A[10 * i + j];
         c->k++;
      }
}

int64_t or_p2op_nelem( const or_p2op* op )
{
   const int64_t mm = (int64_t) 1 << op->L;
   return (int64_t) op->m->nc * mm * mm * mm;
}

/* all element matrices as COO triplets (100 per micro-cell) */
void or_p2op_coo( const or_p2op* op, int64_t* rows, int64_t* cols, double* vals )
{
   coo_ctx c = { rows, cols, vals, 0 };
   for_each_element( op, coo_cb, &c );
}
