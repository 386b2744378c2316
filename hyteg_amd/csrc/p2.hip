// P2 elementwise div(k grad) operator with icosahedral-shell blending:
// host set-up, HIP kernels and C-ABI entry points.
//
// A P2 function of level L is stored on the lattice of level L+1.  One thread
// handles one micro-cell of level L: it gathers the 10 P2 values, evaluates
// the element action with the 11-point Xiao-Gimbutas rule at blended points
// (p2_div_k_grad_blending_q4, P2ElementwiseOperator::gemv,
// P2ElementwiseOperator.cpp:67-233) and adds the 10 results into the cell's
// lattice; contributions on macro-faces/edges/vertices are then summed over
// the incident cells (communicateAdditively<Cell, Face/Edge/Vertex>).
//
// The element action is evaluated as  y_i = sum_q w_q k(F(x_q)) |det M_q|
// (M_q^-T grad phi_i) . (sum_j x_j M_q^-T grad phi_j),  M_q = DF(x_q) J, which is
// the element matrix of the reference applied to the element vector without
// forming the 10x10 matrix.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>

#include "abi_util.hpp"
#include "halo.hpp"
#include "hyteg_amd.h"
#include "p2tile.hpp"
#include "runtime.hpp"

namespace hg {

// quadrature points (barycentric w.r.t. the element vertices in the order
// getMicroVerticesFromMicroCell lists them) and weights (reference tet, sum 1/6)
__constant__ double cQP[11][4];
__constant__ double cQW[11];
__constant__ double cQ4[11][4];  // 4 lam_k at quadrature point q
__constant__ double cQ4m[11][4]; // 4 lam_k - 1
__constant__ float  cQWf[11], cQ4f[11][4], cQ4mf[11][4]; // the same in FP32
__constant__ double cQL[11][3];  // lam_1, lam_2, lam_3 at quadrature point q
__constant__ float  cQLf[11][3];

template < class R >
struct QuadConst;
template <>
struct QuadConst< double >
{
   __device__ static const double* w() { return cQW; }
   __device__ static const double* a4( int q ) { return cQ4[q]; }
   __device__ static const double* b4( int q ) { return cQ4m[q]; }
   __device__ static const double* lam( int q ) { return cQL[q]; }
};
template <>
struct QuadConst< float >
{
   __device__ static const float* w() { return cQWf; }
   __device__ static const float* a4( int q ) { return cQ4f[q]; }
   __device__ static const float* b4( int q ) { return cQ4mf[q]; }
   __device__ static const float* lam( int q ) { return cQLf[q]; }
};

// Xiao-Gimbutas degree-4 rule on the tetrahedron, 11 points; the weights are
// fixed by exactness for all polynomials of degree <= 4
static const double kQP[11][4] = {
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

// weights from the moment equations sum_q w_q xi^a = a1! a2! a3! / (|a| + 3)!,
// |a| <= 4 (35 equations, 11 unknowns, consistent), solved in the least-squares
// sense with Householder reflections
static void quadratureWeights( double w[11] )
{
   auto fact = []( int n ) {
      double f = 1;
      for ( int i = 2; i <= n; ++i )
         f *= i;
      return f;
   };
   std::vector< std::array< double, 12 > > A;
   for ( int a = 0; a <= 4; ++a )
      for ( int b = 0; a + b <= 4; ++b )
         for ( int c = 0; a + b + c <= 4; ++c )
         {
            std::array< double, 12 > r{};
            for ( int q = 0; q < 11; ++q )
               r[q] = std::pow( kQP[q][1], a ) * std::pow( kQP[q][2], b ) * std::pow( kQP[q][3], c );
            r[11] = fact( a ) * fact( b ) * fact( c ) / fact( a + b + c + 3 );
            A.push_back( r );
         }
   const size_t m = A.size();
   for ( int k = 0; k < 11; ++k )
   {
      double nrm = 0.0;
      for ( size_t r = k; r < m; ++r )
         nrm += A[r][k] * A[r][k];
      nrm = std::sqrt( nrm );
      std::vector< double > v( m, 0.0 );
      for ( size_t r = k; r < m; ++r )
         v[r] = A[r][k];
      v[k] += A[k][k] > 0 ? nrm : -nrm;
      double vv = 0.0;
      for ( size_t r = k; r < m; ++r )
         vv += v[r] * v[r];
      for ( int j = k; j < 12; ++j )
      {
         double d = 0.0;
         for ( size_t r = k; r < m; ++r )
            d += v[r] * A[r][j];
         d *= 2.0 / vv;
         for ( size_t r = k; r < m; ++r )
            A[r][j] -= d * v[r];
      }
   }
   for ( int i = 10; i >= 0; --i )
   {
      double s = A[i][11];
      for ( int j = i + 1; j < 11; ++j )
         s -= A[i][j] * w[j];
      w[i] = s / A[i][i];
   }
}

struct CellGeom
{
   double v0[3], e1[3], e2[3], e3[3]; // v0 and edge vectors v_k - v0 of the macro-cell
   double blend[14];                  // ray, ref, thr, for, radRay, radRef
   int    kind;                       // 0 affine, 1 shell
};

struct KParams
{
   int    nm;
   double a[32], om[32][3], ph[32];
};


struct P2Operator : OperatorBase
{
   Storage*               st = nullptr;
   int                    minL = 0, maxL = 0;
   bool                   storeCoefficients = true;
   int                    nm                = 0; // number of sine modes of k (0: k = 1)
   DevBuf< CellGeom >     geom;
   std::map< int, DevBuf< P2CellGeom > >           tgeom;  // by P2 level
   std::map< int, DevBuf< P2CellGeomT< float > > > tgeomF; // by P2 level, for the FP32 kernels
   DevBuf< KParams >      kpar;
   std::map< int, DevBuf< double > > invDiag; // by P2 level, lattice L+1 storage
   // k at the quadrature points (tile-major), by (level, value type of the table)
   std::map< std::pair< int, int >, DevBuf< char > > kq;
   // reduced element form of the FP32 / FP16 applies (the preconditioner of the
   // mixed-precision refinement): on levels >= reducedFrom every micro-cell is
   // integrated as the affine tetrahedron of its four blended vertices with k
   // constant (its value at the blended centroid), by the 4-point degree-2 rule
   bool                                   reduced     = true;
   int                                    reducedFrom = 2;
   std::map< int, DevBuf< __half > >      kc;       // k per micro-cell (tile-major), by P2 level
   std::map< int, DevBuf< float > >       gm;       // reduced-form metric per micro-cell, by P2 level
   std::map< int, DevBuf< float > >       accum;    // FP32 accumulator of the FP16 applies, by P2 level
   // batched operators (hg_p2_divkgrad_create_batched): 2^lvSlots coefficient
   // slots for as many interleaved vectors; kpar holds one KParams per slot, a
   // level's coefficient table holds the slots' tables one after another
   // (tableStride values each) and its inverse diagonal is interleaved
   int                                    nslot = 1, lvSlots = 0;
   std::map< int, int64_t >               tableStride; // by P2 level
   std::map< int, DevBuf< double > >      invScratch;  // one slot's diagonal, by P2 level
   DevBuf< int32_t >                      slotOn;      // 1: the apply computes the slot, 0: its result is 0
};

__device__ __forceinline__ double kEval( const KParams& K, const double x[3] )
{
   if ( K.nm == 0 )
      return 1.0;
   double s = 0.0;
   for ( int m = 0; m < K.nm; ++m )
      s += K.a[m] * sin( K.om[m][0] * x[0] + K.om[m][1] * x[1] + K.om[m][2] * x[2] + K.ph[m] );
   return exp( s );
}

// shell blending (IcosahedralShellMap::evalF / evalDF): radial projection with
// the radius linear in the barycentric coordinate of refVertex
__device__ __forceinline__ void shellMap( const double* B, const double x[3], double F[3], double DF[9] )
{
   const double a[3] = { B[3] - B[0], B[4] - B[1], B[5] - B[2] };
   const double t[3] = { B[6] - B[0], B[7] - B[1], B[8] - B[2] };
   const double f[3] = { B[9] - B[0], B[10] - B[1], B[11] - B[2] };
   const double d[3] = { x[0] - B[0], x[1] - B[1], x[2] - B[2] };
   const double n[3] = { t[1] * f[2] - t[2] * f[1], t[2] * f[0] - t[0] * f[2], t[0] * f[1] - t[1] * f[0] };
   const double volT = a[0] * n[0] + a[1] * n[1] + a[2] * n[2];
   const double lam  = ( d[0] * n[0] + d[1] * n[1] + d[2] * n[2] ) / volT;
   const double rho  = sqrt( x[0] * x[0] + x[1] * x[1] + x[2] * x[2] );
   const double rad  = B[12] + fabs( lam ) * ( B[13] - B[12] );
   const double q    = rad / rho;
   const double s    = ( lam < 0 ? -1.0 : 1.0 ) * ( B[13] - B[12] ) / volT / rho;
   for ( int i = 0; i < 3; ++i )
   {
      F[i] = x[i] * q;
      for ( int j = 0; j < 3; ++j )
         DF[3 * i + j] = q * ( ( i == j ? 1.0 : 0.0 ) - x[i] * x[j] / ( rho * rho ) ) + x[i] * s * n[j];
   }
}


// locate micro-cell gid: returns cell c, type t and anchor (x,y,z)
__device__ __forceinline__ bool microOf( int64_t gid, int64_t mm, int64_t nc, int64_t& c, int& t, int64_t a[3] )
{
   int64_t per = 0;
   for ( int k = 0; k < 6; ++k )
      per += tet( mm - MicroCells::rowShift[k] );
   if ( gid >= per * nc )
      return false;
   c         = gid / per;
   int64_t j = gid - c * per;
   t         = 0;
   while ( j >= tet( mm - MicroCells::rowShift[t] ) )
   {
      j -= tet( mm - MicroCells::rowShift[t] );
      ++t;
   }
   decodeMicro( mm - MicroCells::rowShift[t], j, a[0], a[1], a[2] );
   return true;
}

// physical gradients of the 4 barycentric coordinates D[k] and |det M| at a
// quadrature point, plus the blended point F
__device__ __forceinline__ void elementGeometry( const CellGeom& g, const double p[4][3], const double lam[4],
                                                 double D[4][3], double& absdet, double F[3] )
{
   double J[9];
   for ( int r = 0; r < 3; ++r )
      for ( int k = 0; k < 3; ++k )
         J[3 * r + k] = p[k + 1][r] - p[0][r];
   double x[3];
   for ( int d = 0; d < 3; ++d )
      x[d] = lam[0] * p[0][d] + lam[1] * p[1][d] + lam[2] * p[2][d] + lam[3] * p[3][d];
   double M[9];
   if ( g.kind )
   {
      double DF[9];
      shellMap( g.blend, x, F, DF );
      for ( int i = 0; i < 3; ++i )
         for ( int j = 0; j < 3; ++j )
            M[3 * i + j] = DF[3 * i] * J[j] + DF[3 * i + 1] * J[3 + j] + DF[3 * i + 2] * J[6 + j];
   }
   else
   {
      for ( int i = 0; i < 9; ++i )
         M[i] = J[i];
      F[0] = x[0], F[1] = x[1], F[2] = x[2];
   }
   const double c00 = M[4] * M[8] - M[5] * M[7], c01 = M[5] * M[6] - M[3] * M[8], c02 = M[3] * M[7] - M[4] * M[6];
   const double det = M[0] * c00 + M[1] * c01 + M[2] * c02;
   const double id  = 1.0 / det;
   // rows of M^-1 are the gradients of barycentric coordinates 1..3
   D[1][0] = c00 * id;
   D[1][1] = ( M[2] * M[7] - M[1] * M[8] ) * id;
   D[1][2] = ( M[1] * M[5] - M[2] * M[4] ) * id;
   D[2][0] = c01 * id;
   D[2][1] = ( M[0] * M[8] - M[2] * M[6] ) * id;
   D[2][2] = ( M[2] * M[3] - M[0] * M[5] ) * id;
   D[3][0] = c02 * id;
   D[3][1] = ( M[1] * M[6] - M[0] * M[7] ) * id;
   D[3][2] = ( M[0] * M[4] - M[1] * M[3] ) * id;
   for ( int d = 0; d < 3; ++d )
      D[0][d] = -( D[1][d] + D[2][d] + D[3][d] );
   absdet = fabs( det );
}

// physical gradients of the 10 P2 basis functions at barycentric point lam
__device__ __forceinline__ void p2Gradients( const double lam[4], const double D[4][3], double G[10][3] )
{
   for ( int i = 0; i < 4; ++i )
      for ( int d = 0; d < 3; ++d )
         G[i][d] = ( 4.0 * lam[i] - 1.0 ) * D[i][d];
   for ( int e = 0; e < 6; ++e )
   {
      const int a = cP2E[e][0], b = cP2E[e][1];
      for ( int d = 0; d < 3; ++d )
         G[4 + e][d] = 4.0 * ( lam[a] * D[b][d] + lam[b] * D[a][d] );
   }
}

// vertex coordinates (computational space) and the 10 lattice positions (level L+1)
__device__ __forceinline__ void microElement( const CellGeom& g, int t, const int64_t a[3], int64_t mm, int64_t c,
                                              int64_t sc1, int64_t w1, double p[4][3], int64_t pos[10] )
{
   const double h = 1.0 / double( mm );
   int64_t      v[4][3];
   for ( int k = 0; k < 4; ++k )
   {
      for ( int d = 0; d < 3; ++d )
         v[k][d] = a[d] + cVof[t][k][d];
      for ( int d = 0; d < 3; ++d )
         p[k][d] = g.v0[d] + g.e1[d] * ( h * double( v[k][0] ) ) + g.e2[d] * ( h * double( v[k][1] ) ) +
                   g.e3[d] * ( h * double( v[k][2] ) );
      pos[k] = c * sc1 + tetIndex( w1, 2 * v[k][0], 2 * v[k][1], 2 * v[k][2] );
   }
   for ( int e = 0; e < 6; ++e )
   {
      const int64_t* va = v[cP2E[e][0]];
      const int64_t* vb = v[cP2E[e][1]];
      pos[4 + e]        = c * sc1 + tetIndex( w1, va[0] + vb[0], va[1] + vb[1], va[2] + vb[2] );
   }
}


__device__ __forceinline__ double kAt( const P2CellGeom& g, const KParams& K, const double x[3], double ls )
{
   if ( g.kind )
   {
      double qq, w[3], beta;
      shellScalars( g, x, ls, qq, w, beta );
      const double F[3] = { qq * x[0], qq * x[1], qq * x[2] };
      return kEval( K, F );
   }
   return kEval( K, x );
}


// k(F(x_q)) for every quadrature point of every micro-cell, stored tile-major:
// table[(tile * 11 + q) * 384 + thread] (zero for threads outside the macro-cell)
// (computed in FP64 and rounded to the table's value type K)
template < class K >
__global__ __launch_bounds__( kTileThreads ) void kP2TileCoefficients( const P2CellGeom* __restrict__ geom,
                                                                      const KParams* __restrict__ kp, int64_t mm,
                                                                      int64_t ntiles, K* __restrict__ table )
{
   const int64_t lb = blockIdx.x;
   if ( lb >= ntiles )
      return;
   const TileCoord            tc = tileOf( lb, mm );
   const P2CellGeom&          g  = geom[tc.c];
   const TileThread< double > th = tileThread( g, tc, mm );
   for ( int q = 0; q < 11; ++q )
   {
      const double* xq   = g.XQ[th.t][q];
      const double  x[3] = { th.p0[0] + xq[0], th.p0[1] + xq[1], th.p0[2] + xq[2] };
      const double  v    = th.valid ? kAt( g, *kp, x, th.ls0 + xq[3] ) : 0.0;
      table[( lb * 11 + q ) * kTileThreads + threadIdx.x] = fromAcc< K >( typename Acc< K >::type( v ) );
   }
}


// ---- reduced element form (low-precision applies) ---------------------------
// k at the blended centroid of every micro-cell, stored tile-major like the
// quadrature-point table: table[tile * 384 + thread] (FP64 evaluation, FP16 value)
__global__ __launch_bounds__( kTileThreads ) void kP2TileCentroidK( const P2CellGeom* __restrict__ geom,
                                                                   const KParams* __restrict__ kp, int64_t mm,
                                                                   int64_t ntiles, __half* __restrict__ table )
{
   const int64_t lb = blockIdx.x;
   if ( lb >= ntiles )
      return;
   const TileCoord            tc = tileOf( lb, mm );
   const P2CellGeom&          g  = geom[tc.c];
   const TileThread< double > th = tileThread( g, tc, mm );
   const double*              J  = g.J[th.t];
   double                     x[3];
   for ( int d = 0; d < 3; ++d )
      x[d] = th.p0[d] + 0.25 * ( J[3 * d] + J[3 * d + 1] + J[3 * d + 2] );
   const double ls = g.nl[0] * x[0] + g.nl[1] * x[1] + g.nl[2] * x[2] - g.lam0;
   table[lb * kTileThreads + threadIdx.x] = __float2half( th.valid ? float( kAt( g, *kp, x, ls ) ) : 0.0f );
}

// the 4-point degree-2 rule: one barycentric coordinate kB, the others kA
constexpr float kQ2A = 0.1381966011250105f, kQ2B = 0.5854101966249685f;

// y += A_e u of the reduced form: constant metric G (scaled by k |det| w) and
// the 4 points of the degree-2 rule, in the barycentric form of p2ElementAction
__device__ __forceinline__ void p2ReducedAction( const float G[6], const float u[10], float y[10] )
{
#pragma unroll
   for ( int q = 0; q < 4; ++q )
   {
      float l[4];
#pragma unroll
      for ( int k = 0; k < 4; ++k )
         l[k] = k == q ? kQ2B : kQ2A;
      const float g0 = ( 4.0f * l[0] - 1.0f ) * u[0] + 4.0f * ( l[1] * u[4] + l[2] * u[5] + l[3] * u[6] );
      const float g1 = ( 4.0f * l[1] - 1.0f ) * u[1] + 4.0f * ( l[0] * u[4] + l[2] * u[7] + l[3] * u[8] );
      const float g2 = ( 4.0f * l[2] - 1.0f ) * u[2] + 4.0f * ( l[0] * u[5] + l[1] * u[7] + l[3] * u[9] );
      const float g3 = ( 4.0f * l[3] - 1.0f ) * u[3] + 4.0f * ( l[0] * u[6] + l[1] * u[8] + l[2] * u[9] );
      const float d1 = g1 - g0, d2 = g2 - g0, d3 = g3 - g0;
      const float e1 = G[0] * d1 + G[1] * d2 + G[2] * d3;
      const float e2 = G[1] * d1 + G[3] * d2 + G[4] * d3;
      const float e3 = G[2] * d1 + G[4] * d2 + G[5] * d3;
      const float e0 = -( e1 + e2 + e3 );
      y[0] += ( 4.0f * l[0] - 1.0f ) * e0;
      y[1] += ( 4.0f * l[1] - 1.0f ) * e1;
      y[2] += ( 4.0f * l[2] - 1.0f ) * e2;
      y[3] += ( 4.0f * l[3] - 1.0f ) * e3;
      y[4] += 4.0f * ( l[1] * e0 + l[0] * e1 );
      y[5] += 4.0f * ( l[2] * e0 + l[0] * e2 );
      y[6] += 4.0f * ( l[3] * e0 + l[0] * e3 );
      y[7] += 4.0f * ( l[2] * e1 + l[1] * e2 );
      y[8] += 4.0f * ( l[3] * e1 + l[1] * e3 );
      y[9] += 4.0f * ( l[3] * e2 + l[2] * e3 );
   }
}

// Reduced-form apply, FP32 arithmetic on FP32 or FP16 values.  A tile's source
// box is staged in LDS, the blended positions of its 5x5x5 micro-vertices
// (relative to its first one, from FP64 evaluations) too; every thread forms its
// micro-cell's metric from those, integrates, and sums into its wavefront's
// LDS result box (waveBoxAccumulate).  FP32
// destinations receive the box with FP32 atomics.  FP16 destinations keep every
// sum in FP32: box-interior points belong to this tile alone and are stored once
// (or added to, rmw), box-surface points (a lattice coordinate divisible by 8,
// shared with neighbouring tiles) are added with FP32 atomics into `accum`,
// which kP2FinishAccum rounds into the FP16 destination afterwards.
// blended positions of a tile's 5x5x5 micro-vertices relative to its first one
// (FP64), staged in LDS by the first 125 threads
__device__ __forceinline__ void stageBlendedVertices( const P2CellGeom& g, const TileCoord& tc, double ( *sv )[3] )
{
   constexpr int nv = ( kTile + 1 ) * ( kTile + 1 ) * ( kTile + 1 );
   if ( threadIdx.x >= nv )
      return;
   auto blended = [&]( int i, int j, int k, double F[3] ) {
      const double vx = double( tc.x0 + i ), vy = double( tc.y0 + j ), vz = double( tc.z0 + k );
      double       x[3];
#pragma unroll
      for ( int d = 0; d < 3; ++d )
         x[d] = g.v0[d] + g.hE[3 * d] * vx + g.hE[3 * d + 1] * vy + g.hE[3 * d + 2] * vz;
      double q = 1.0;
      if ( g.kind )
      {
         const double ls = g.nl[0] * x[0] + g.nl[1] * x[1] + g.nl[2] * x[2] - g.lam0;
         q               = ( g.r0 + fabs( ls ) * g.dr ) / sqrt( x[0] * x[0] + x[1] * x[1] + x[2] * x[2] );
      }
#pragma unroll
      for ( int d = 0; d < 3; ++d )
         F[d] = q * x[d];
   };
   const int i = threadIdx.x % ( kTile + 1 ), j = ( threadIdx.x / ( kTile + 1 ) ) % ( kTile + 1 ),
             k = threadIdx.x / ( ( kTile + 1 ) * ( kTile + 1 ) );
   double F[3], F0[3];
   blended( i, j, k, F );
   blended( 0, 0, 0, F0 );
#pragma unroll
   for ( int d = 0; d < 3; ++d )
      sv[threadIdx.x][d] = F[d] - F0[d];
}

// the reduced-form metric of micro-cell (type t, anchor ax, ay, az) from the
// staged vertices: edge vectors c_k = P_k - P_0 of the blended vertices, n_k
// their cross products (det J * rows of J^-1), G = (k w / |det J|) n_i . n_j,
// formed in FP64 (MI355X issues FP64 and non-packed FP32 at one rate) and
// rounded once
__device__ __forceinline__ void reducedMetric( const double ( *sv )[3], int t, int ax, int ay, int az, float ke,
                                               float G[6] )
{
   double P[4][3];
#pragma unroll
   for ( int k = 0; k < 4; ++k )
   {
      const int vi = ( ax + cVof[t][k][0] ) + ( kTile + 1 ) * ( ( ay + cVof[t][k][1] ) + ( kTile + 1 ) * ( az + cVof[t][k][2] ) );
#pragma unroll
      for ( int d = 0; d < 3; ++d )
         P[k][d] = sv[vi][d];
   }
   double c[3][3], n[3][3];
#pragma unroll
   for ( int k = 0; k < 3; ++k )
#pragma unroll
      for ( int d = 0; d < 3; ++d )
         c[k][d] = P[k + 1][d] - P[0][d];
#pragma unroll
   for ( int k = 0; k < 3; ++k )
   {
      const double* b1 = c[( k + 1 ) % 3];
      const double* b2 = c[( k + 2 ) % 3];
      n[k][0]          = b1[1] * b2[2] - b1[2] * b2[1];
      n[k][1]          = b1[2] * b2[0] - b1[0] * b2[2];
      n[k][2]          = b1[0] * b2[1] - b1[1] * b2[0];
   }
   const double det = c[0][0] * n[0][0] + c[0][1] * n[0][1] + c[0][2] * n[0][2];
   const double s   = ( double( ke ) * ( 1.0 / 24.0 ) ) / fabs( det );
   G[0]             = float( s * ( n[0][0] * n[0][0] + n[0][1] * n[0][1] + n[0][2] * n[0][2] ) );
   G[1]             = float( s * ( n[0][0] * n[1][0] + n[0][1] * n[1][1] + n[0][2] * n[1][2] ) );
   G[2]             = float( s * ( n[0][0] * n[2][0] + n[0][1] * n[2][1] + n[0][2] * n[2][2] ) );
   G[3]             = float( s * ( n[1][0] * n[1][0] + n[1][1] * n[1][1] + n[1][2] * n[1][2] ) );
   G[4]             = float( s * ( n[1][0] * n[2][0] + n[1][1] * n[2][1] + n[1][2] * n[2][2] ) );
   G[5]             = float( s * ( n[2][0] * n[2][0] + n[2][1] * n[2][1] + n[2][2] * n[2][2] ) );
}

// the reduced-form metric of every micro-cell of a level, stored once
// (tile-major [tile][6][384], FP32): the reduced applies then read 24 bytes per
// micro-cell instead of forming the blended metric in every apply
__global__ __launch_bounds__( kTileThreads ) void kP2TileReducedMetric( const P2CellGeom* __restrict__ geom,
                                                                        const __half* __restrict__ kc, int64_t mm,
                                                                        int64_t ntiles,
                                                                        const uint64_t* __restrict__ tiles,
                                                                        float* __restrict__ gm )
{
   __shared__ double sv[( kTile + 1 ) * ( kTile + 1 ) * ( kTile + 1 )][3];
   const int64_t    lb = xcdBlock( ntiles );
   if ( lb >= ntiles )
      return;
   int64_t         full;
   const TileCoord tc = tiles ? tileFromList( tiles, lb, mm, full ) : tileOf( lb, mm );
   const float     ke = __half2float( kc[lb * kTileThreads + threadIdx.x] );
   stageBlendedVertices( geom[tc.c], tc, sv );
   __syncthreads();
   const int t = __builtin_amdgcn_readfirstlane( int( threadIdx.x ) >> 6 );
   const int a = int( threadIdx.x ) & 63;
   float     G[6];
   reducedMetric( sv, t, a & 3, ( a >> 2 ) & 3, a >> 4, ke, G );
#pragma unroll
   for ( int i = 0; i < 6; ++i )
      gm[( lb * 6 + i ) * kTileThreads + threadIdx.x] = G[i];
}

template < class T, bool STORED_G >
__global__ __launch_bounds__( kTileThreads ) void kP2TileApplyReduced( const P2CellGeom* __restrict__ geom,
                                                                      const __half* __restrict__ kc,
                                                                      const float* __restrict__ gm, int64_t mm,
                                                                      int64_t ntiles, int64_t sc1, int64_t w1,
                                                                      const T* __restrict__ src, T* __restrict__ dst,
                                                                      float* __restrict__ accum, int rmw, float alpha,
                                                                      const uint64_t* __restrict__ tiles )
{
   __shared__ float  sb[kBox3];
   __shared__ double sv[STORED_G ? 1 : ( kTile + 1 ) * ( kTile + 1 ) * ( kTile + 1 )][3];
   __shared__ float  sy[kWaveBoxes * kBox3];
   const int64_t    lb = xcdBlock( ntiles );
   if ( lb >= ntiles )
      return;
   int64_t         full;
   const TileCoord tc   = tiles ? tileFromList( tiles, lb, mm, full ) : tileOf( lb, mm );
   const int64_t   base = tc.c * sc1;
   const int32_t   top  = int32_t( 2 * mm );
   float           G[6];
   float           ke = 0.0f;
   if constexpr ( STORED_G )
   {
#pragma unroll
      for ( int i = 0; i < 6; ++i )
         G[i] = gm[( lb * 6 + i ) * kTileThreads + threadIdx.x];
   }
   else
      ke = __half2float( kc[lb * kTileThreads + threadIdx.x] );
   const T* srcC = src + base;
   {
      // the thread's (at most two) box points, both loads in flight before the LDS writes
      T bv[2];
#pragma unroll
      for ( int j = 0; j < 2; ++j )
      {
         const int     p  = int( threadIdx.x ) + j * kTileThreads;
         const int32_t lx = p % kBox, ly = ( p / kBox ) % kBox, lz = p / ( kBox * kBox );
         const int32_t X = 2 * int32_t( tc.x0 ) + lx, Y = 2 * int32_t( tc.y0 ) + ly, Z = 2 * int32_t( tc.z0 ) + lz;
         bv[j]           = ( p < kBox3 && X + Y + Z <= top ) ? srcC[tetRowStart32( int32_t( w1 ), Y, Z ) + X] : T( 0 );
      }
#pragma unroll
      for ( int j = 0; j < 2; ++j )
      {
         const int p = int( threadIdx.x ) + j * kTileThreads;
         if ( p < kBox3 )
            sb[p] = float( toAcc( bv[j] ) );
      }
   }
   if constexpr ( !STORED_G )
      stageBlendedVertices( geom[tc.c], tc, sv );
   __syncthreads();
   const int t  = __builtin_amdgcn_readfirstlane( int( threadIdx.x ) >> 6 );
   const int a  = int( threadIdx.x ) & 63;
   const int ax = a & 3, ay = ( a >> 2 ) & 3, az = a >> 4;
   const bool valid = tc.x0 + ax + tc.y0 + ay + tc.z0 + az <= mm - 1 - MicroCells::rowShift[t];
   float u[10];
   int   pos[10];
#pragma unroll
   for ( int k = 0; k < 4; ++k )
      pos[k] = 2 * ( ax + cVof[t][k][0] ) + kBox * ( 2 * ( ay + cVof[t][k][1] ) + kBox * 2 * ( az + cVof[t][k][2] ) );
#pragma unroll
   for ( int e = 0; e < 6; ++e )
      pos[4 + e] = ( pos[cP2E[e][0]] + pos[cP2E[e][1]] ) / 2;
#pragma unroll
   for ( int i = 0; i < 10; ++i )
      u[i] = valid ? sb[pos[i]] : 0.0f;
   if constexpr ( !STORED_G )
      reducedMetric( sv, t, ax, ay, az, ke, G );
   {
      float y[10];
#pragma unroll
      for ( int i = 0; i < 10; ++i )
         y[i] = 0.0f;
      if ( valid )
         p2ReducedAction( G, u, y );
      waveBoxAccumulate( sy, valid, pos, y );
   }
   __syncthreads();
   T* dstC = dst + base;
   for ( int p = threadIdx.x; p < kBox3; p += kTileThreads )
   {
      const int32_t lx = p % kBox, ly = ( p / kBox ) % kBox, lz = p / ( kBox * kBox );
      const int32_t X = 2 * int32_t( tc.x0 ) + lx, Y = 2 * int32_t( tc.y0 ) + ly, Z = 2 * int32_t( tc.z0 ) + lz;
      if ( X + Y + Z > top )
         continue;
      const int64_t o = tetRowStart32( int32_t( w1 ), Y, Z ) + X;
      const float   v = alpha * waveBoxSum( sy, p );
      if constexpr ( std::is_same< T, float >::value )
         atomicAdd( dstC + o, v );
      else
      {
         const bool surface = lx == 0 || ly == 0 || lz == 0 || lx == kBox - 1 || ly == kBox - 1 || lz == kBox - 1;
         if ( surface )
            atomicAdd( accum + base + o, v );
         else
            dstC[o] = __float2half( rmw ? __half2float( dstC[o] ) + v : v );
      }
   }
}

// FP16 destinations of kP2TileApplyReduced: the tile-surface points (a lattice
// coordinate divisible by 8) receive their FP32 sums, which are reset to zero.
// One thread per (surface point of the list, cell).
__global__ void kP2FinishAccum( const int32_t* __restrict__ pts, int64_t npts, int64_t sc1, int rmw,
                                float* __restrict__ accum, __half* __restrict__ dst )
{
   const int64_t i = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( i >= npts )
      return;
   const int64_t o = int64_t( blockIdx.y ) * sc1 + pts[i];
   const float   v = accum[o];
   accum[o]        = 0.0f;
   dst[o]          = __float2half( rmw ? __half2float( dst[o] ) + v : v );
}

// the lattice's list of tile-surface points (built on first use)
static void surfacePointList( Context* ctx, Lattice& l )
{
   if ( l.surfacePoints.p )
      return;
   constexpr int32_t       s = 2 * kTile;
   const int32_t           w = int32_t( l.w );
   std::vector< int32_t > pts;
   for ( int32_t Z = 0; Z < w; ++Z )
      for ( int32_t Y = 0; Y + Z < w; ++Y )
         for ( int32_t X = 0; X + Y + Z < w; ++X )
            if ( X % s == 0 || Y % s == 0 || Z % s == 0 )
               pts.push_back( tetRowStart32( w, Y, Z ) + X );
   l.surfacePoints.upload( ctx, pts );
}

// y = A_e u for the micro-cell of th: the element action of
// p2_div_k_grad_blending_q4 (11-point rule); kq(q) gives k at quadrature point q
// of the micro-cell.
//
// With barycentric coordinates lam_0..3 (lam_0 = 1 - lam_1 - lam_2 - lam_3), the
// P2 function u = sum_k u_k lam_k (2 lam_k - 1) + sum_(i<j) 4 u_ij lam_i lam_j has
// the computational-space gradient  g = sum_(k=1..3) d_k grad lam_k  with
// d_k = du/dlam_k - du/dlam_0, which is affine in (lam_1, lam_2, lam_3):
//   d = d0 + 4 B lam',  d0_k = 4 u_0k - 3 u_0 - u_k,
//   B_kk = u_k + u_0 - 2 u_0k,  B_kj = u_kj - u_0j - u_0k + u_0 (k != j),
// so g(q) = v0 + V lam'_q with v0 = Dc^T d0 and V = 4 Dc^T B, formed once per
// micro-cell.  The flux m_q = w_q k_q |det DF J| (DF^-1 DF^-T) g(q) enters the
// result only through Tm = sum_q m_q lam'_q^T and Sm = sum_q m_q; with
// T = Dc Tm, S = Dc Sm (the sums of lam'_j (grad lam_k . m) and grad lam_k . m),
// the element result is the transpose of the map u -> d applied to them:
//   y_0 = 4 sum(T) - 3 sum(S),  y_k = 4 T_kk - S_k,
//   y_0k = 4 (S_k - sum_j T_kj - sum_j T_jk),  y_kj = 4 (T_kj + T_jk).
// Per quadrature point that leaves 9 FMA for g and 12 for the sums, besides
// the blending.  The shell map F(x) = q(x) x has DF = q I + x w^T (w = grad q),
// inverted in closed form (Sherman-Morrison); det DF = q^2 beta with
// beta = q + w . x = s (n . x), s = sign(lam) dr / |x|, so the powers of q cancel.
template < class R, class KQ >
__device__ __forceinline__ void p2ElementAction( const P2CellGeomT< R >& g, const TileThread< R >& th, const R u[10],
                                                 R y[10], KQ&& kq )
{
   using QC      = QuadConst< R >;
   const R*  Dc  = g.Dc[th.t];
   const R   adet = g.adet[th.t];
   R         v0[3], V[3][3];
   {
      // d0 and B (symmetric) from the element values (edges 4..9 in the order
      // (0,1), (0,2), (0,3), (1,2), (1,3), (2,3))
      const R d0[3] = { R( 4 ) * u[4] - R( 3 ) * u[0] - u[1], R( 4 ) * u[5] - R( 3 ) * u[0] - u[2],
                        R( 4 ) * u[6] - R( 3 ) * u[0] - u[3] };
      const R b00 = u[1] + u[0] - R( 2 ) * u[4], b11 = u[2] + u[0] - R( 2 ) * u[5], b22 = u[3] + u[0] - R( 2 ) * u[6];
      const R b01 = u[7] - u[5] - u[4] + u[0], b02 = u[8] - u[6] - u[4] + u[0], b12 = u[9] - u[6] - u[5] + u[0];
      const R B[3][3] = { { b00, b01, b02 }, { b01, b11, b12 }, { b02, b12, b22 } };
#pragma unroll
      for ( int d = 0; d < 3; ++d )
      {
         v0[d] = Dc[d] * d0[0] + Dc[3 + d] * d0[1] + Dc[6 + d] * d0[2];
#pragma unroll
         for ( int j = 0; j < 3; ++j )
            V[d][j] = R( 4 ) * ( Dc[d] * B[0][j] + Dc[3 + d] * B[1][j] + Dc[6 + d] * B[2][j] );
      }
   }
   R Tm[3][3], Sm[3];
#pragma unroll
   for ( int d = 0; d < 3; ++d )
   {
      Sm[d] = R( 0 );
#pragma unroll
      for ( int j = 0; j < 3; ++j )
         Tm[d][j] = R( 0 );
   }
#pragma unroll 1
   for ( int q = 0; q < 11; ++q )
   {
      const R* lq = QC::lam( q );
      R        v[3];
#pragma unroll
      for ( int d = 0; d < 3; ++d )
         v[d] = v0[d] + V[d][0] * lq[0] + V[d][1] * lq[1] + V[d][2] * lq[2];
      const R* xq = g.XQ[th.t][q];
      R        cc = QC::w()[q] * adet * kq( q );
      if ( g.kind )
      {
         const R x[3] = { th.p0[0] + xq[0], th.p0[1] + xq[1], th.p0[2] + xq[2] };
         const R ls   = th.ls0 + xq[3];
         const R r2   = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
         const R irho = fastRsqrt( r2 );
         const R qq   = ( g.r0 + fabs( ls ) * g.dr ) * irho;
         const R s    = copysign( g.dr, ls ) * irho;
         const R qr2  = qq * ( irho * irho );
         R       w[3];
#pragma unroll
         for ( int d = 0; d < 3; ++d )
            w[d] = s * g.nl[d] - qr2 * x[d];
         const R beta = s * ( ls + g.lam0 );
         const R ib   = fastRcp( beta );
         const R t1   = ( x[0] * v[0] + x[1] * v[1] + x[2] * v[2] ) * ib;
#pragma unroll
         for ( int d = 0; d < 3; ++d )
            v[d] -= w[d] * t1;
         const R t2 = ( w[0] * v[0] + w[1] * v[1] + w[2] * v[2] ) * ib;
#pragma unroll
         for ( int d = 0; d < 3; ++d )
            v[d] -= x[d] * t2;
         cc *= fabs( beta );
      }
#pragma unroll
      for ( int d = 0; d < 3; ++d )
      {
         const R m = cc * v[d];
         Sm[d] += m;
#pragma unroll
         for ( int j = 0; j < 3; ++j )
            Tm[d][j] += lq[j] * m;
      }
   }
   R T[3][3], S[3];
#pragma unroll
   for ( int k = 0; k < 3; ++k )
   {
      S[k] = Dc[3 * k] * Sm[0] + Dc[3 * k + 1] * Sm[1] + Dc[3 * k + 2] * Sm[2];
#pragma unroll
      for ( int j = 0; j < 3; ++j )
         T[k][j] = Dc[3 * k] * Tm[0][j] + Dc[3 * k + 1] * Tm[1][j] + Dc[3 * k + 2] * Tm[2][j];
   }
   const R rs[3] = { T[0][0] + T[0][1] + T[0][2], T[1][0] + T[1][1] + T[1][2], T[2][0] + T[2][1] + T[2][2] };
   const R cs[3] = { T[0][0] + T[1][0] + T[2][0], T[0][1] + T[1][1] + T[2][1], T[0][2] + T[1][2] + T[2][2] };
   y[0]          = R( 4 ) * ( rs[0] + rs[1] + rs[2] ) - R( 3 ) * ( S[0] + S[1] + S[2] );
#pragma unroll
   for ( int k = 0; k < 3; ++k )
   {
      y[1 + k] = R( 4 ) * T[k][k] - S[k];
      y[4 + k] = R( 4 ) * ( S[k] - rs[k] - cs[k] );
   }
   y[7] = R( 4 ) * ( T[0][1] + T[1][0] );
   y[8] = R( 4 ) * ( T[0][2] + T[2][0] );
   y[9] = R( 4 ) * ( T[1][2] + T[2][1] );
}

// y = A x on the cell lattices of level L+1 (halo positions included): the
// element action of p2_div_k_grad_blending_q4 evaluated in barycentric form,
//   gamma_k = sum_j u_j dphi_j/dlam_k,  g = sum_k gamma_k grad_c lam_k,
//   y_i += sum_q w_q k_q |det DF J| dphi_i/dlam_k grad_c lam_k . (DF^-1 DF^-T g)
// (grad_c: computational space), equal to the reference's element matrix times
// the element vector.
// T: value type of src / dst; K: value type of the coefficient table; the
// arithmetic runs in Acc<T> (FP64 for double, FP32 for float and half).
// BATCH: 2^lv interleaved vectors (vector blockIdx.y, its coefficient table at
// table + blockIdx.y * tstride); otherwise one vector.
template < class T, class K, bool STORED, bool BATCH = false >
__device__ __forceinline__ void p2TileApplyBody( const P2CellGeomT< typename Acc< T >::type >* __restrict__ geom,
                                                 const P2CellGeom* __restrict__ geom64, const KParams* __restrict__ kp,
                                                 const K* __restrict__ table, int64_t mm, int64_t ntiles, int64_t sc1,
                                                 int64_t w1, const T* __restrict__ src, T* __restrict__ dst,
                                                 double alpha, const uint64_t* __restrict__ tiles, int lvB = 0,
                                                 int64_t tstride = 0, const int32_t* __restrict__ on = nullptr )
{
   const int lv = BATCH ? lvB : 0;
   if constexpr ( BATCH )
   {
      if ( on && !on[blockIdx.y] )
         return; // an inactive slot: its (zeroed) result stays zero
      table += int64_t( blockIdx.y ) * tstride;
      src += blockIdx.y;
      dst += blockIdx.y;
   }
   using R  = typename Acc< T >::type;
   using QC = QuadConst< R >;
   // Element results are summed with LDS atomics into a result box of their own
   // in FP64 (cleared while the loads are in flight, so a tile needs two
   // workgroup barriers), and in per-wavefront result boxes (waveBoxAccumulate)
   // in FP32, where the LDS atomics cost about half the kernel time (in FP64 the
   // per-wavefront boxes are not faster).  LDS: the tile's source box and the
   // staged coefficients (+ the FP64 result box; the FP32 result boxes reuse the
   // staging space once every wavefront has integrated).
   constexpr bool kAtomic = std::is_same< R, double >::value;
   constexpr int  kStage  = kBox3 + ( STORED ? 11 * kTileThreads : 0 );
   constexpr int  kLds    = kAtomic ? kStage + kBox3 : ( kStage > kWaveBoxes * kBox3 ? kStage : kWaveBoxes * kBox3 );
   __shared__ R   lds[kLds];
   R* const       sb = lds;
   R* const       sk = lds + kBox3;
   R* const       sa = lds + kStage; // FP64: the result box, apart from the source box
   const int64_t li = xcdBlock( ntiles );
   if ( li >= ntiles )
      return;
   int64_t         lb;
   const TileCoord tc   = tiles ? tileFromList( tiles, li, mm, lb ) : tileOfInterior( li, mm, lb );
   const int64_t   base = tc.c * sc1;
   const int32_t   top  = int32_t( 2 * mm );
   // the coefficient loads are issued first (their latency overlaps the box
   // load) and parked in LDS, which keeps them out of the registers of the
   // quadrature loop
   // and the thread's (at most two) source-box values right behind them, so
   // that the two load latencies overlap
   static_assert( kBox3 <= 2 * kTileThreads, "two box points per thread" );
   K kv[STORED ? 11 : 1];
   if constexpr ( STORED )
   {
#pragma unroll
      for ( int q = 0; q < 11; ++q )
         kv[q] = table[( lb * 11 + q ) * kTileThreads + threadIdx.x];
   }
   const T* srcC = src + ( base << lv );
   T        bv[2];
   bool     bin[2];
#pragma unroll
   for ( int j = 0; j < 2; ++j )
   {
      const int     p  = int( threadIdx.x ) + j * kTileThreads;
      const int32_t lx = p % kBox, ly = ( p / kBox ) % kBox, lz = p / ( kBox * kBox );
      const int32_t X = 2 * int32_t( tc.x0 ) + lx, Y = 2 * int32_t( tc.y0 ) + ly, Z = 2 * int32_t( tc.z0 ) + lz;
      bin[j]          = p < kBox3 && X + Y + Z <= top;
      bv[j]           = bin[j] ? srcC[int64_t( tetRowStart32( int32_t( w1 ), Y, Z ) + X ) << lv] : T( 0 );
   }
   if constexpr ( kAtomic )
   {
      // the result box is cleared while the loads are in flight
#pragma unroll
      for ( int j = 0; j < 2; ++j )
      {
         const int p = int( threadIdx.x ) + j * kTileThreads;
         if ( p < kBox3 )
            sa[p] = R( 0 );
      }
   }
   if constexpr ( STORED )
   {
#pragma unroll
      for ( int q = 0; q < 11; ++q )
         sk[q * kTileThreads + threadIdx.x] = R( toAcc( kv[q] ) );
   }
#pragma unroll
   for ( int j = 0; j < 2; ++j )
   {
      const int p = int( threadIdx.x ) + j * kTileThreads;
      if ( p < kBox3 )
         sb[p] = bin[j] ? R( toAcc( bv[j] ) ) : R( 0 );
   }
   __syncthreads();
   const P2CellGeomT< R >& g  = geom[tc.c];
   const TileThread< R >   th = tileThread( g, tc, mm );
   R                       u[10];
   {
      int pos[10];
      elementBoxPositions( th, pos );
#pragma unroll
      for ( int i = 0; i < 10; ++i )
         u[i] = th.valid ? sb[pos[i]] : R( 0 );
   }
   R y[10];
   if ( th.valid )
   {
      p2ElementAction( g, th, u, y, [&]( int q ) -> R {
         if constexpr ( STORED )
            return sk[q * kTileThreads + threadIdx.x];
         else
         {
            // k from the FP64 geometry of the same point
            const P2CellGeom&          g64  = geom64[tc.c];
            const TileThread< double > t64  = tileThread( g64, tc, mm );
            const double*              xq64 = g64.XQ[th.t][q];
            const double x64[3] = { t64.p0[0] + xq64[0], t64.p0[1] + xq64[1], t64.p0[2] + xq64[2] };
            return R( kAt( g64, *kp, x64, t64.ls0 + xq64[3] ) );
         }
      } );
   }
   {
      int pos[10];
      elementBoxPositions( th, pos );
      if constexpr ( kAtomic )
      {
         if ( th.valid )
         {
#pragma unroll
            for ( int i = 0; i < 10; ++i )
               atomicAdd( &sa[pos[i]], y[i] );
         }
      }
      else
      {
         __syncthreads(); // the source box and the coefficients are no longer read
         waveBoxAccumulate( lds, th.valid, pos, y );
      }
   }
   __syncthreads();
   // every box point is added with an atomic that returns nothing: box-surface
   // points are shared with neighbouring tiles, and a read-modify-write of the
   // interior points would make the wave wait for its outstanding atomics too
   // (the vector-memory counter covers loads and atomics alike)
   T* dstC = dst + ( base << lv );
   for ( int p = threadIdx.x; p < kBox3; p += kTileThreads )
   {
      const int32_t lx = p % kBox, ly = ( p / kBox ) % kBox, lz = p / ( kBox * kBox );
      const int32_t X = 2 * int32_t( tc.x0 ) + lx, Y = 2 * int32_t( tc.y0 ) + ly, Z = 2 * int32_t( tc.z0 ) + lz;
      if ( X + Y + Z <= top )
         sharedAdd( dstC + ( int64_t( tetRowStart32( int32_t( w1 ), Y, Z ) + X ) << lv ),
                    R( alpha ) * ( kAtomic ? sa[p] : waveBoxSum( lds, p ) ) );
   }
}

// The last tile layer of a macro-cell: one wavefront per tile, one lane per
// micro-cell inside the macro-cell (cDiagPairs), the type varying over the lanes;
// same element action as the tile kernels, box staging, accumulation and
// write-back over the 165 box points inside the macro-cell (cDiagBox).
template < class T, class K, bool STORED >
__global__ __launch_bounds__( 64 ) void kP2TileApplyDiag( const P2CellGeomT< typename Acc< T >::type >* __restrict__ geom,
                                                          const P2CellGeom* __restrict__ geom64,
                                                          const KParams* __restrict__ kp, const K* __restrict__ table,
                                                          int64_t mm, int64_t ndiag, int64_t sc1, int64_t w1,
                                                          const T* __restrict__ src, T* __restrict__ dst, double alpha,
                                                          const uint64_t* __restrict__ tiles, int lv = 0,
                                                          int64_t tstride = 0, const int32_t* __restrict__ on = nullptr )
{
   using R = typename Acc< T >::type;
   if ( on && !on[blockIdx.y] )
      return; // an inactive slot: its (zeroed) result stays zero
   // 2^lv interleaved vectors: vector blockIdx.y with its own coefficient table
   table += int64_t( blockIdx.y ) * tstride;
   src += blockIdx.y;
   dst += blockIdx.y;
   __shared__ R  sb[kBox3];
   const int64_t ld = xcdBlock( ndiag );
   if ( ld >= ndiag )
      return;
   int64_t         lb;
   const TileCoord tc   = tiles ? tileFromList( tiles, ld, mm, lb ) : tileOfDiagonal( ld, mm, lb );
   const int64_t   base = tc.c * sc1;
   const T*        srcC = src + ( base << lv );
   // the lane's (at most three) of the 165 box points inside the macro-cell:
   // all loads issued before the first LDS write
   constexpr int kPer = kDiagBoxP / 64;
   int           bp[kPer];
   T             bv[kPer];
#pragma unroll
   for ( int j = 0; j < kPer; ++j )
   {
      bp[j]           = cDiagBox[int( threadIdx.x ) + 64 * j];
      const int     p  = bp[j] < 0 ? 0 : bp[j];
      const int32_t lx = p % kBox, ly = ( p / kBox ) % kBox, lz = p / ( kBox * kBox );
      const int32_t X = 2 * int32_t( tc.x0 ) + lx, Y = 2 * int32_t( tc.y0 ) + ly, Z = 2 * int32_t( tc.z0 ) + lz;
      bv[j]           = bp[j] >= 0 ? srcC[int64_t( tetRowStart32( int32_t( w1 ), Y, Z ) + X ) << lv] : T( 0 );
   }
#pragma unroll
   for ( int j = 0; j < kPer; ++j )
      if ( bp[j] >= 0 )
         sb[bp[j]] = R( toAcc( bv[j] ) );
   __syncthreads();
   const P2CellGeomT< R >& g    = geom[tc.c];
   const int32_t           pair = cDiagPairs[threadIdx.x];
   const int               t = pair & 7, a = pair >> 3;
   const TileThread< R >   th = tileThreadAt( g, tc, mm, t, a );
   int                     pos[10];
   elementBoxPositions( th, pos );
   R u[10];
#pragma unroll
   for ( int i = 0; i < 10; ++i )
      u[i] = th.valid ? sb[pos[i]] : R( 0 );
   __syncthreads();
#pragma unroll
   for ( int j = 0; j < kPer; ++j )
      if ( bp[j] >= 0 )
         sb[bp[j]] = R( 0 );
   __syncthreads();
   if ( th.valid )
   {
      R y[10];
      p2ElementAction( g, th, u, y, [&]( int q ) -> R {
         if constexpr ( STORED )
            return R( toAcc( table[( lb * 11 + q ) * kTileThreads + t * 64 + a] ) );
         else
         {
            const P2CellGeom&          g64  = geom64[tc.c];
            const TileThread< double > t64  = tileThreadAt( g64, tc, mm, t, a );
            const double*              xq64 = g64.XQ[t][q];
            const double x64[3] = { t64.p0[0] + xq64[0], t64.p0[1] + xq64[1], t64.p0[2] + xq64[2] };
            return R( kAt( g64, *kp, x64, t64.ls0 + xq64[3] ) );
         }
      } );
#pragma unroll
      for ( int i = 0; i < 10; ++i )
         atomicAdd( &sb[pos[i]], y[i] );
   }
   __syncthreads();
   T* dstC = dst + ( base << lv );
#pragma unroll
   for ( int j = 0; j < kPer; ++j )
   {
      if ( bp[j] < 0 )
         continue;
      const int     p  = bp[j];
      const int32_t lx = p % kBox, ly = ( p / kBox ) % kBox, lz = p / ( kBox * kBox );
      const int32_t X = 2 * int32_t( tc.x0 ) + lx, Y = 2 * int32_t( tc.y0 ) + ly, Z = 2 * int32_t( tc.z0 ) + lz;
      sharedAdd( dstC + ( int64_t( tetRowStart32( int32_t( w1 ), Y, Z ) + X ) << lv ), R( alpha ) * sb[p] );
   }
}

// stored coefficients, the compiler's own register budget: the FP64 kernel
// holds 94 VGPRs (5 waves per SIMD) without spills; a budget of 80 (6 waves)
// spills the micro-cell's vertex and shell coordinate inside the quadrature loop
template < class T, class K >
__global__ __launch_bounds__( kTileThreads ) void kP2TileApplyStored(
    const P2CellGeomT< typename Acc< T >::type >* __restrict__ geom, const P2CellGeom* __restrict__ geom64,
    const KParams* __restrict__ kp, const K* __restrict__ table, int64_t mm, int64_t ntiles, int64_t sc1, int64_t w1,
    const T* __restrict__ src, T* __restrict__ dst, double alpha, const uint64_t* __restrict__ tiles )
{
   p2TileApplyBody< T, K, true >( geom, geom64, kp, table, mm, ntiles, sc1, w1, src, dst, alpha, tiles );
}
// the same over 2^lv interleaved vectors with one coefficient table each
// (grid.y = vectors)
template < class T, class K >
__global__ __launch_bounds__( kTileThreads ) void kP2TileApplyStoredBatched(
    const P2CellGeomT< typename Acc< T >::type >* __restrict__ geom, const K* __restrict__ table, int64_t tstride,
    int64_t mm, int64_t ntiles, int64_t sc1, int64_t w1, const T* __restrict__ src, T* __restrict__ dst, double alpha,
    const uint64_t* __restrict__ tiles, int lv, const int32_t* __restrict__ on )
{
   p2TileApplyBody< T, K, true, true >( geom, nullptr, nullptr, table, mm, ntiles, sc1, w1, src, dst, alpha, tiles, lv,
                                        tstride, on );
}
// coefficients evaluated in place (the compiler's own register budget)
template < class T, class K >
__global__ __launch_bounds__( kTileThreads ) void kP2TileApplyEval(
    const P2CellGeomT< typename Acc< T >::type >* __restrict__ geom, const P2CellGeom* __restrict__ geom64,
    const KParams* __restrict__ kp, const K* __restrict__ table, int64_t mm, int64_t ntiles, int64_t sc1, int64_t w1,
    const T* __restrict__ src, T* __restrict__ dst, double alpha, const uint64_t* __restrict__ tiles )
{
   p2TileApplyBody< T, K, false >( geom, geom64, kp, table, mm, ntiles, sc1, w1, src, dst, alpha, tiles );
}

__global__ void kP2ElementDiag( const CellGeom* __restrict__ geom, const KParams* __restrict__ kp, int64_t mm,
                                int64_t nc, int64_t sc1, int64_t w1, double* __restrict__ dst )
{
   int64_t c, a[3];
   int     t;
   if ( !microOf( int64_t( blockIdx.x ) * blockDim.x + threadIdx.x, mm, nc, c, t, a ) )
      return;
   const CellGeom& g = geom[c];
   double          p[4][3];
   int64_t         pos[10];
   microElement( g, t, a, mm, c, sc1, w1, p, pos );
   double ye[10] = {};
   for ( int q = 0; q < 11; ++q )
   {
      const double lam[4] = { cQP[q][0], cQP[q][1], cQP[q][2], cQP[q][3] };
      double       D[4][3], absdet, F[3], G[10][3];
      elementGeometry( g, p, lam, D, absdet, F );
      p2Gradients( lam, D, G );
      const double wq = cQW[q] * kEval( *kp, F ) * absdet;
      for ( int i = 0; i < 10; ++i )
         ye[i] += wq * ( G[i][0] * G[i][0] + G[i][1] * G[i][1] + G[i][2] * G[i][2] );
   }
   for ( int i = 0; i < 10; ++i )
      atomicAdd( dst + pos[i], ye[i] );
}


static int64_t microCellsPerCell( int64_t mm )
{
   int64_t per = 0;
   for ( int k = 0; k < 6; ++k )
      per += tet( mm - MicroCells::rowShift[k] );
   return per;
}

static void p2CheckFunction( const P2Operator& op, const Function& f, int L )
{
   if ( f.st != op.st || f.off != 1 || f.lv != op.lvSlots )
      throw Error( -1, "P2 operator needs P2 functions on its storage with one vector per coefficient slot" );
   if ( L < op.minL || L > op.maxL || L < f.minL || L > f.maxL )
      throw Error( -1, "level outside the operator's or function's level range" );
}

// Levels 0 and 1 (mm <= 2: one tile per macro-cell and at most 8 micro-cells in
// it, so the tile kernel would keep at most 8 of its 384 threads busy): one
// wavefront per macro-cell.  Lane l < 24 takes micro-cell type l / 4 at the
// anchor (0,0,0), (1,0,0), (0,1,0) or (0,0,1) (the only anchors holding
// micro-cells there); the cell's lattice (at most 35 points) is staged in the
// tile kernel's LDS box layout, and the element action and the coefficient
// table are the tile kernel's.
constexpr int kSmallMaxMM = 2;
template < class T, class K, bool STORED >
__global__ __launch_bounds__( 64 ) void kP2SmallApply( const P2CellGeomT< typename Acc< T >::type >* __restrict__ geom,
                                                       const P2CellGeom* __restrict__ geom64,
                                                       const KParams* __restrict__ kp, const K* __restrict__ table,
                                                       int64_t mm, int64_t nc, int64_t sc1, int64_t w1,
                                                       const T* __restrict__ src, T* __restrict__ dst, double alpha,
                                                       int lv = 0, int64_t tstride = 0,
                                                       const int32_t* __restrict__ on = nullptr )
{
   using R = typename Acc< T >::type;
   if ( on && !on[blockIdx.y] )
      return; // an inactive slot: its (zeroed) result stays zero
   // 2^lv interleaved vectors: vector blockIdx.y with its own coefficient table
   table += int64_t( blockIdx.y ) * tstride;
   src += blockIdx.y;
   dst += blockIdx.y;
   __shared__ R  sb[kBox3];
   const int64_t c = blockIdx.x;
   if ( c >= nc )
      return;
   const TileCoord tc{ c, 0, 0, 0 };
   const int32_t   w    = int32_t( w1 );
   const int32_t   top  = w - 1;
   const T*        srcC = src + ( ( c * sc1 ) << lv );
   for ( int p = threadIdx.x; p < w * w * w; p += 64 )
   {
      const int32_t X = p % w, Y = ( p / w ) % w, Z = p / ( w * w );
      if ( X + Y + Z <= top )
         sb[X + kBox * ( Y + kBox * Z )] = R( toAcc( srcC[int64_t( tetRowStart32( w, Y, Z ) + X ) << lv] ) );
   }
   __syncthreads();
   const int               l = int( threadIdx.x );
   const int               t = l < 24 ? l >> 2 : 0;
   const int               a = ( l & 3 ) == 0 ? 0 : 1 << ( 2 * ( ( l & 3 ) - 1 ) ); // 0, 1, 4, 16
   const P2CellGeomT< R >& g  = geom[c];
   TileThread< R >         th = tileThreadAt( g, tc, mm, t, a );
   th.valid                   = th.valid && l < 24;
   R u[10];
   int pos[10];
   elementBoxPositions( th, pos );
#pragma unroll
   for ( int i = 0; i < 10; ++i )
      u[i] = th.valid ? sb[pos[i]] : R( 0 );
   __syncthreads();
   for ( int p = threadIdx.x; p < w * w * w; p += 64 )
   {
      const int32_t X = p % w, Y = ( p / w ) % w, Z = p / ( w * w );
      if ( X + Y + Z <= top )
         sb[X + kBox * ( Y + kBox * Z )] = R( 0 );
   }
   __syncthreads();
   if ( th.valid )
   {
      R y[10];
      p2ElementAction( g, th, u, y, [&]( int q ) -> R {
         if constexpr ( STORED )
            return R( toAcc( table[( c * 11 + q ) * kTileThreads + 64 * t + a] ) );
         else
         {
            const P2CellGeom&          g64  = geom64[c];
            const TileThread< double > t64  = tileThreadAt( g64, tc, mm, t, a );
            const double*              xq64 = g64.XQ[t][q];
            const double x64[3] = { t64.p0[0] + xq64[0], t64.p0[1] + xq64[1], t64.p0[2] + xq64[2] };
            return R( kAt( g64, *kp, x64, t64.ls0 + xq64[3] ) );
         }
      } );
#pragma unroll
      for ( int i = 0; i < 10; ++i )
         atomicAdd( &sb[pos[i]], y[i] );
   }
   __syncthreads();
   // the wave owns the cell's storage (halo copies included): no other wave adds here
   T* dstC = dst + ( ( c * sc1 ) << lv );
   for ( int p = threadIdx.x; p < w * w * w; p += 64 )
   {
      const int32_t X = p % w, Y = ( p / w ) % w, Z = p / ( w * w );
      if ( X + Y + Z <= top )
      {
         T& o = dstC[int64_t( tetRowStart32( w, Y, Z ) + X ) << lv];
         o    = fromAcc< T >( R( toAcc( o ) ) + R( alpha ) * sb[X + kBox * ( Y + kBox * Z )] );
      }
   }
}

// k at the quadrature points of level L (built on first use when the operator
// stores coefficients; nullptr when it evaluates k in the apply kernel)
// k at the quadrature points of level L in value type K (built on first use
// when the operator stores coefficients; nullptr when it evaluates k in the
// apply kernel)
template < class K >
static const K* p2Coefficients( Context* ctx, P2Operator& op, int L, int dtype )
{
   if ( op.nslot == 1 && ( !op.storeCoefficients || op.nm == 0 ) )
      return nullptr;
   auto it = op.kq.find( { L, dtype } );
   if ( it != op.kq.end() )
      return (const K*) it->second.p;
   const int64_t mm     = int64_t( 1 ) << L;
   const int64_t ntiles = tilesPerCell( mm ) * int64_t( op.st->topo.nc );
   const int64_t per    = ntiles * 11 * kTileThreads;
   auto&         buf    = op.kq[{ L, dtype }];
   buf.alloc( ctx, size_t( per * op.nslot ) * sizeof( K ) );
   op.tableStride[L] = per;
   for ( int sl = 0; sl < op.nslot; ++sl )
      kP2TileCoefficients< K ><<< gridFor( ntiles, 1 ), kTileThreads, 0, ctx->stream >>>(
          op.tgeom.at( L ).p, op.kpar.p + sl, mm, ntiles, (K*) buf.p + sl * per );
   HG_CHECK_HIP( hipGetLastError() );
   return (const K*) buf.p;
}

template < class T >
static const P2CellGeomT< typename Acc< T >::type >* p2Geometry( P2Operator& op, int L )
{
   if constexpr ( std::is_same< typename Acc< T >::type, double >::value )
      return op.tgeom.at( L ).p;
   else
      return op.tgeomF.at( L ).p;
}

// k at the blended centroids of the micro-cells of level L (tile-major, FP16),
// built on first use
static const __half* p2CentroidK( Context* ctx, P2Operator& op, int L )
{
   auto it = op.kc.find( L );
   if ( it != op.kc.end() )
      return it->second.p;
   const int64_t mm     = int64_t( 1 ) << L;
   const int64_t ntiles = tilesPerCell( mm ) * int64_t( op.st->topo.nc );
   auto&         buf    = op.kc[L];
   buf.alloc( ctx, size_t( ntiles ) * kTileThreads );
   kP2TileCentroidK<<< gridFor( ntiles, 1 ), kTileThreads, 0, ctx->stream >>>( op.tgeom.at( L ).p, op.kpar.p, mm, ntiles,
                                                                               buf.p );
   HG_CHECK_HIP( hipGetLastError() );
   return buf.p;
}

// The reduced-form metric table of level L (6 FP32 values per micro-cell slot,
// 14.4 GB for the level-6 shell), built on first use when it stays within
// kReducedMetricMaxBytes; otherwise (null) the applies form the metric in the
// kernel from the blended vertices, with the same values.
constexpr int64_t kReducedMetricMaxBytes = int64_t( 32 ) << 30;
static const float* p2ReducedMetric( Context* ctx, P2Operator& op, Lattice& lat, int L, const __half* kc )
{
   auto it = op.gm.find( L );
   if ( it != op.gm.end() )
      return it->second.p;
   const int64_t mm     = int64_t( 1 ) << L;
   const int64_t ntiles = tilesPerCell( mm ) * int64_t( op.st->topo.nc );
   const int64_t n      = ntiles * 6 * kTileThreads;
   if ( n * int64_t( sizeof( float ) ) > kReducedMetricMaxBytes )
      return nullptr;
   auto& buf = op.gm[L];
   buf.alloc( ctx, size_t( n ) );
   kP2TileReducedMetric<<< xcdGrid( ntiles ), kTileThreads, 0, ctx->stream >>>(
       op.tgeom.at( L ).p, kc, mm, ntiles, tileList( ctx, lat, mm, kTilesAll ), buf.p );
   HG_CHECK_HIP( hipGetLastError() );
   return buf.p;
}

// the reduced-form apply on a tile level (FP32 arithmetic): the cell-lattice
// part; the additive push to the owners follows in p2Apply
template < class T >
static void p2ApplyReduced( Context* ctx, P2Operator& op, Lattice& lat, int L, const T* s, T* d, int flag, int update,
                            double alpha )
{
   const int64_t mm     = int64_t( 1 ) << L;
   const int64_t ntiles = tilesPerCell( mm ) * lat.nc;
   // a Replace that includes the cell interiors overwrites every cell-lattice
   // point; otherwise the halo copies start from zero and the rest is added to
   const int     rmw    = !( update == HG_REPLACE && ( flag & kInner ) );
   Storage&      st     = *op.st;
   if ( rmw && lat.haloPerCell() > 0 )
      kZeroHalos< T ><<< gridFor( lat.haloPerCell() * lat.nc ), kBlock, 0, ctx->stream >>>( BndGeom::of( st, lat ),
                                                                                         lat.nc, d );
   else if ( !rmw && std::is_same< T, float >::value )
      kFillRange< T ><<< gridFor( lat.nc * lat.sc ), kBlock, 0, ctx->stream >>>( lat.nc * lat.sc, d, 0.0 );
   HG_CHECK_HIP( hipGetLastError() );
   const __half* kc    = p2CentroidK( ctx, op, L );
   float*        accum = nullptr;
   if constexpr ( std::is_same< T, __half >::value )
   {
      auto& buf = op.accum[L];
      if ( !buf.p )
         buf.alloc( ctx, size_t( lat.nc * lat.sc ) ); // zero-filled once; every finish pass resets what it read
      accum = buf.p;
      surfacePointList( ctx, lat );
   }
   const uint64_t* tiles = tileList( ctx, lat, mm, kTilesAll );
   const float*    gm    = p2ReducedMetric( ctx, op, lat, L, kc );
   ctx->beginTiming();
   if ( gm )
      kP2TileApplyReduced< T, true ><<< xcdGrid( ntiles ), kTileThreads, 0, ctx->stream >>>(
          op.tgeom.at( L ).p, kc, gm, mm, ntiles, lat.sc, lat.w, s, d, accum, rmw, float( alpha ), tiles );
   else
      kP2TileApplyReduced< T, false ><<< xcdGrid( ntiles ), kTileThreads, 0, ctx->stream >>>(
          op.tgeom.at( L ).p, kc, nullptr, mm, ntiles, lat.sc, lat.w, s, d, accum, rmw, float( alpha ), tiles );
   HG_CHECK_HIP( hipGetLastError() );
   ctx->endTiming();
   if constexpr ( std::is_same< T, __half >::value )
   {
      const int64_t npts = int64_t( lat.surfacePoints.n );
      kP2FinishAccum<<< dim3( unsigned( ( npts + 255 ) / 256 ), unsigned( lat.nc ) ), 256, 0, ctx->stream >>>(
          lat.surfacePoints.p, npts, lat.sc, rmw, accum, d );
      HG_CHECK_HIP( hipGetLastError() );
   }
}

// dst = alpha A src (update Replace) or dst + alpha A src (update Add) on the
// DoFs of type flag: P2ElementwiseOperator::gemv with beta 0 or 1
// (P2ElementwiseOperator.cpp:110-233)
// the apply of a batched operator: FP64 interleaved vectors, vector v with
// the coefficients of slot v, the same kernels (and per-vector results) as the
// single-vector apply with the stored table
static void p2ApplyBatched( Context* ctx, P2Operator& op, Lattice& lat, int L, Function& src, Function& dst, int flag,
                            int update, double alpha )
{
   if ( src.dtype != 0 )
      throw Error( HG_ERR_UNSUPPORT, "batched P2 apply: FP64 functions only" );
   if ( update != HG_REPLACE || !( flag & kInner ) )
      throw Error( HG_ERR_UNSUPPORT, "batched P2 apply: update Replace with Inner in the flag only" );
   Storage&       st = *op.st;
   const int      lv = op.lvSlots;
   const unsigned nv = 1u << lv;
   double*        s  = (double*) src.ptr( L );
   double*        d  = (double*) dst.ptr( L );
   const int64_t  mm = int64_t( 1 ) << L;
   pullHalos( ctx, st, lat, s, 0, lv );
   kFillRange< double ><<< gridFor( ( lat.nc * lat.sc ) << lv ), kBlock, 0, ctx->stream >>>( ( lat.nc * lat.sc ) << lv, d,
                                                                                          0.0 );
   HG_CHECK_HIP( hipGetLastError() );
   const double*    table   = p2Coefficients< double >( ctx, op, L, 0 );
   const int64_t    tstride = op.tableStride.at( L );
   const auto*      geom    = op.tgeom.at( L ).p;
   const int64_t    nint    = tet( tileEdges( mm ) - 1 ) * lat.nc;
   const int64_t    ndiag   = tri( tileEdges( mm ) ) * lat.nc;
   const uint64_t*  tint    = mm > kSmallMaxMM ? tileList( ctx, lat, mm, kTilesInterior ) : nullptr;
   const uint64_t*  tlast   = mm > kSmallMaxMM ? tileList( ctx, lat, mm, kTilesLast ) : nullptr;
   ctx->beginTiming();
   if ( mm <= kSmallMaxMM )
      kP2SmallApply< double, double, true ><<< dim3( unsigned( lat.nc ), nv ), 64, 0, ctx->stream >>>(
          geom, geom, nullptr, table, mm, lat.nc, lat.sc, lat.w, s, d, alpha, lv, tstride, op.slotOn.p );
   else
   {
      if ( mm % kTile != 0 )
         throw Error( HG_ERR_STATE, "P2 tile apply: micro-edge count not a multiple of the tile width" );
      if ( nint > 0 )
         kP2TileApplyStoredBatched< double, double ><<< dim3( xcdGrid( nint ), nv ), kTileThreads, 0, ctx->stream >>>(
             geom, table, tstride, mm, nint, lat.sc, lat.w, s, d, alpha, tint, lv, op.slotOn.p );
      HG_CHECK_HIP( hipGetLastError() );
      kP2TileApplyDiag< double, double, true ><<< dim3( xcdGrid( ndiag ), nv ), 64, 0, ctx->stream >>>(
          geom, geom, nullptr, table, mm, ndiag, lat.sc, lat.w, s, d, alpha, tlast, lv, tstride, op.slotOn.p );
   }
   HG_CHECK_HIP( hipGetLastError() );
   ctx->endTiming();
   pushAdditive( ctx, st, lat, d, 0, kAll ^ flag, true, lv );
}

void p2Apply( Context* ctx, P2Operator& op, Function& src, Function& dst, int L, int flag, int update,
              double alpha = 1.0 )
{
   p2CheckFunction( op, src, L );
   p2CheckFunction( op, dst, L );
   if ( src.dtype != dst.dtype )
      throw Error( HG_ERR_ARG, "apply needs src and dst of one value type" );
   if ( &src == &dst )
      throw Error( -1, "apply needs distinct src and dst" );
   Storage&      st  = *op.st;
   Lattice&      lat = st.lattice( ctx, L + 1 );
   void*         s   = src.ptr( L );
   void*         d   = dst.ptr( L );
   const int64_t mm  = int64_t( 1 ) << L;
   if ( lat.w > 1024 )
      throw Error( HG_ERR_UNSUPPORT, "P2 apply supports levels up to 9 (32-bit lattice offsets)" );
   if ( op.nslot > 1 )
   {
      p2ApplyBatched( ctx, op, lat, L, src, dst, flag, update, alpha );
      return;
   }
   pullHalos( ctx, st, lat, s, src.dtype );
   dispatchType( src.dtype, [&]( auto tag ) {
      using T = decltype( tag );
      // FP64 functions use an FP64 table, FP32 an FP32 one, FP16 an FP16 one
      using K = T;
      if constexpr ( !std::is_same< T, double >::value )
      {
         if ( op.reduced && L >= op.reducedFrom && mm > kSmallMaxMM )
         {
            p2ApplyReduced< T >( ctx, op, lat, L, (const T*) s, (T*) d, flag, update, alpha );
            return;
         }
      }
      // zero the halos (all) and, for Replace with Inner in flag, the cell interiors
      if ( update == HG_REPLACE && ( flag & kInner ) )
         kFillRange< T ><<< gridFor( lat.nc * lat.sc ), kBlock, 0, ctx->stream >>>( lat.nc * lat.sc, (T*) d, 0.0 );
      else if ( lat.haloPerCell() > 0 )
         kZeroHalos< T ><<< gridFor( lat.haloPerCell() * lat.nc ), kBlock, 0, ctx->stream >>>( BndGeom::of( st, lat ),
                                                                                            lat.nc, (T*) d );
      HG_CHECK_HIP( hipGetLastError() );
      const int64_t ntiles = tilesPerCell( mm ) * lat.nc;
      const int64_t nint   = tet( tileEdges( mm ) - 1 ) * lat.nc; // interior tiles
      const int64_t ndiag  = tri( tileEdges( mm ) ) * lat.nc;     // last-layer tiles
      const K*      table  = p2Coefficients< K >( ctx, op, L, src.dtype );
      const auto*   geom   = p2Geometry< T >( op, L );
      const uint64_t* tint  = mm > kSmallMaxMM ? tileList( ctx, lat, mm, kTilesInterior ) : nullptr;
      const uint64_t* tlast = mm > kSmallMaxMM ? tileList( ctx, lat, mm, kTilesLast ) : nullptr;
      ctx->beginTiming();
      if ( mm <= kSmallMaxMM )
      {
         if ( table )
            kP2SmallApply< T, K, true ><<< unsigned( lat.nc ), 64, 0, ctx->stream >>>(
                geom, op.tgeom.at( L ).p, op.kpar.p, table, mm, lat.nc, lat.sc, lat.w, (const T*) s, (T*) d, alpha );
         else
            kP2SmallApply< T, K, false ><<< unsigned( lat.nc ), 64, 0, ctx->stream >>>(
                geom, op.tgeom.at( L ).p, op.kpar.p, nullptr, mm, lat.nc, lat.sc, lat.w, (const T*) s, (T*) d, alpha );
      }
      else
      {
         if ( mm % kTile != 0 )
            throw Error( HG_ERR_STATE, "P2 tile apply: micro-edge count not a multiple of the tile width" );
         if ( nint > 0 && table )
            kP2TileApplyStored< T, K ><<< xcdGrid( nint ), kTileThreads, 0, ctx->stream >>>(
                geom, op.tgeom.at( L ).p, op.kpar.p, table, mm, nint, lat.sc, lat.w, (const T*) s, (T*) d, alpha,
                tint );
         else if ( nint > 0 )
            kP2TileApplyEval< T, K ><<< xcdGrid( nint ), kTileThreads, 0, ctx->stream >>>(
                geom, op.tgeom.at( L ).p, op.kpar.p, nullptr, mm, nint, lat.sc, lat.w, (const T*) s, (T*) d, alpha,
                tint );
         HG_CHECK_HIP( hipGetLastError() );
         if ( table )
            kP2TileApplyDiag< T, K, true ><<< xcdGrid( ndiag ), 64, 0, ctx->stream >>>(
                geom, op.tgeom.at( L ).p, op.kpar.p, table, mm, ndiag, lat.sc, lat.w, (const T*) s, (T*) d, alpha,
                tlast );
         else
            kP2TileApplyDiag< T, K, false ><<< xcdGrid( ndiag ), 64, 0, ctx->stream >>>(
                geom, op.tgeom.at( L ).p, op.kpar.p, nullptr, mm, ndiag, lat.sc, lat.w, (const T*) s, (T*) d, alpha,
                tlast );
      }
      HG_CHECK_HIP( hipGetLastError() );
      ctx->endTiming();
   } );
   pushAdditive( ctx, st, lat, d, dst.dtype, kAll ^ flag, update == HG_REPLACE );
}

// inverse diagonal of coefficient slot sl into out (one vector, lattice L+1)
static void p2SlotInverseDiagonal( Context* ctx, P2Operator& op, int L, int sl, double* out )
{
   Storage&      st  = *op.st;
   Lattice&      lat = st.lattice( ctx, L + 1 );
   const int64_t mm  = int64_t( 1 ) << L;
   const int64_t n   = microCellsPerCell( mm ) * lat.nc;
   kP2ElementDiag<<< gridFor( n, 128 ), 128, 0, ctx->stream >>>( op.geom.p, op.kpar.p + sl, mm, lat.nc, lat.sc, lat.w,
                                                                out );
   HG_CHECK_HIP( hipGetLastError() );
   pushAdditive( ctx, st, lat, out, 0, 0, true );
   kInvertNonZero<<< gridFor( lat.nstore ), kBlock, 0, ctx->stream >>>( lat.nstore, out );
   HG_CHECK_HIP( hipGetLastError() );
}

// dst[(i << lv) + k] = src[i], i < n
__global__ void kInsertVector( int64_t n, const double* __restrict__ src, int lv, int k, double* __restrict__ dst )
{
   const int64_t i = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( i < n )
      dst[( i << lv ) + k] = src[i];
}

// slot sl of a batched operator's inverse diagonal on level L (built before)
static void p2RefreshSlotDiagonal( Context* ctx, P2Operator& op, int L, int sl )
{
   Lattice& lat = op.st->lattice( ctx, L + 1 );
   auto&    tmp = op.invScratch[L];
   if ( !tmp.p )
      tmp.alloc( ctx, lat.nstore );
   kFillRange< double ><<< gridFor( lat.nstore ), kBlock, 0, ctx->stream >>>( lat.nstore, tmp.p, 0.0 );
   HG_CHECK_HIP( hipGetLastError() );
   p2SlotInverseDiagonal( ctx, op, L, sl, tmp.p );
   kInsertVector<<< gridFor( lat.nstore ), kBlock, 0, ctx->stream >>>( lat.nstore, tmp.p, op.lvSlots, sl,
                                                                      op.invDiag.at( L ).p );
   HG_CHECK_HIP( hipGetLastError() );
}

double* p2InverseDiagonal( Context* ctx, P2Operator& op, int L )
{
   auto it = op.invDiag.find( L );
   if ( it != op.invDiag.end() )
      return it->second.p;
   Storage& st  = *op.st;
   Lattice& lat = st.lattice( ctx, L + 1 );
   auto&    buf = op.invDiag[L];
   buf.alloc( ctx, size_t( lat.nstore ) << op.lvSlots );
   if ( op.nslot == 1 )
      p2SlotInverseDiagonal( ctx, op, L, 0, buf.p );
   else
      for ( int sl = 0; sl < op.nslot; ++sl )
         p2RefreshSlotDiagonal( ctx, op, L, sl );
   return buf.p;
}

// new coefficients for slot sl of a batched operator: its KParams, and its
// part of every coefficient table and inverse diagonal built so far
static void p2SetSlot( Context* ctx, P2Operator& op, int sl, const KParams& kp )
{
   {
      DeviceGateLock gate;
      HG_CHECK_HIP( hipMemcpyAsync( op.kpar.p + sl, &kp, sizeof( KParams ), hipMemcpyHostToDevice, ctx->stream ) );
      HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
   }
   for ( auto& [key, buf] : op.kq )
   {
      if ( key.second != 0 )
         continue; // batched operators build FP64 tables only
      const int     L      = key.first;
      const int64_t mm     = int64_t( 1 ) << L;
      const int64_t ntiles = tilesPerCell( mm ) * int64_t( op.st->topo.nc );
      kP2TileCoefficients< double ><<< gridFor( ntiles, 1 ), kTileThreads, 0, ctx->stream >>>(
          op.tgeom.at( L ).p, op.kpar.p + sl, mm, ntiles, (double*) buf.p + sl * op.tableStride.at( L ) );
      HG_CHECK_HIP( hipGetLastError() );
   }
   for ( auto& [L, buf] : op.invDiag )
      p2RefreshSlotDiagonal( ctx, op, L, sl );
}

template < class T >
__global__ void kMult( int64_t ncell, int64_t nb, const uint8_t* __restrict__ bndFlag, int flag,
                       const T* __restrict__ a, const T* __restrict__ b, int lv, T* __restrict__ dst )
{
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t i   = tid >> lv;
   const int64_t n   = ( flag & kInner ? ncell : 0 ) + nb;
   if ( i >= n )
      return;
   const int64_t pt = ( flag & kInner ) ? i : ncell + i;
   if ( pt >= ncell && !( bndFlag[pt - ncell] & flag ) )
      return;
   const int64_t p = ( pt << lv ) + ( tid & ( ( int64_t( 1 ) << lv ) - 1 ) );
   dst[p]          = fromAcc< T >( toAcc( a[p] ) * toAcc( b[p] ) );
}


// One step of ChebyshevSmoother::solve (ChebyshevSmoother.hpp:186-231) in one
// pass over the DoFs: r = b - s on the first step (s = A x) and r = s after it
// (s = A t), then t = invDiag * r on the flagged DoFs and 0 on all others, and
// x += c t on the flagged DoFs.  Each intermediate is rounded to the storage
// type where the separate assign / setToZero / multElementwise / add kernels
// store it, with the same operations (no contraction), so the result is theirs
// bit for bit.
// cv (optional): the step size of each of the 2^lv vectors, in place of c
template < class T >
__global__ void kChebyshevStep( int64_t ncell, int64_t nb, const uint8_t* __restrict__ bndFlag, int flag, int first,
                                double c, const T* __restrict__ b, const T* __restrict__ s,
                                const T* __restrict__ invDiag, int lv, T* __restrict__ t, T* __restrict__ x,
                                const double* __restrict__ cv )
{
#pragma clang fp contract( off )
   using A           = typename Acc< T >::type;
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t i   = tid >> lv;
   if ( i >= ncell + nb )
      return;
   if ( cv )
      c = cv[tid & ( ( int64_t( 1 ) << lv ) - 1 )];
   const int64_t p  = tid;
   const bool    on = i < ncell ? ( flag & kInner ) != 0 : ( bndFlag[i - ncell] & flag ) != 0;
   if ( !on )
   {
      t[p] = fromAcc< T >( A( 0 ) );
      return;
   }
   T r = s[p];
   if ( first )
   {
      A v = A( 0 );
      v += A( 1.0 ) * toAcc( b[p] );
      v += A( -1.0 ) * toAcc( r );
      r = fromAcc< T >( v );
   }
   const T tv = fromAcc< T >( toAcc( invDiag[p] ) * toAcc( r ) );
   t[p]       = tv;
   A v        = A( 0 );
   v += A( c ) * toAcc( tv );
   v += toAcc( x[p] );
   x[p] = fromAcc< T >( v );
}

} // namespace hg

using namespace hg;

extern "C" {

int hg_storage_shell( hg_ctx* h, int ntan, int nrad, double rmin, double rmax, hg_handle* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !out )
         throw Error( HG_ERR_ARG, "null output" );
      auto st  = std::make_unique< Storage >();
      st->topo = Topology::shell( ntan, nrad, rmin, rmax );
      st->init( ctx );
      *out                 = ctx->nextHandle++;
      ctx->storages[*out] = std::move( st );
   } );
}

int hg_storage_shell_bc( hg_ctx* h, int ntan, int nrad, double rmin, double rmax, int innerFlag, int outerFlag,
                         hg_handle* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      const int bnd = kDirichlet | kNeumann | kFreeslip;
      if ( !out || ( innerFlag & ~bnd ) || ( outerFlag & ~bnd ) || !innerFlag || !outerFlag ||
           ( innerFlag & ( innerFlag - 1 ) ) || ( outerFlag & ( outerFlag - 1 ) ) )
         throw Error( HG_ERR_ARG, "boundary flags must each be one of Dirichlet, Neumann, FreeSlip" );
      auto st  = std::make_unique< Storage >();
      st->topo = Topology::shell( ntan, nrad, rmin, rmax );
      st->topo.setShellBoundaryFlags( 0.5 * ( rmin + rmax ), innerFlag, outerFlag );
      st->init( ctx );
      *out                 = ctx->nextHandle++;
      ctx->storages[*out] = std::move( st );
   } );
}

int hg_storage_blend( hg_ctx* h, hg_handle s, double* out14 )
{
   return hg_guard( h, [&]( Context* ctx ) {
      const Topology& t = ctx->storage( s ).topo;
      if ( !out14 || t.blend.empty() )
         throw Error( HG_ERR_ARG, "storage has no blending" );
      for ( int c = 0; c < t.nc; ++c )
      {
         const ShellCellMap& b = t.blend[c];
         double*             o = out14 + 14 * size_t( c );
         for ( int d = 0; d < 3; ++d )
         {
            o[d]     = b.rayVertex[d];
            o[3 + d] = b.refVertex[d];
            o[6 + d] = b.thrVertex[d];
            o[9 + d] = b.forVertex[d];
         }
         o[12] = b.radRayVertex;
         o[13] = b.radRefVertex;
      }
   } );
}

int hg_storage_cell_vertices( hg_ctx* h, hg_handle s, int64_t* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      const Topology& t = ctx->storage( s ).topo;
      if ( !out )
         throw Error( HG_ERR_ARG, "null output" );
      for ( int c = 0; c < t.nc; ++c )
         for ( int k = 0; k < 4; ++k )
            out[4 * size_t( c ) + k] = t.cv[c][k];
   } );
}

int hg_storage_vertex_coords( hg_ctx* h, hg_handle s, double* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      const Topology& t = ctx->storage( s ).topo;
      if ( !out )
         throw Error( HG_ERR_ARG, "null output" );
      std::copy( t.X.begin(), t.X.end(), out );
   } );
}

int hg_p2_ndofs( hg_ctx* h, hg_handle s, int level, int64_t* out )
{
   return hg_p1_ndofs( h, s, level + 1, out );
}

int hg_p2_dof_coords( hg_ctx* h, hg_handle s, int level, double* xyz )
{
   return hg_p1_dof_coords( h, s, level + 1, xyz );
}

int hg_p2_dof_flags( hg_ctx* h, hg_handle s, int level, int32_t* flags )
{
   return hg_p1_dof_flags( h, s, level + 1, flags );
}

int hg_p2_divkgrad_create( hg_ctx* h, hg_handle s, int minLevel, int maxLevel, int nm, const double* a, const double* omega,
                           const double* phi, hg_handle* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !out || minLevel < 0 || maxLevel < minLevel || maxLevel > 11 || nm < 0 || nm > 32 ||
           ( nm > 0 && ( !a || !omega || !phi ) ) )
         throw Error( HG_ERR_ARG, "invalid operator parameters" );
      // the quadrature constants are uploaded once per process (several host
      // threads may create operators at once)
      static std::once_flag quadOnce;
      std::call_once( quadOnce, [] {
         double w[11];
         quadratureWeights( w );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cQP ), kQP, sizeof( kQP ) ) );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cQW ), w, sizeof( w ) ) );
         double q4[11][4], q4m[11][4];
         for ( int q = 0; q < 11; ++q )
            for ( int k = 0; k < 4; ++k )
            {
               q4[q][k]  = 4.0 * kQP[q][k];
               q4m[q][k] = 4.0 * kQP[q][k] - 1.0;
            }
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cQ4 ), q4, sizeof( q4 ) ) );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cQ4m ), q4m, sizeof( q4m ) ) );
         float wf[11], q4f[11][4], q4mf[11][4];
         for ( int q = 0; q < 11; ++q )
         {
            wf[q] = float( w[q] );
            for ( int k = 0; k < 4; ++k )
            {
               q4f[q][k]  = float( q4[q][k] );
               q4mf[q][k] = float( q4m[q][k] );
            }
         }
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cQWf ), wf, sizeof( wf ) ) );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cQ4f ), q4f, sizeof( q4f ) ) );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cQ4mf ), q4mf, sizeof( q4mf ) ) );
         double ql[11][3];
         float  qlf[11][3];
         for ( int q = 0; q < 11; ++q )
            for ( int k = 0; k < 3; ++k )
            {
               ql[q][k]  = kQP[q][k + 1];
               qlf[q][k] = float( kQP[q][k + 1] );
            }
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cQL ), ql, sizeof( ql ) ) );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cQLf ), qlf, sizeof( qlf ) ) );
         const std::vector< int32_t > pairs = diagonalPairs();
         if ( pairs.size() != 64 )
            throw Error( -1, "last-layer tile pairs: expected 64" );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cDiagPairs ), pairs.data(), sizeof( int32_t ) * 64 ) );
         const std::vector< int32_t > boxp = diagonalBoxPoints();
         if ( std::count_if( boxp.begin(), boxp.end(), []( int32_t v ) { return v >= 0; } ) != kDiagBox )
            throw Error( -1, "last-layer box points: expected 165" );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cDiagBox ), boxp.data(), sizeof( int32_t ) * kDiagBoxP ) );
      } );
      Storage&        st = ctx->storage( s );
      const Topology& t  = st.topo;
      auto            op = std::make_unique< P2Operator >();
      op->st             = &st;
      op->minL           = minLevel;
      op->maxL           = maxLevel;
      std::vector< CellGeom > g( t.nc );
      for ( int c = 0; c < t.nc; ++c )
      {
         const double* v[4];
         for ( int k = 0; k < 4; ++k )
            v[k] = &t.X[3 * t.cv[c][k]];
         for ( int d = 0; d < 3; ++d )
         {
            g[c].v0[d] = v[0][d];
            g[c].e1[d] = v[1][d] - v[0][d];
            g[c].e2[d] = v[2][d] - v[0][d];
            g[c].e3[d] = v[3][d] - v[0][d];
         }
         g[c].kind = t.blend.empty() ? 0 : 1;
         if ( g[c].kind )
         {
            const ShellCellMap& b = t.blend[c];
            for ( int d = 0; d < 3; ++d )
            {
               g[c].blend[d]     = b.rayVertex[d];
               g[c].blend[3 + d] = b.refVertex[d];
               g[c].blend[6 + d] = b.thrVertex[d];
               g[c].blend[9 + d] = b.forVertex[d];
            }
            g[c].blend[12] = b.radRayVertex;
            g[c].blend[13] = b.radRefVertex;
         }
      }
      op->geom.upload( ctx, g );
      const std::vector< P2CellGeom > tg = tileGeometry( t );
      for ( int L = minLevel; L <= maxLevel; ++L )
      {
         std::vector< P2CellGeom > gl = tg;
         tileGeometryLevel( gl, int64_t( 1 ) << L, kQP, 11 );
         op->tgeom[L].upload( ctx, gl );
         op->tgeomF[L].upload( ctx, toFloat( gl ) );
      }
      std::vector< KParams > kp( 1 );
      kp[0]    = KParams{};
      kp[0].nm = nm;
      for ( int m = 0; m < nm; ++m )
      {
         kp[0].a[m]  = a[m];
         kp[0].ph[m] = phi[m];
         for ( int d = 0; d < 3; ++d )
            kp[0].om[m][d] = omega[3 * m + d];
      }
      op->kpar.upload( ctx, kp );
      op->nm = nm;
      for ( int L = minLevel; L <= maxLevel; ++L )
         st.lattice( ctx, L + 1 );
      *out                  = ctx->nextHandle++;
      ctx->operators[*out] = std::move( op );
   } );
}

int hg_p2_divkgrad_create_batched( hg_ctx* h, hg_handle s, int minLevel, int maxLevel, int nslot, hg_handle* out )
{
   int lv = 0;
   while ( ( 1 << lv ) < nslot )
      ++lv;
   if ( nslot < 2 || ( 1 << lv ) != nslot || lv > kMaxLv || !out )
      return hg_guard( h, []( Context* ) {
         throw Error( HG_ERR_ARG, "batched P2 operator: the slot count must be a power of two from 2 to 64" );
      } );
   const int rc = hg_p2_divkgrad_create( h, s, minLevel, maxLevel, 0, nullptr, nullptr, nullptr, out );
   if ( rc != HG_OK )
      return rc;
   const int rc2 = hg_guard( h, [&]( Context* ctx ) {
      P2Operator& op = ctx->op< P2Operator >( *out );
      op.nslot       = nslot;
      op.lvSlots     = lv;
      op.reduced     = false;
      std::vector< KParams > kp( size_t( nslot ), KParams{} );
      op.kpar.upload( ctx, kp );
      op.slotOn.upload( ctx, std::vector< int32_t >( size_t( nslot ), 1 ) );
   } );
   if ( rc2 != HG_OK )
   {
      (void) hg_operator_destroy( h, *out );
      *out = 0;
   }
   return rc2;
}

int hg_p2_divkgrad_slot_active( hg_ctx* h, hg_handle opH, int slot, int active )
{
   return hg_guard( h, [&]( Context* ctx ) {
      P2Operator& op = ctx->op< P2Operator >( opH );
      if ( op.nslot < 2 || slot < 0 || slot >= op.nslot )
         throw Error( HG_ERR_ARG, "slot_active needs a batched operator and a slot index in range" );
      if ( tlsCaptures > 0 )
         throw Error( HG_ERR_STATE, "slot_active during a stream capture" );
      const int32_t  v = active ? 1 : 0;
      DeviceGateLock gate;
      HG_CHECK_HIP( hipMemcpyAsync( op.slotOn.p + slot, &v, sizeof( v ), hipMemcpyHostToDevice, ctx->stream ) );
      HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
   } );
}

int hg_p2_divkgrad_set_slot( hg_ctx* h, hg_handle opH, int slot, int nm, const double* a, const double* omega,
                             const double* phi )
{
   return hg_guard( h, [&]( Context* ctx ) {
      P2Operator& op = ctx->op< P2Operator >( opH );
      if ( op.nslot < 2 || slot < 0 || slot >= op.nslot || nm < 0 || nm > 32 || ( nm > 0 && ( !a || !omega || !phi ) ) )
         throw Error( HG_ERR_ARG, "set_slot needs a batched operator, a slot index in range and up to 32 modes" );
      KParams kp{};
      kp.nm = nm;
      for ( int m = 0; m < nm; ++m )
      {
         kp.a[m]  = a[m];
         kp.ph[m] = phi[m];
         for ( int d = 0; d < 3; ++d )
            kp.om[m][d] = omega[3 * m + d];
      }
      p2SetSlot( ctx, op, slot, kp );
   } );
}

int hg_p2_apply( hg_ctx* h, hg_handle op, hg_handle src, hg_handle dst, int level, int flag, int update )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( update != HG_REPLACE && update != HG_ADD )
         throw Error( HG_ERR_ARG, "invalid update type" );
      p2Apply( ctx, ctx->op< P2Operator >( op ), ctx->function( src ), ctx->function( dst ), level, flag, update );
   } );
}

int hg_p2_gemv( hg_ctx* h, hg_handle opH, double alpha, hg_handle src, double beta, hg_handle dst, int level, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      P2Operator& op = ctx->op< P2Operator >( opH );
      Function&   s  = ctx->function( src );
      Function&   d  = ctx->function( dst );
      if ( beta == 0.0 )
      {
         p2Apply( ctx, op, s, d, level, flag, HG_REPLACE, alpha );
         return;
      }
      p2CheckFunction( op, d, level );
      if ( beta != 1.0 )
         scaleFlagged( ctx, d, level, flag, beta );
      p2Apply( ctx, op, s, d, level, flag, HG_ADD, alpha );
   } );
}

int hg_p2_smooth_jac( hg_ctx* h, hg_handle opH, hg_handle dst, hg_handle rhs, hg_handle src, double relax, int level,
                      int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      P2Operator& op = ctx->op< P2Operator >( opH );
      Function&   d  = ctx->function( dst );
      Function&   r  = ctx->function( rhs );
      Function&   s  = ctx->function( src );
      p2CheckFunction( op, r, level );
      if ( r.dtype != d.dtype || &r == &d || &s == &d )
         throw Error( HG_ERR_ARG, "smooth_jac needs dst, rhs, src of one value type, dst distinct" );
      const double* inv = p2InverseDiagonal( ctx, op, level );
      p2Apply( ctx, op, s, d, level, flag, HG_REPLACE );
      jacobiUpdate( ctx, op.st->lattice( ctx, level + 1 ), d.dtype, 0, flag, relax, r.ptr( level ), s.ptr( level ), inv,
                    d.ptr( level ) );
   } );
}

int hg_p2_low_precision_form( hg_ctx* h, hg_handle op, int reduced, int fromLevel )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( fromLevel < 0 )
         throw Error( HG_ERR_ARG, "low_precision_form: negative level" );
      P2Operator& o = ctx->op< P2Operator >( op );
      o.reduced     = reduced != 0;
      o.reducedFrom = fromLevel;
   } );
}

int hg_p2_store_coefficients( hg_ctx* h, hg_handle op, int enable )
{
   return hg_guard( h, [&]( Context* ctx ) {
      P2Operator& o = ctx->op< P2Operator >( op );
      o.storeCoefficients = enable != 0;
      if ( !o.storeCoefficients )
         o.kq.clear();
   } );
}

int hg_p2_inverse_diagonal( hg_ctx* h, hg_handle op, int level, hg_handle dst )
{
   return hg_guard( h, [&]( Context* ctx ) {
      P2Operator& o = ctx->op< P2Operator >( op );
      Function&   f = ctx->function( dst );
      p2CheckFunction( o, f, level );
      const double* inv = p2InverseDiagonal( ctx, o, level );
      Lattice&      lat = o.st->lattice( ctx, level + 1 );
      const int64_t n   = lat.nstore << o.lvSlots;
      if ( f.dtype == 0 )
         HG_CHECK_HIP( hipMemcpyAsync( f.ptr( level ), inv, n * sizeof( double ), hipMemcpyDeviceToDevice, ctx->stream ) );
      else
         castCopy( ctx, inv, 0, f.ptr( level ), f.dtype, n );
   } );
}

int hg_function_mult( hg_ctx* h, hg_handle dst, hg_handle a, hg_handle b, int level, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& d  = ctx->function( dst );
      Function& fa = ctx->function( a );
      Function& fb = ctx->function( b );
      for ( Function* f : { &d, &fa, &fb } )
         if ( f->st != d.st || f->dtype != d.dtype || f->off != d.off || f->lv != d.lv || level < f->minL ||
              level > f->maxL )
            throw Error( HG_ERR_ARG, "multElementwise needs functions of one storage, space, value type and vector count" );
      Lattice&      l = d.lat( ctx, level );
      const int64_t n = ( flag & kInner ? l.nc * l.sc : 0 ) + l.nb;
      if ( n == 0 )
         return;
      dispatchType( d.dtype, [&]( auto tag ) {
         using T = decltype( tag );
         kMult< T ><<< gridFor( n << d.lv ), kBlock, 0, ctx->stream >>>( l.nc * l.sc, l.nb, l.bndFlag.p, flag,
                                                                         (const T*) fa.ptr( level ),
                                                                         (const T*) fb.ptr( level ), d.lv,
                                                                         (T*) d.ptr( level ) );
      } );
      HG_CHECK_HIP( hipGetLastError() );
   } );
}

static int chebyshevStep( hg_ctx* h, hg_handle x, hg_handle t, hg_handle s, hg_handle b, hg_handle invDiag, double c,
                          int slot0, int level, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& fx = ctx->function( x );
      Function& ft = ctx->function( t );
      Function& fs = ctx->function( s );
      Function& fd = ctx->function( invDiag );
      Function* fb = b ? &ctx->function( b ) : nullptr;
      for ( Function* f : { &fx, &ft, &fs, &fd, fb } )
         if ( f && ( f->st != fx.st || f->dtype != fx.dtype || f->off != fx.off || f->lv != fx.lv || level < f->minL ||
                     level > f->maxL ) )
            throw Error( HG_ERR_ARG, "chebyshev_step needs functions of one storage, space, value type and vector count" );
      if ( &fx == &ft || &ft == &fs || &fx == &fs )
         throw Error( HG_ERR_ARG, "chebyshev_step needs distinct x, t and s" );
      Lattice&      l = fx.lat( ctx, level );
      const int64_t n = l.nc * l.sc + l.nb;
      if ( n == 0 )
         return;
      const double* cv = nullptr;
      if ( slot0 >= 0 )
      {
         if ( slot0 + fx.nvec() > kSlots )
            throw Error( HG_ERR_ARG, "chebyshev_step_slots: slot range outside the context's slots" );
         cv = ctx->slots.p + slot0;
      }
      dispatchType( fx.dtype, [&]( auto tag ) {
         using T = decltype( tag );
         kChebyshevStep< T ><<< gridFor( n << fx.lv ), kBlock, 0, ctx->stream >>>(
             l.nc * l.sc, l.nb, l.bndFlag.p, flag, fb ? 1 : 0, c, fb ? (const T*) fb->ptr( level ) : nullptr,
             (const T*) fs.ptr( level ), (const T*) fd.ptr( level ), fx.lv, (T*) ft.ptr( level ), (T*) fx.ptr( level ),
             cv );
      } );
      HG_CHECK_HIP( hipGetLastError() );
   } );
}

int hg_function_chebyshev_step( hg_ctx* h, hg_handle x, hg_handle t, hg_handle s, hg_handle b, hg_handle invDiag,
                                double c, int level, int flag )
{
   return chebyshevStep( h, x, t, s, b, invDiag, c, -1, level, flag );
}

int hg_function_chebyshev_step_slots( hg_ctx* h, hg_handle x, hg_handle t, hg_handle s, hg_handle b, hg_handle invDiag,
                                      int slot0, int level, int flag )
{
   if ( slot0 < 0 )
      return hg_guard( h, []( Context* ) { throw Error( HG_ERR_ARG, "chebyshev_step_slots: invalid slot" ); } );
   return chebyshevStep( h, x, t, s, b, invDiag, 0.0, slot0, level, flag );
}

} // extern "C"
