// P2-P1 Taylor-Hood Stokes operator with the blended, variable-viscosity
// full-viscous velocity block (P2P1ElementwiseBlendingFullViscousStokesOperator,
// P2P1ElementwiseBlendingStokesOperator.hpp:125-216): host set-up, HIP kernels
// and C-ABI entry points.
//
// One tiled kernel evaluates any combination of the three blocks of
//   [ A  B^T ] [u]
//   [ B   0  ] [p]
// per micro-cell: A = 2 eta eps(u):eps(v) - 2/3 eta div u div v with the
// 5-point degree-3 rule (p2_full_stokesvar_*_blending_q3), B = -(q, div u) and
// B^T with the 4-point degree-2 rule (p2_to_p1_div_blending_q2,
// p1_to_p2_divt_blending_q2); eta = exp(c0 + c1 |F(x)|) at the blended points.
// The element matrices are never formed: per quadrature point the velocity
// gradient is built from the 30 element values, turned into the stress and
// tested against the 30 basis gradients, the same sums the reference's 9 + 6
// sub-operators form one block at a time.
//
// The shell map is used in rank-one form, DF = q I + x w^T, beta = q + w.x:
// DF^-T v = (v - w (x.v)/beta)/q, DF^-1 s = (s - x (w.s)/beta)/q and
// |det DF| = q^2 |beta|, so the powers of q cancel in the viscous block and one
// power remains in the divergence blocks.
#include <array>
#include <type_traits>
#include <cmath>
#include <mutex>

#include "abi_util.hpp"
#include "halo.hpp"
#include "hyteg_amd.h"
#include "p2tile.hpp"
#include "runtime.hpp"

namespace hg {

namespace {
// Stokes quadrature: points 0..4 are the degree-3 rule of the viscous block
// (centroid with weight -2/15, then the points with barycentric coordinate 1/2
// at vertex 3, 2, 1, 0 with weight 3/40), points 5..8 the degree-2 rule of the
// divergence blocks ((a,a,a,b) permutations, weight 1/24)
__constant__ double cSQ4[9][4];  // 4 lam_k
__constant__ double cSQ4m[9][4]; // 4 lam_k - 1
__constant__ double cSQL[9][4];  // lam_k
__constant__ double cSQW[9];

constexpr double kQa = 0.1381966011250105, kQb = 0.5854101966249685;
const double     kSQP[9][4] = { { 0.25, 0.25, 0.25, 0.25 },
                                { 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 0.5 },
                                { 1.0 / 6.0, 1.0 / 6.0, 0.5, 1.0 / 6.0 },
                                { 1.0 / 6.0, 0.5, 1.0 / 6.0, 1.0 / 6.0 },
                                { 0.5, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0 },
                                { kQa, kQa, kQa, kQb },
                                { kQa, kQa, kQb, kQa },
                                { kQa, kQb, kQa, kQa },
                                { kQb, kQa, kQa, kQa } };
const double     kSQW[9]    = { -2.0 / 15.0, 3.0 / 40.0, 3.0 / 40.0,  3.0 / 40.0, 3.0 / 40.0,
                                1.0 / 24.0,  1.0 / 24.0, 1.0 / 24.0,  1.0 / 24.0 };

constexpr int kPBox  = kTile + 1; // pressure box: the tile's level-L vertices
constexpr int kPBox3 = kPBox * kPBox * kPBox;

// FP32 copies of the quadrature tables for the reduced-precision velocity block
__constant__ float cSQ4f[9][4];
__constant__ float cSQ4mf[9][4];
__constant__ float cSQLf[9][4];
__constant__ float cSQWf[9];

template < class R >
struct SQ;
template <>
struct SQ< double >
{
   static __device__ __forceinline__ const double* a4( int q ) { return cSQ4[q]; }
   static __device__ __forceinline__ const double* b4( int q ) { return cSQ4m[q]; }
   static __device__ __forceinline__ const double* lam( int q ) { return cSQL[q]; }
   static __device__ __forceinline__ double        w( int q ) { return cSQW[q]; }
};
template <>
struct SQ< float >
{
   static __device__ __forceinline__ const float* a4( int q ) { return cSQ4f[q]; }
   static __device__ __forceinline__ const float* b4( int q ) { return cSQ4mf[q]; }
   static __device__ __forceinline__ const float* lam( int q ) { return cSQLf[q]; }
   static __device__ __forceinline__ float        w( int q ) { return cSQWf[q]; }
};

__device__ __forceinline__ double rexp( double v ) { return exp( v ); }
__device__ __forceinline__ float  rexp( float v ) { return expf( v ); }

// barycentric derivatives g_k = sum_j u_j dphi_j / dlam_k of a P2 function
template < class R >
__device__ __forceinline__ void baryGrad( const R* a4, const R* b4, const R u[10], R g[4] )
{
   g[0] = b4[0] * u[0] + a4[1] * u[4] + a4[2] * u[5] + a4[3] * u[6];
   g[1] = b4[1] * u[1] + a4[0] * u[4] + a4[2] * u[7] + a4[3] * u[8];
   g[2] = b4[2] * u[2] + a4[0] * u[5] + a4[1] * u[7] + a4[3] * u[9];
   g[3] = b4[3] * u[3] + a4[0] * u[6] + a4[1] * u[8] + a4[2] * u[9];
}

// computational-space gradient sum_k (g_k - g_0) grad lam_k
template < class R >
__device__ __forceinline__ void compGrad( const R* Dc, const R g[4], R v[3] )
{
   const R d1 = g[1] - g[0], d2 = g[2] - g[0], d3 = g[3] - g[0];
#pragma unroll
   for ( int d = 0; d < 3; ++d )
      v[d] = d1 * Dc[d] + d2 * Dc[3 + d] + d3 * Dc[6 + d];
}

// y_j += grad phi_j . tau for the 10 P2 functions (tau in computational space)
template < class R >
__device__ __forceinline__ void testGrad( const R* a4, const R* b4, const R* Dc, const R tau[3], R y[10] )
{
   const R e1 = Dc[0] * tau[0] + Dc[1] * tau[1] + Dc[2] * tau[2];
   const R e2 = Dc[3] * tau[0] + Dc[4] * tau[1] + Dc[5] * tau[2];
   const R e3 = Dc[6] * tau[0] + Dc[7] * tau[1] + Dc[8] * tau[2];
   const R e0 = -( e1 + e2 + e3 );
   y[0] += b4[0] * e0;
   y[1] += b4[1] * e1;
   y[2] += b4[2] * e2;
   y[3] += b4[3] * e3;
   y[4] += a4[1] * e0 + a4[0] * e1;
   y[5] += a4[2] * e0 + a4[0] * e2;
   y[6] += a4[3] * e0 + a4[0] * e3;
   y[7] += a4[2] * e1 + a4[1] * e2;
   y[8] += a4[3] * e1 + a4[1] * e3;
   y[9] += a4[3] * e2 + a4[2] * e3;
}

// quadrature point q of the thread's micro-cell: computational point x, the
// shell scalars (q = 1, w = 0, beta = 1 without blending), 1/beta and |F(x)|
template < class R >
struct QPoint
{
   R x[3], qq, w[3], beta, ib, r;
};

template < class R >
__device__ __forceinline__ QPoint< R > qpoint( const P2CellGeomT< R >& g, const TileThread< R >& th, int q )
{
   QPoint< R > P;
   const R*    xq = g.XQ[th.t][q];
#pragma unroll
   for ( int d = 0; d < 3; ++d )
      P.x[d] = th.p0[d] + xq[d];
   if ( g.kind )
   {
      const R ls = th.ls0 + xq[3];
      shellScalars( g, P.x, ls, P.qq, P.w, P.beta );
      P.ib = fastRcp( P.beta );
      P.r  = g.r0 + fabs( ls ) * g.dr;
   }
   else
   {
      P.qq = P.beta = P.ib = R( 1 );
      P.w[0] = P.w[1] = P.w[2] = R( 0 );
      P.r                      = sqrt( P.x[0] * P.x[0] + P.x[1] * P.x[1] + P.x[2] * P.x[2] );
   }
   return P;
}

// y (=|+=) the selected blocks on the cell lattices (halo positions included):
// MASK bit 1: yu = A u, bit 2: yu += B^T p, bit 4: yp = B u; MASK 8 alone:
// yp = M_k p, the P1 mass with k = 1/eta (the Schur-complement approximation
// of the block-factorisation preconditioner), 4-point degree-2 rule.
// T is the value type of the functions (double; float or half for the
// reduced-precision velocity block), the arithmetic runs in Acc<T>.
template < int MASK, class T >
__global__ __launch_bounds__( kTileThreads ) void kStokesTileApply(
    const P2CellGeomT< typename Acc< T >::type >* __restrict__ geom, typename Acc< T >::type c0,
    typename Acc< T >::type c1, int64_t mm, int64_t ntiles, int64_t sc1, int64_t w1, int64_t sc0, int64_t w0,
    const T* __restrict__ u0, const T* __restrict__ u1, const T* __restrict__ u2, const T* __restrict__ pin,
    T* __restrict__ y0, T* __restrict__ y1, T* __restrict__ y2, T* __restrict__ yp, const uint64_t* __restrict__ tiles )
{
   using R  = typename Acc< T >::type;
   using QC = SQ< R >;
   constexpr bool DO_A = MASK & 1, DO_BT = MASK & 2, DO_B = MASK & 4, DO_M = MASK & 8;
   constexpr bool NEED_U = DO_A || DO_B, OUT_U = DO_A || DO_BT, NEED_P = DO_BT || DO_M, OUT_P = DO_B || DO_M;
   // LDS boxes of the tile's input values and, apart from them, the result
   // boxes (cleared while the input loads are in flight: two workgroup barriers
   // per tile)
   __shared__ R su[3][kBox3];
   __shared__ R sp[kPBox3];
   __shared__ R au[OUT_U ? 3 : 1][OUT_U ? kBox3 : 1];
   __shared__ R ap[OUT_P ? kPBox3 : 1];
   const int64_t lb = xcdBlock( ntiles );
   if ( lb >= ntiles )
      return;
   int64_t         full;
   const TileCoord tc    = tiles ? tileFromList( tiles, lb, mm, full ) : tileOf( lb, mm );
   const int64_t   base1 = tc.c * sc1, base0 = tc.c * sc0;
   const int32_t   top1 = int32_t( 2 * mm ), top0 = int32_t( mm );
   const T*        uin[3] = { u0, u1, u2 };
   // the thread's (at most two) velocity box points and one pressure box point:
   // all loads issued before the first LDS write, so their latencies overlap
   static_assert( kBox3 <= 2 * kTileThreads && kPBox3 <= kTileThreads, "box points per thread" );
   T    uv[2][3], pvl = T( 0 );
   bool uok[2] = { false, false }, pok = false;
   if constexpr ( NEED_U )
   {
#pragma unroll
      for ( int j = 0; j < 2; ++j )
      {
         const int     p  = int( threadIdx.x ) + j * kTileThreads;
         const int32_t lx = p % kBox, ly = ( p / kBox ) % kBox, lz = p / ( kBox * kBox );
         const int32_t X = 2 * int32_t( tc.x0 ) + lx, Y = 2 * int32_t( tc.y0 ) + ly, Z = 2 * int32_t( tc.z0 ) + lz;
         uok[j]          = p < kBox3 && X + Y + Z <= top1;
         const int64_t idx = base1 + tetRowStart32( int32_t( w1 ), Y, Z ) + X;
#pragma unroll
         for ( int c = 0; c < 3; ++c )
            uv[j][c] = uok[j] ? uin[c][idx] : T( 0 );
      }
   }
   if constexpr ( NEED_P )
   {
      const int     p  = int( threadIdx.x );
      const int32_t lx = p % kPBox, ly = ( p / kPBox ) % kPBox, lz = p / ( kPBox * kPBox );
      const int32_t X = int32_t( tc.x0 ) + lx, Y = int32_t( tc.y0 ) + ly, Z = int32_t( tc.z0 ) + lz;
      pok             = p < kPBox3 && X + Y + Z <= top0;
      pvl             = pok ? pin[base0 + tetRowStart32( int32_t( w0 ), Y, Z ) + X] : T( 0 );
   }
   if constexpr ( OUT_U )
   {
#pragma unroll
      for ( int j = 0; j < 2; ++j )
      {
         const int p = int( threadIdx.x ) + j * kTileThreads;
         if ( p < kBox3 )
            au[0][p] = au[1][p] = au[2][p] = R( 0 );
      }
   }
   if constexpr ( OUT_P )
   {
      if ( int( threadIdx.x ) < kPBox3 )
         ap[threadIdx.x] = R( 0 );
   }
   if constexpr ( NEED_U )
   {
#pragma unroll
      for ( int j = 0; j < 2; ++j )
      {
         const int p = int( threadIdx.x ) + j * kTileThreads;
         if ( p < kBox3 )
         {
#pragma unroll
            for ( int c = 0; c < 3; ++c )
               su[c][p] = uok[j] ? R( toAcc( uv[j][c] ) ) : R( 0 );
         }
      }
   }
   if constexpr ( NEED_P )
   {
      if ( int( threadIdx.x ) < kPBox3 )
         sp[threadIdx.x] = pok ? R( toAcc( pvl ) ) : R( 0 );
   }
   __syncthreads();
   const P2CellGeomT< R >& g  = geom[tc.c];
   const TileThread< R >   th = tileThread( g, tc, mm );
   int                     pos[10], ppos[4];
   elementBoxPositions( th, pos );
#pragma unroll
   for ( int k = 0; k < 4; ++k )
      ppos[k] = ( th.ax + cVof[th.t][k][0] ) + kPBox * ( ( th.ay + cVof[th.t][k][1] ) + kPBox * ( th.az + cVof[th.t][k][2] ) );
   R u[3][10], pv[4];
   if constexpr ( NEED_U )
   {
#pragma unroll
      for ( int c = 0; c < 3; ++c )
#pragma unroll
         for ( int i = 0; i < 10; ++i )
            u[c][i] = th.valid ? su[c][pos[i]] : R( 0 );
   }
   if constexpr ( NEED_P )
   {
#pragma unroll
      for ( int k = 0; k < 4; ++k )
         pv[k] = th.valid ? sp[ppos[k]] : R( 0 );
   }
   if ( th.valid )
   {
      const R* Dc   = g.Dc[th.t];
      const R  adet = g.adet[th.t];
      R        y[3][10], yq[4];
#pragma unroll
      for ( int c = 0; c < 3; ++c )
#pragma unroll
         for ( int i = 0; i < 10; ++i )
            y[c][i] = R( 0 );
#pragma unroll
      for ( int k = 0; k < 4; ++k )
         yq[k] = R( 0 );
      if constexpr ( DO_A )
      {
#pragma unroll 1
         for ( int q = 0; q < 5; ++q )
         {
            const R*          a4 = QC::a4( q );
            const R*          b4 = QC::b4( q );
            const QPoint< R > P  = qpoint( g, th, q );
            // q DF^-T (computational gradient) of each component: rows of q grad u
            R G[3][3];
#pragma unroll
            for ( int c = 0; c < 3; ++c )
            {
               R gm[4];
               baryGrad( a4, b4, u[c], gm );
               compGrad( Dc, gm, G[c] );
               const R t = ( P.x[0] * G[c][0] + P.x[1] * G[c][1] + P.x[2] * G[c][2] ) * P.ib;
#pragma unroll
               for ( int d = 0; d < 3; ++d )
                  G[c][d] -= P.w[d] * t;
            }
            const R eta = rexp( c0 + c1 * P.r );
            const R tr  = ( R( 2 ) / R( 3 ) ) * eta * ( G[0][0] + G[1][1] + G[2][2] );
            const R cc  = QC::w( q ) * adet * fabs( P.beta );
#pragma unroll
            for ( int c = 0; c < 3; ++c )
            {
               // row c of the stress (times q), then q DF^-1 of it
               R s[3];
#pragma unroll
               for ( int d = 0; d < 3; ++d )
                  s[d] = eta * ( G[c][d] + G[d][c] ) - ( c == d ? tr : R( 0 ) );
               const R t = ( P.w[0] * s[0] + P.w[1] * s[1] + P.w[2] * s[2] ) * P.ib;
#pragma unroll
               for ( int d = 0; d < 3; ++d )
                  s[d] = cc * ( s[d] - P.x[d] * t );
               testGrad( a4, b4, Dc, s, y[c] );
            }
         }
      }
      if constexpr ( DO_M )
      {
#pragma unroll 1
         for ( int q = 5; q < 9; ++q )
         {
            const R*          lam = QC::lam( q );
            const QPoint< R > P   = qpoint( g, th, q );
            // |det D(F J)| = q^2 |beta| adet; k = 1/eta at the blended point
            const R cc = QC::w( q ) * adet * fabs( P.beta ) * P.qq * P.qq * rexp( -( c0 + c1 * P.r ) );
            const R s  = cc * ( lam[0] * pv[0] + lam[1] * pv[1] + lam[2] * pv[2] + lam[3] * pv[3] );
#pragma unroll
            for ( int k = 0; k < 4; ++k )
               yq[k] += lam[k] * s;
         }
      }
      if constexpr ( DO_B || DO_BT )
      {
#pragma unroll 1
         for ( int q = 5; q < 9; ++q )
         {
            const R*          a4  = QC::a4( q );
            const R*          b4  = QC::b4( q );
            const R*          lam = QC::lam( q );
            const QPoint< R > P   = qpoint( g, th, q );
            const R           cc  = QC::w( q ) * adet * fabs( P.beta ) * P.qq;
            if constexpr ( DO_B )
            {
               // q div u = trace of q grad u
               R dq = R( 0 );
#pragma unroll
               for ( int c = 0; c < 3; ++c )
               {
                  R gm[4], v[3];
                  baryGrad( a4, b4, u[c], gm );
                  compGrad( Dc, gm, v );
                  dq += v[c] - P.w[c] * ( P.x[0] * v[0] + P.x[1] * v[1] + P.x[2] * v[2] ) * P.ib;
               }
               const R s = -cc * dq;
#pragma unroll
               for ( int k = 0; k < 4; ++k )
                  yq[k] += lam[k] * s;
            }
            if constexpr ( DO_BT )
            {
               const R s = -cc * ( lam[0] * pv[0] + lam[1] * pv[1] + lam[2] * pv[2] + lam[3] * pv[3] );
#pragma unroll
               for ( int c = 0; c < 3; ++c )
               {
                  // q DF^-1 e_c, scaled
                  R tau[3];
#pragma unroll
                  for ( int d = 0; d < 3; ++d )
                     tau[d] = s * ( ( c == d ? R( 1 ) : R( 0 ) ) - P.x[d] * P.w[c] * P.ib );
                  testGrad( a4, b4, Dc, tau, y[c] );
               }
            }
         }
      }
      if constexpr ( OUT_U )
      {
#pragma unroll
         for ( int c = 0; c < 3; ++c )
#pragma unroll
            for ( int i = 0; i < 10; ++i )
               atomicAdd( &au[c][pos[i]], y[c][i] );
      }
      if constexpr ( OUT_P )
      {
#pragma unroll
         for ( int k = 0; k < 4; ++k )
            atomicAdd( &ap[ppos[k]], yq[k] );
      }
   }
   __syncthreads();
   if constexpr ( OUT_U )
   {
      T* yout[3] = { y0, y1, y2 };
      for ( int p = threadIdx.x; p < kBox3; p += kTileThreads )
      {
         const int32_t lx = p % kBox, ly = ( p / kBox ) % kBox, lz = p / ( kBox * kBox );
         const int32_t X = 2 * int32_t( tc.x0 ) + lx, Y = 2 * int32_t( tc.y0 ) + ly, Z = 2 * int32_t( tc.z0 ) + lz;
         if ( X + Y + Z > top1 )
            continue;
         // atomics for every point: a read-modify-write of the box interior would
         // wait for the outstanding atomics of the box surface
         const int64_t idx = base1 + tetRowStart32( int32_t( w1 ), Y, Z ) + X;
#pragma unroll
         for ( int c = 0; c < 3; ++c )
            sharedAdd( yout[c] + idx, au[c][p] );
      }
   }
   if constexpr ( OUT_P )
      for ( int p = threadIdx.x; p < kPBox3; p += kTileThreads )
      {
         const int32_t lx = p % kPBox, ly = ( p / kPBox ) % kPBox, lz = p / ( kPBox * kPBox );
         const int32_t X = int32_t( tc.x0 ) + lx, Y = int32_t( tc.y0 ) + ly, Z = int32_t( tc.z0 ) + lz;
         if ( X + Y + Z > top0 )
            continue;
         sharedAdd( yp + base0 + tetRowStart32( int32_t( w0 ), Y, Z ) + X, ap[p] );
      }
}

// diagonal of the three diagonal blocks of A:
//   sum_q w |det M| eta (|grad phi_j|^2 + (d_c phi_j)^2 / 3)
// added into the cell lattices (halo positions included)
__global__ __launch_bounds__( kTileThreads ) void kStokesTileDiag( const P2CellGeom* __restrict__ geom, double c0,
                                                                   double c1, int64_t mm, int64_t ntiles, int64_t sc1,
                                                                   int64_t w1, double* __restrict__ d0,
                                                                   double* __restrict__ d1, double* __restrict__ d2 )
{
   const int64_t lb = blockIdx.x;
   if ( lb >= ntiles )
      return;
   const TileCoord            tc = tileOf( lb, mm );
   const P2CellGeom&          g  = geom[tc.c];
   const TileThread< double > th = tileThread( g, tc, mm );
   if ( !th.valid )
      return;
   const double* Dc   = g.Dc[th.t];
   const double  adet = g.adet[th.t];
   double        dg[3][10] = {};
   for ( int q = 0; q < 5; ++q )
   {
      const double*          a4 = cSQ4[q];
      const double*          b4 = cSQ4m[q];
      const QPoint< double > P  = qpoint( g, th, q );
      const double           cc = cSQW[q] * adet * fabs( P.beta ) * exp( c0 + c1 * P.r );
      // computational gradients of lam_0..3
      double gl[4][3];
      for ( int d = 0; d < 3; ++d )
      {
         gl[1][d] = Dc[d];
         gl[2][d] = Dc[3 + d];
         gl[3][d] = Dc[6 + d];
         gl[0][d] = -( gl[1][d] + gl[2][d] + gl[3][d] );
      }
      for ( int j = 0; j < 10; ++j )
      {
         double v[3];
         if ( j < 4 )
            for ( int d = 0; d < 3; ++d )
               v[d] = b4[j] * gl[j][d];
         else
         {
            const int a = cP2E[j - 4][0], b = cP2E[j - 4][1];
            for ( int d = 0; d < 3; ++d )
               v[d] = a4[b] * gl[a][d] + a4[a] * gl[b][d];
         }
         const double t = ( P.x[0] * v[0] + P.x[1] * v[1] + P.x[2] * v[2] ) * P.ib;
         for ( int d = 0; d < 3; ++d )
            v[d] -= P.w[d] * t;
         const double n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
         for ( int c = 0; c < 3; ++c )
            dg[c][j] += cc * ( n2 + v[c] * v[c] / 3.0 );
      }
   }
   const int64_t gx = tc.x0 + th.ax, gy = tc.y0 + th.ay, gz = tc.z0 + th.az;
   int64_t       vx[4][3];
   for ( int k = 0; k < 4; ++k )
   {
      vx[k][0] = gx + cVof[th.t][k][0];
      vx[k][1] = gy + cVof[th.t][k][1];
      vx[k][2] = gz + cVof[th.t][k][2];
   }
   double* dout[3] = { d0, d1, d2 };
   for ( int j = 0; j < 10; ++j )
   {
      int64_t X, Y, Z;
      if ( j < 4 )
         X = 2 * vx[j][0], Y = 2 * vx[j][1], Z = 2 * vx[j][2];
      else
      {
         const int a = cP2E[j - 4][0], b = cP2E[j - 4][1];
         X = vx[a][0] + vx[b][0], Y = vx[a][1] + vx[b][1], Z = vx[a][2] + vx[b][2];
      }
      const int64_t idx = tc.c * sc1 + tetRowStart32( int32_t( w1 ), int32_t( Y ), int32_t( Z ) ) + X;
      for ( int c = 0; c < 3; ++c )
         atomicAdd( dout[c] + idx, dg[c][j] );
   }
}

// P1 diagonal operators on the affine micro-cells of level L (kind 0: the
// PSPG diagonal -(|T|^(2/3)/12) |T| |grad lam_k|^2, kind 1: the lumped mass
// |T|/4), added into the level-L cell lattices
__global__ __launch_bounds__( kTileThreads ) void kP1AffineDiag( const P2CellGeom* __restrict__ geom, int kind,
                                                                 int64_t mm, int64_t ntiles, int64_t sc0, int64_t w0,
                                                                 double* __restrict__ dst )
{
   const int64_t lb = blockIdx.x;
   if ( lb >= ntiles )
      return;
   const TileCoord            tc = tileOf( lb, mm );
   const P2CellGeom&          g  = geom[tc.c];
   const TileThread< double > th = tileThread( g, tc, mm );
   if ( !th.valid )
      return;
   const double* Dc  = g.Dc[th.t];
   const double  vol = g.adet[th.t] / 6.0;
   const double  s   = -cbrt( vol * vol ) / 12.0 * vol;
   for ( int k = 0; k < 4; ++k )
   {
      double gk[3];
      for ( int d = 0; d < 3; ++d )
         gk[d] = k == 0 ? -( Dc[d] + Dc[3 + d] + Dc[6 + d] ) : Dc[3 * ( k - 1 ) + d];
      const double  v   = kind == 0 ? s * ( gk[0] * gk[0] + gk[1] * gk[1] + gk[2] * gk[2] ) : 0.25 * vol;
      const int64_t X   = tc.x0 + th.ax + cVof[th.t][k][0];
      const int64_t Y   = tc.y0 + th.ay + cVof[th.t][k][1];
      const int64_t Z   = tc.z0 + th.az + cVof[th.t][k][2];
      const int64_t idx = tc.c * sc0 + tetRowStart32( int32_t( w0 ), int32_t( Y ), int32_t( Z ) ) + X;
      atomicAdd( dst + idx, v );
   }
}
} // namespace

// v += value on the flagged DoFs of a single-vector FP64 function
__global__ void kAddConst( int64_t ncell, int64_t nb, const uint8_t* __restrict__ bndFlag, int flag, double value,
                           double* __restrict__ v )
{
   const int64_t i = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t n = ( flag & kInner ? ncell : 0 ) + nb;
   if ( i >= n )
      return;
   const int64_t pt = ( flag & kInner ) ? i : ncell + i;
   if ( pt >= ncell && !( bndFlag[pt - ncell] & flag ) )
      return;
   v[pt] += value;
}

// v = 1/v on the flagged DoFs of a single-vector FP64 function (zeros stay)
__global__ void kInvertFlagged( int64_t ncell, int64_t nb, const uint8_t* __restrict__ bndFlag, int flag,
                                double* __restrict__ v )
{
   const int64_t i = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t n = ( flag & kInner ? ncell : 0 ) + nb;
   if ( i >= n )
      return;
   const int64_t pt = ( flag & kInner ) ? i : ncell + i;
   if ( pt >= ncell && !( bndFlag[pt - ncell] & flag ) )
      return;
   if ( v[pt] != 0.0 )
      v[pt] = 1.0 / v[pt];
}

// u -= (u . n) n on the owned boundary DoFs whose type is in flag
// (P2ProjectNormalOperator::project, p1functionspace/freeslip/VertexDoFProjectNormal.hpp:43-120)
template < class T >
__global__ void kProjectNormal( int64_t nb, int64_t ncell, const uint8_t* __restrict__ bndFlag, int flag,
                                const double* __restrict__ nrm, T* __restrict__ u0, T* __restrict__ u1,
                                T* __restrict__ u2 )
{
   using R         = typename Acc< T >::type;
   const int64_t b = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( b >= nb || !( bndFlag[b] & flag ) )
      return;
   const double* n    = nrm + 3 * b;
   const int64_t p    = ncell + b;
   const R       v[3] = { R( toAcc( u0[p] ) ), R( toAcc( u1[p] ) ), R( toAcc( u2[p] ) ) };
   const R       d    = R( n[0] ) * v[0] + R( n[1] ) * v[1] + R( n[2] ) * v[2];
   u0[p]              = fromAcc< T >( v[0] - d * R( n[0] ) );
   u1[p]              = fromAcc< T >( v[1] - d * R( n[1] ) );
   u2[p]              = fromAcc< T >( v[2] - d * R( n[2] ) );
}

struct StokesOperator : OperatorBase
{
   Storage*                                            st   = nullptr;
   int                                                 minL = 0, maxL = 0;
   double                                              c0 = 0.0, c1 = 0.0;
   std::map< int, DevBuf< P2CellGeom > >               tgeom;   // by P2 level
   std::map< int, DevBuf< P2CellGeomT< float > > >    tgeom32; // the same in FP32 (reduced-precision velocity block)
   std::map< int, std::array< DevBuf< double >, 3 > > invDiag; // velocity, lattice L+1
   std::map< int, DevBuf< double > >                   pspgInv, lumpedInv; // lattice L
};

namespace {

void checkVelocity( const StokesOperator& op, const Function& f, int L, int dtype = 0 )
{
   if ( f.st != op.st || f.off != 1 || f.lv != 0 || f.dtype != dtype )
      throw Error( HG_ERR_ARG, dtype == 0 ? "Stokes velocity components must be FP64 P2 functions on the operator's storage" :
                                            "the velocity components of a reduced-precision apply must share one type" );
   if ( L < op.minL || L > op.maxL || L < f.minL || L > f.maxL )
      throw Error( HG_ERR_ARG, "level outside the operator's or function's level range" );
}

void checkPressure( const StokesOperator& op, const Function& f, int L )
{
   if ( f.st != op.st || f.off != 0 || f.lv != 0 || f.dtype != 0 )
      throw Error( HG_ERR_ARG, "Stokes pressure must be an FP64 P1 function on the operator's storage" );
   if ( L < op.minL || L > op.maxL || L < f.minL || L > f.maxL )
      throw Error( HG_ERR_ARG, "level outside the operator's or function's level range" );
}

// pressure DoFs are all Inner (P2P1TaylorHoodFunction gives the pressure
// createAllInnerBC): a velocity flag selects every pressure DoF or none
int pressureFlag( int flag ) { return ( flag & kInner ) ? kAll : 0; }

template < class T >
void zeroTarget( Context* ctx, Storage& st, Lattice& lat, T* d, bool replaceAll )
{
   if ( replaceAll )
      kFillRange< T ><<< gridFor( lat.nc * lat.sc ), kBlock, 0, ctx->stream >>>( lat.nc * lat.sc, d, 0.0 );
   else if ( lat.haloPerCell() > 0 )
      kZeroHalos< T ><<< gridFor( lat.haloPerCell() * lat.nc ), kBlock, 0, ctx->stream >>>( BndGeom::of( st, lat ),
                                                                                         lat.nc, d );
   HG_CHECK_HIP( hipGetLastError() );
}

template < int MASK >
void launchStokes( Context* ctx, StokesOperator& op, int L, int64_t ntiles, Lattice& l1, Lattice& l0, double* const u[3],
                   const double* p, double* const y[3], double* yp )
{
   const uint64_t* tiles = tileList( ctx, l1, int64_t( 1 ) << L, kTilesAll );
   kStokesTileApply< MASK, double ><<< xcdGrid( ntiles ), kTileThreads, 0, ctx->stream >>>(
       op.tgeom.at( L ).p, op.c0, op.c1, int64_t( 1 ) << L, ntiles, l1.sc, l1.w, l0.sc, l0.w, u[0], u[1], u[2], p, y[0],
       y[1], y[2], yp, tiles );
   HG_CHECK_HIP( hipGetLastError() );
}

// the velocity block alone on FP32 or FP16 velocity functions (FP32 arithmetic)
void stokesViscousLowPrecision( Context* ctx, StokesOperator& op, Function* const src[4], Function* const dst[4], int L,
                                int flag, int update )
{
   const int dtype = src[0]->dtype;
   for ( int c = 0; c < 3; ++c )
   {
      checkVelocity( op, *src[c], L, dtype );
      checkVelocity( op, *dst[c], L, dtype );
      for ( int k = 0; k < 3; ++k )
         if ( src[c] == dst[k] )
            throw Error( HG_ERR_ARG, "apply needs distinct src and dst functions" );
   }
   Storage&      st = *op.st;
   Lattice&      l1 = st.lattice( ctx, L + 1 );
   Lattice&      l0 = st.lattice( ctx, L );
   const int64_t mm = int64_t( 1 ) << L;
   if ( l1.w > 1024 )
      throw Error( HG_ERR_UNSUPPORT, "Stokes apply supports levels up to 9 (32-bit lattice offsets)" );
   const int64_t ntiles = tilesPerCell( mm ) * l1.nc;
   dispatchType( dtype, [&]( auto tag ) {
      using T = decltype( tag );
      if constexpr ( std::is_same< T, double >::value )
         throw Error( HG_ERR_ARG, "FP64 velocity goes through the FP64 apply" );
      else
      {
         T* u[3];
         T* y[3];
         for ( int c = 0; c < 3; ++c )
         {
            u[c] = (T*) src[c]->ptr( L );
            pullHalos( ctx, st, l1, u[c], dtype );
            y[c] = (T*) dst[c]->ptr( L );
            zeroTarget( ctx, st, l1, y[c], update == HG_REPLACE && ( flag & kInner ) );
         }
         const uint64_t* tiles = tileList( ctx, l1, mm, kTilesAll );
         ctx->beginTiming();
         kStokesTileApply< 1, T ><<< xcdGrid( ntiles ), kTileThreads, 0, ctx->stream >>>(
             op.tgeom32.at( L ).p, float( op.c0 ), float( op.c1 ), mm, ntiles, l1.sc, l1.w, l0.sc, l0.w, u[0], u[1], u[2],
             nullptr, y[0], y[1], y[2], nullptr, tiles );
         HG_CHECK_HIP( hipGetLastError() );
         ctx->endTiming();
         for ( int c = 0; c < 3; ++c )
            pushAdditive( ctx, st, l1, y[c], dtype, kAll ^ flag, update == HG_REPLACE );
      }
   } );
}

void stokesApply( Context* ctx, StokesOperator& op, Function* const src[4], Function* const dst[4], int L, int flag,
                  int mask, int update )
{
   if ( mask < 1 || mask > 8 )
      throw Error( HG_ERR_ARG, "block mask must be 1..7 (1: A u, 2: B^T p, 4: B u) or 8 (the 1/eta pressure mass)" );
   const bool needU = mask & 5, needP = mask & 10, outU = mask & 3, outP = mask & 12;
   for ( int c = 0; c < 4; ++c )
      if ( ( ( c < 3 ? needU : needP ) && !src[c] ) || ( ( c < 3 ? outU : outP ) && !dst[c] ) )
         throw Error( HG_ERR_ARG, "a function the block mask needs is missing" );
   if ( mask == 1 && src[0]->dtype != 0 )
   {
      stokesViscousLowPrecision( ctx, op, src, dst, L, flag, update );
      return;
   }
   for ( int c = 0; c < 3; ++c )
   {
      if ( needU )
         checkVelocity( op, *src[c], L );
      if ( outU )
         checkVelocity( op, *dst[c], L );
   }
   if ( needP )
      checkPressure( op, *src[3], L );
   if ( outP )
      checkPressure( op, *dst[3], L );
   for ( int c = 0; c < 4; ++c )
      for ( int k = 0; k < 4; ++k )
         if ( ( ( c < 3 && needU ) || ( c == 3 && needP ) ) && ( ( k < 3 && outU ) || ( k == 3 && outP ) ) &&
              src[c] == dst[k] )
            throw Error( HG_ERR_ARG, "apply needs distinct src and dst functions" );
   Storage&      st = *op.st;
   Lattice&      l1 = st.lattice( ctx, L + 1 );
   Lattice&      l0 = st.lattice( ctx, L );
   const int64_t mm = int64_t( 1 ) << L;
   if ( l1.w > 1024 )
      throw Error( HG_ERR_UNSUPPORT, "Stokes apply supports levels up to 9 (32-bit lattice offsets)" );
   double* u[3] = { nullptr, nullptr, nullptr };
   double* y[3] = { nullptr, nullptr, nullptr };
   double* p    = nullptr;
   double* yp   = nullptr;
   for ( int c = 0; c < 3; ++c )
   {
      if ( needU )
      {
         u[c] = (double*) src[c]->ptr( L );
         pullHalos( ctx, st, l1, u[c], 0 );
      }
      if ( outU )
      {
         y[c] = (double*) dst[c]->ptr( L );
         zeroTarget( ctx, st, l1, y[c], update == HG_REPLACE && ( flag & kInner ) );
      }
   }
   const int pflag = pressureFlag( flag );
   if ( needP )
   {
      p = (double*) src[3]->ptr( L );
      pullHalos( ctx, st, l0, p, 0 );
   }
   if ( outP )
   {
      yp = (double*) dst[3]->ptr( L );
      zeroTarget( ctx, st, l0, yp, update == HG_REPLACE && ( pflag & kInner ) );
   }
   const int64_t ntiles = tilesPerCell( mm ) * l1.nc;
   tileList( ctx, l1, mm, kTilesAll ); // built before the timed region on first use
   ctx->beginTiming();
   switch ( mask )
   {
   case 1: launchStokes< 1 >( ctx, op, L, ntiles, l1, l0, u, p, y, yp ); break;
   case 2: launchStokes< 2 >( ctx, op, L, ntiles, l1, l0, u, p, y, yp ); break;
   case 3: launchStokes< 3 >( ctx, op, L, ntiles, l1, l0, u, p, y, yp ); break;
   case 4: launchStokes< 4 >( ctx, op, L, ntiles, l1, l0, u, p, y, yp ); break;
   case 5: launchStokes< 5 >( ctx, op, L, ntiles, l1, l0, u, p, y, yp ); break;
   case 6: launchStokes< 6 >( ctx, op, L, ntiles, l1, l0, u, p, y, yp ); break;
   case 7: launchStokes< 7 >( ctx, op, L, ntiles, l1, l0, u, p, y, yp ); break;
   default: launchStokes< 8 >( ctx, op, L, ntiles, l1, l0, u, p, y, yp ); break;
   }
   ctx->endTiming();
   if ( outU )
      for ( int c = 0; c < 3; ++c )
         pushAdditive( ctx, st, l1, y[c], 0, kAll ^ flag, update == HG_REPLACE );
   if ( outP )
      pushAdditive( ctx, st, l0, yp, 0, kAll ^ pflag, update == HG_REPLACE );
}

const std::array< DevBuf< double >, 3 >& velocityInverseDiagonal( Context* ctx, StokesOperator& op, int L )
{
   auto it = op.invDiag.find( L );
   if ( it != op.invDiag.end() )
      return it->second;
   Storage& st  = *op.st;
   Lattice& lat = st.lattice( ctx, L + 1 );
   auto&    b   = op.invDiag[L];
   for ( auto& d : b )
   {
      d.alloc( ctx, lat.nstore );
      kFillRange< double ><<< gridFor( lat.nstore ), kBlock, 0, ctx->stream >>>( lat.nstore, d.p, 0.0 );
   }
   const int64_t mm     = int64_t( 1 ) << L;
   const int64_t ntiles = tilesPerCell( mm ) * lat.nc;
   kStokesTileDiag<<< gridFor( ntiles, 1 ), kTileThreads, 0, ctx->stream >>>( op.tgeom.at( L ).p, op.c0, op.c1, mm, ntiles,
                                                                             lat.sc, lat.w, b[0].p, b[1].p, b[2].p );
   HG_CHECK_HIP( hipGetLastError() );
   for ( auto& d : b )
   {
      pushAdditive( ctx, st, lat, d.p, 0, 0, true );
      kInvertNonZero<<< gridFor( lat.nstore ), kBlock, 0, ctx->stream >>>( lat.nstore, d.p );
   }
   HG_CHECK_HIP( hipGetLastError() );
   return b;
}

const double* affineInverseDiagonal( Context* ctx, StokesOperator& op, int L, int kind )
{
   auto& m  = kind == 0 ? op.pspgInv : op.lumpedInv;
   auto  it = m.find( L );
   if ( it != m.end() )
      return it->second.p;
   Storage& st  = *op.st;
   Lattice& lat = st.lattice( ctx, L );
   auto&    b   = m[L];
   b.alloc( ctx, lat.nstore );
   kFillRange< double ><<< gridFor( lat.nstore ), kBlock, 0, ctx->stream >>>( lat.nstore, b.p, 0.0 );
   const int64_t mm     = int64_t( 1 ) << L;
   const int64_t ntiles = tilesPerCell( mm ) * lat.nc;
   kP1AffineDiag<<< gridFor( ntiles, 1 ), kTileThreads, 0, ctx->stream >>>( op.tgeom.at( L ).p, kind, mm, ntiles, lat.sc,
                                                                           lat.w, b.p );
   HG_CHECK_HIP( hipGetLastError() );
   pushAdditive( ctx, st, lat, b.p, 0, 0, true );
   kInvertNonZero<<< gridFor( lat.nstore ), kBlock, 0, ctx->stream >>>( lat.nstore, b.p );
   HG_CHECK_HIP( hipGetLastError() );
   return b.p;
}

// x/|x| of the owned boundary DoFs of a lattice: the outer normal of a sphere
// centred at the origin, which the shell map keeps (it only scales radii)
const double* boundaryNormals( Context* ctx, Storage& st, Lattice& lat )
{
   if ( lat.bndNormal.p )
      return lat.bndNormal.p;
   const Topology&       t  = st.topo;
   const int64_t         mm = lat.mm;
   std::vector< double > n( 3 * size_t( lat.nb ), 0.0 );
   std::vector< char >   done( size_t( lat.nb ), 0 );
   for ( int c = 0; c < t.nc; ++c )
   {
      const double* v[4];
      for ( int k = 0; k < 4; ++k )
         v[k] = &t.X[3 * t.cv[c][k]];
      for ( int64_t z = 0; z <= mm; ++z )
         for ( int64_t y = 0; y + z <= mm; ++y )
            for ( int64_t x = 0; x + y + z <= mm; x = ( y > 0 && z > 0 && x == 0 && mm - y - z > 0 ) ? mm - y - z : x + 1 )
            {
               const int64_t id = lat.num.locate( t, c, x, y, z );
               if ( id >= lat.nb || done[id] )
                  continue;
               done[id] = 1;
               double p[3], r = 0.0;
               for ( int d = 0; d < 3; ++d )
               {
                  p[d] = v[0][d] + ( double( x ) / double( mm ) ) * ( v[1][d] - v[0][d] ) +
                         ( double( y ) / double( mm ) ) * ( v[2][d] - v[0][d] ) +
                         ( double( z ) / double( mm ) ) * ( v[3][d] - v[0][d] );
                  r += p[d] * p[d];
               }
               r = std::sqrt( r );
               for ( int d = 0; d < 3; ++d )
                  n[3 * id + d] = p[d] / r;
            }
   }
   lat.bndNormal.upload( ctx, n );
   return lat.bndNormal.p;
}

void copyInto( Context* ctx, const double* src, Function& f, int L, size_t n )
{
   HG_CHECK_HIP( hipMemcpyAsync( f.ptr( L ), src, n * sizeof( double ), hipMemcpyDeviceToDevice, ctx->stream ) );
}

} // namespace
} // namespace hg

using namespace hg;

extern "C" {

int hg_stokes_create( hg_ctx* h, hg_handle s, int minLevel, int maxLevel, double viscC0, double viscC1, hg_handle* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !out || minLevel < 0 || maxLevel < minLevel || maxLevel > 9 || !std::isfinite( viscC0 ) ||
           !std::isfinite( viscC1 ) )
         throw Error( HG_ERR_ARG, "invalid Stokes operator parameters" );
      // uploaded once per process (several host threads may create operators at once)
      static std::once_flag quadOnce;
      std::call_once( quadOnce, [] {
         double q4[9][4], q4m[9][4];
         for ( int q = 0; q < 9; ++q )
            for ( int k = 0; k < 4; ++k )
            {
               q4[q][k]  = 4.0 * kSQP[q][k];
               q4m[q][k] = 4.0 * kSQP[q][k] - 1.0;
            }
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cSQ4 ), q4, sizeof( q4 ) ) );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cSQ4m ), q4m, sizeof( q4m ) ) );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cSQL ), kSQP, sizeof( kSQP ) ) );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cSQW ), kSQW, sizeof( kSQW ) ) );
         float f4[9][4], f4m[9][4], fl[9][4], fw[9];
         for ( int q = 0; q < 9; ++q )
         {
            for ( int k = 0; k < 4; ++k )
            {
               f4[q][k]  = float( q4[q][k] );
               f4m[q][k] = float( q4m[q][k] );
               fl[q][k]  = float( kSQP[q][k] );
            }
            fw[q] = float( kSQW[q] );
         }
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cSQ4f ), f4, sizeof( f4 ) ) );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cSQ4mf ), f4m, sizeof( f4m ) ) );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cSQLf ), fl, sizeof( fl ) ) );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cSQWf ), fw, sizeof( fw ) ) );
      } );
      Storage& st = ctx->storage( s );
      auto     op = std::make_unique< StokesOperator >();
      op->st      = &st;
      op->minL    = minLevel;
      op->maxL    = maxLevel;
      op->c0      = viscC0;
      op->c1      = viscC1;
      const std::vector< P2CellGeom > tg = tileGeometry( st.topo );
      for ( int L = minLevel; L <= maxLevel; ++L )
      {
         std::vector< P2CellGeom > gl = tg;
         tileGeometryLevel( gl, int64_t( 1 ) << L, kSQP, 9 );
         op->tgeom[L].upload( ctx, gl );
         op->tgeom32[L].upload( ctx, toFloat( gl ) );
         st.lattice( ctx, L );
         st.lattice( ctx, L + 1 );
      }
      *out                  = ctx->nextHandle++;
      ctx->operators[*out] = std::move( op );
   } );
}

int hg_stokes_apply( hg_ctx* h, hg_handle op, const hg_handle* src, const hg_handle* dst, int level, int flag, int mask,
                     int update )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( update != HG_REPLACE && update != HG_ADD )
         throw Error( HG_ERR_ARG, "invalid update type" );
      if ( !src || !dst )
         throw Error( HG_ERR_ARG, "null function list" );
      Function* s[4] = { nullptr, nullptr, nullptr, nullptr };
      Function* d[4] = { nullptr, nullptr, nullptr, nullptr };
      for ( int k = 0; k < 4; ++k )
      {
         const bool needS = k < 3 ? ( mask & 5 ) : ( mask & 10 );
         const bool needD = k < 3 ? ( mask & 3 ) : ( mask & 12 );
         if ( needS )
            s[k] = &ctx->function( src[k] );
         if ( needD )
            d[k] = &ctx->function( dst[k] );
      }
      stokesApply( ctx, ctx->op< StokesOperator >( op ), s, d, level, flag, mask, update );
   } );
}

int hg_stokes_velocity_inverse_diagonal( hg_ctx* h, hg_handle op, int level, const hg_handle* dst )
{
   return hg_guard( h, [&]( Context* ctx ) {
      StokesOperator& o = ctx->op< StokesOperator >( op );
      if ( !dst )
         throw Error( HG_ERR_ARG, "null function list" );
      const auto& inv = velocityInverseDiagonal( ctx, o, level );
      Lattice&    lat = o.st->lattice( ctx, level + 1 );
      for ( int c = 0; c < 3; ++c )
      {
         Function& f = ctx->function( dst[c] );
         checkVelocity( o, f, level );
         copyInto( ctx, inv[c].p, f, level, size_t( lat.nstore ) );
      }
   } );
}

int hg_stokes_pressure_diagonal( hg_ctx* h, hg_handle op, int level, int kind, hg_handle dst )
{
   return hg_guard( h, [&]( Context* ctx ) {
      StokesOperator& o = ctx->op< StokesOperator >( op );
      if ( kind != 0 && kind != 1 )
         throw Error( HG_ERR_ARG, "kind must be 0 (PSPG inverse diagonal) or 1 (inverse lumped mass)" );
      Function& f = ctx->function( dst );
      checkPressure( o, f, level );
      const double* inv = affineInverseDiagonal( ctx, o, level, kind );
      copyInto( ctx, inv, f, level, size_t( o.st->lattice( ctx, level ).nstore ) );
   } );
}

int hg_p2_project_normal( hg_ctx* h, const hg_handle* u, int level, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !u )
         throw Error( HG_ERR_ARG, "null function list" );
      // the normal is defined on boundary DoFs only (P2ProjectNormalOperator aborts on a non-boundary face)
      if ( flag & ( kInner | ~kAll ) )
         throw Error( HG_ERR_ARG, "normal projection flag must select boundary types only" );
      Function* f[3];
      for ( int c = 0; c < 3; ++c )
      {
         f[c] = &ctx->function( u[c] );
         if ( f[c]->off != 1 || f[c]->lv != 0 || f[c]->dtype != f[0]->dtype || f[c]->st != f[0]->st ||
              level < f[c]->minL || level > f[c]->maxL )
            throw Error( HG_ERR_ARG, "normal projection needs three P2 functions of one storage and one type" );
      }
      Storage& st  = *f[0]->st;
      Lattice& lat = st.lattice( ctx, level + 1 );
      if ( lat.nb == 0 )
         return;
      const double* nrm = boundaryNormals( ctx, st, lat );
      dispatchType( f[0]->dtype, [&]( auto tag ) {
         using T = decltype( tag );
         kProjectNormal< T ><<< gridFor( lat.nb ), kBlock, 0, ctx->stream >>>(
             lat.nb, lat.nc * lat.sc, lat.bndFlag.p, flag, nrm, (T*) f[0]->ptr( level ), (T*) f[1]->ptr( level ),
             (T*) f[2]->ptr( level ) );
      } );
      HG_CHECK_HIP( hipGetLastError() );
   } );
}

int hg_function_sum( hg_ctx* h, hg_handle fn, int level, int flag, double* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& f = ctx->function( fn );
      if ( !out || f.lv != 0 || level < f.minL || level > f.maxL )
         throw Error( HG_ERR_ARG, "sum needs a single-vector function and a level in its range" );
      Lattice& lat = f.lat( ctx, level );
      if ( !lat.ones.p )
      {
         lat.ones.alloc( ctx, lat.nstore );
         kFillRange< double ><<< gridFor( lat.nstore ), kBlock, 0, ctx->stream >>>( lat.nstore, lat.ones.p, 1.0 );
         HG_CHECK_HIP( hipGetLastError() );
      }
      if ( f.dtype != 0 )
         throw Error( HG_ERR_UNSUPPORT, "sum supports FP64 functions" );
      *out = dotDevice( ctx, lat, f.ptr( level ), lat.ones.p, 0, flag );
   } );
}

int hg_function_add_const( hg_ctx* h, hg_handle fn, double value, int level, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& f = ctx->function( fn );
      if ( f.lv != 0 || f.dtype != 0 || level < f.minL || level > f.maxL )
         throw Error( HG_ERR_ARG, "add of a constant needs a single-vector FP64 function and a level in its range" );
      Lattice&      l = f.lat( ctx, level );
      const int64_t n = ( flag & kInner ? l.nc * l.sc : 0 ) + l.nb;
      if ( n == 0 )
         return;
      kAddConst<<< gridFor( n ), kBlock, 0, ctx->stream >>>( l.nc * l.sc, l.nb, l.bndFlag.p, flag, value,
                                                             (double*) f.ptr( level ) );
      HG_CHECK_HIP( hipGetLastError() );
   } );
}

int hg_function_invert( hg_ctx* h, hg_handle fn, int level, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& f = ctx->function( fn );
      if ( f.lv != 0 || f.dtype != 0 || level < f.minL || level > f.maxL )
         throw Error( HG_ERR_ARG, "invertElementwise needs a single-vector FP64 function and a level in its range" );
      Lattice&      l = f.lat( ctx, level );
      const int64_t n = ( flag & kInner ? l.nc * l.sc : 0 ) + l.nb;
      if ( n == 0 )
         return;
      kInvertFlagged<<< gridFor( n ), kBlock, 0, ctx->stream >>>( l.nc * l.sc, l.nb, l.bndFlag.p, flag,
                                                                  (double*) f.ptr( level ) );
      HG_CHECK_HIP( hipGetLastError() );
   } );
}

} // extern "C"
