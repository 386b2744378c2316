// Shared machinery of the tiled P2 element kernels (the P2 div(k grad)
// operator and the P2-P1 Stokes operator): micro-cell tables, per-cell tile
// geometry with the icosahedral-shell blending in rank-one form, the tile
// decomposition of a macro-cell, the LDS box positions of a micro-cell's P2
// DoFs and the shared-point accumulation.
#pragma once

#include <cmath>
#include <vector>

#include "halo.hpp"
#include "runtime.hpp"

namespace hg {

namespace {
__constant__ int    cVof[6][4][3] = {
    { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } }, { { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 }, { 1, 0, 1 } },
    { { 1, 0, 0 }, { 0, 1, 0 }, { 1, 0, 1 }, { 0, 0, 1 } }, { { 1, 1, 0 }, { 1, 1, 1 }, { 0, 1, 1 }, { 1, 0, 1 } },
    { { 1, 0, 1 }, { 0, 1, 1 }, { 0, 0, 1 }, { 0, 1, 0 } }, { { 0, 1, 0 }, { 1, 1, 0 }, { 1, 0, 1 }, { 0, 1, 1 } } };
__constant__ int cP2E[6][2] = { { 0, 1 }, { 0, 2 }, { 0, 3 }, { 1, 2 }, { 1, 3 }, { 2, 3 } };
} // namespace

// Per macro-cell data of the tiled element kernels.  Micro-cell of type t with
// anchor a has the computational-space vertices p_k = v0 + hE (a + vof_t[k]);
// J[t] = hE V_t (columns p_k - p_0), Dc[t] = rows of J[t]^-1 (gradients of the
// barycentric coordinates 1..3), adet[t] = |det J[t]|.  The shell map of the cell
// (IcosahedralShellMap.hpp:109-146,183-236) is F(x) = q(x) x with
// q = (r0 + |lam| dr) / |x|, lam = ((x - B0) . n) / volT, so that
// DF = q I + x w^T,  w = s n - (q / |x|^2) x,  s = sign(lam) dr / (volT |x|):
// a rank-one update of a multiple of the identity, inverted in closed form.
// R is the arithmetic type of the kernel that reads it (double, or float for
// the reduced-precision apply).
template < class R >
struct P2CellGeomT
{
   R   v0[3], hE[9];
   R   J[6][9], Dc[6][9], adet[6];
   R   XQ[6][11][4]; // J[t] (lam_1, lam_2, lam_3)_q and its lam-increment nl . XQ
   R   nl[3], lam0;  // lam(x) = nl . x - lam0,  nl = n / volT,  lam0 = nl . B0
   R   r0, dr;       // radius at the ray vertex and its increment to the ref vertex
   int kind;         // 0 affine, 1 shell
};
using P2CellGeom = P2CellGeomT< double >;

// micro-cell number j of type t in a macro-cell lattice of 2^L micro-edges
__device__ __forceinline__ void decodeMicro( int64_t wt, int64_t j, int64_t& x, int64_t& y, int64_t& z )
{
   // slice z holds tri(wt - z) micro-cells
   int64_t lo = 0, hi = wt - 1;
   const int64_t total = tet( wt );
   while ( lo < hi )
   {
      const int64_t mid = ( lo + hi + 1 ) / 2;
      if ( total - tet( wt - mid ) <= j )
         lo = mid;
      else
         hi = mid - 1;
   }
   z = lo;
   decodeTri( wt - z, j - ( total - tet( wt - z ) ), x, y );
}

// ---- tiled element kernels ------------------------------------------------
// A workgroup handles one tile of kTile^3 micro-cell anchors of one macro-cell
// (all 6 micro-cell types: 384 threads, one micro-cell each, one type per
// wavefront).  The tile's P2 values sit on a (2 kTile + 1)^3 box of the level
// L+1 lattice; the box is staged through LDS, element results are summed in
// LDS (one result box per wavefront), and the box is added back once with
// atomics that return nothing (box-surface points are shared with neighbouring
// tiles).
constexpr int kTile        = 4;
constexpr int kBox         = 2 * kTile + 1;
constexpr int kBox3        = kBox * kBox * kBox;
constexpr int kTileThreads = 6 * kTile * kTile * kTile;

__host__ __device__ inline int64_t tileEdges( int64_t mm ) { return ( mm - 1 ) / kTile + 1; }
__host__ __device__ inline int64_t tilesPerCell( int64_t mm ) { return tet( tileEdges( mm ) ); }

struct TileCoord
{
   int64_t c, x0, y0, z0;
};

__device__ __forceinline__ TileCoord tileOf( int64_t lb, int64_t mm )
{
   const int64_t T = tileEdges( mm );
   TileCoord     tc;
   const int64_t tpc = tilesPerCell( mm );
   tc.c              = lb / tpc;
   int64_t ti, tj, tk;
   decodeMicro( T, lb - tc.c * tpc, ti, tj, tk );
   tc.x0 = kTile * ti;
   tc.y0 = kTile * tj;
   tc.z0 = kTile * tk;
   return tc;
}

// Tiles whose anchor sum ti + tj + tk is at most T - 2 (T = tileEdges) lie
// inside the macro-cell up to a few micro-cells; they form the tile tetrahedron
// of edge T - 1.  The last layer (ti + tj + tk = T - 1) holds 64 micro-cells per
// tile (of 384 anchor/type pairs) and is handled by a one-wavefront kernel.
// tileOfInterior decodes interior tile lb and returns its number `full` in the
// numbering of all tiles (tileOf), which the coefficient tables use.
__device__ __forceinline__ TileCoord tileOfInterior( int64_t lb, int64_t mm, int64_t& full )
{
   const int64_t T = tileEdges( mm );
   TileCoord     tc;
   const int64_t tpi = tet( T - 1 );
   tc.c              = lb / tpi;
   int64_t ti, tj, tk;
   decodeMicro( T - 1, lb - tc.c * tpi, ti, tj, tk );
   tc.x0 = kTile * ti;
   tc.y0 = kTile * tj;
   tc.z0 = kTile * tk;
   full  = tc.c * tet( T ) + ( tet( T ) - tet( T - tk ) ) + triIndex( T - tk, ti, tj );
   return tc;
}
// last-layer tile lb (of tri(T) per cell) and its number in the numbering of all tiles
__device__ __forceinline__ TileCoord tileOfDiagonal( int64_t lb, int64_t mm, int64_t& full )
{
   const int64_t T = tileEdges( mm );
   TileCoord     tc;
   const int64_t per = tri( T );
   tc.c              = lb / per;
   int64_t ti, tj;
   decodeTri( T, lb - tc.c * per, ti, tj );
   const int64_t tk = T - 1 - ti - tj;
   tc.x0            = kTile * ti;
   tc.y0            = kTile * tj;
   tc.z0            = kTile * tk;
   full             = tc.c * tet( T ) + ( tet( T ) - tet( T - tk ) ) + triIndex( T - tk, ti, tj );
   return tc;
}
// ---- tile lists -------------------------------------------------------------
// Decoding a tile number by search costs a wavefront several hundred scalar
// instructions (64-bit divisions) before its first load can issue.  The apply
// kernels read their tile from a per-lattice list instead, one scalar load:
// entry = c | m << 24 | ti << 40 | tj << 48 | tk << 56 (cell, number of the tile
// in the cell in the numbering of all tiles, anchor / kTile).  List kTilesAll
// follows tileOf, kTilesInterior tileOfInterior, kTilesLast tileOfDiagonal.
// Lists exist while tet(T) < 2^16 and the cell count is below 2^24 (T: tile
// edge count); otherwise the kernels decode.
enum TileListKind
{
   kTilesAll      = 0,
   kTilesInterior = 1,
   kTilesLast     = 2
};
inline const uint64_t* tileList( Context* ctx, Lattice& l, int64_t mm, int kind )
{
   const int64_t T = tileEdges( mm );
   if ( tet( T ) >= ( int64_t( 1 ) << 16 ) || l.nc >= ( int64_t( 1 ) << 24 ) || T > 255 )
      return nullptr;
   DevBuf< uint64_t >& buf = l.tileLists[kind];
   if ( !buf.p )
   {
      std::vector< uint64_t > e;
      auto put = [&]( int64_t c, int64_t ti, int64_t tj, int64_t tk ) {
         const int64_t m = tet( T ) - tet( T - tk ) + triIndex( T - tk, ti, tj );
         e.push_back( uint64_t( c ) | ( uint64_t( m ) << 24 ) | ( uint64_t( ti ) << 40 ) | ( uint64_t( tj ) << 48 ) |
                      ( uint64_t( tk ) << 56 ) );
      };
      for ( int64_t c = 0; c < l.nc; ++c )
      {
         if ( kind == kTilesLast )
         {
            for ( int64_t tj = 0; tj < T; ++tj )
               for ( int64_t ti = 0; ti + tj < T; ++ti )
                  put( c, ti, tj, T - 1 - ti - tj );
            continue;
         }
         const int64_t E = kind == kTilesAll ? T : T - 1; // tiles with ti + tj + tk < E
         for ( int64_t tk = 0; tk < E; ++tk )
            for ( int64_t tj = 0; tj + tk < E; ++tj )
               for ( int64_t ti = 0; ti + tj + tk < E; ++ti )
                  put( c, ti, tj, tk );
      }
      if ( e.empty() )
         return nullptr;
      buf.upload( ctx, e );
   }
   return buf.p;
}
// tile i of a list (a uniform index: the entry comes with a scalar load through
// the constant address space); full = its number in the numbering of all tiles
__device__ __forceinline__ TileCoord tileFromList( const uint64_t* list, int64_t i, int64_t mm, int64_t& full )
{
   using ConstU64   = const __attribute__( ( address_space( 4 ) ) ) uint64_t;
   const uint64_t e = ( (ConstU64*) list )[i];
   TileCoord      tc;
   tc.c  = int64_t( e & 0xFFFFFFu );
   full  = tc.c * tilesPerCell( mm ) + int64_t( ( e >> 24 ) & 0xFFFFu );
   tc.x0 = kTile * int64_t( ( e >> 40 ) & 255u );
   tc.y0 = kTile * int64_t( ( e >> 48 ) & 255u );
   tc.z0 = kTile * int64_t( e >> 56 );
   return tc;
}

// the 64 (type, anchor) pairs of a last-layer tile inside the macro-cell, for
// lattices whose micro-edge count is a multiple of kTile: anchor sum at most
// 3 - rowShift(type); entry = type | anchor << 3
// The box points such a tile's micro-cells touch are those inside the
// macro-cell: lx + ly + lz <= 2 (kTile - 1) + 2 = 8, tet(9) = 165 of the 729
// (cDiagBox, box indices lx + 9 ly + 81 lz, padded to 192 with -1).
constexpr int kDiagBox  = 165;
constexpr int kDiagBoxP = 192;
namespace {
__constant__ int32_t cDiagPairs[64];
__constant__ int32_t cDiagBox[kDiagBoxP];
} // namespace
inline std::vector< int32_t > diagonalPairs()
{
   std::vector< int32_t > out;
   for ( int t = 0; t < 6; ++t )
      for ( int a = 0; a < 64; ++a )
         if ( ( a & 3 ) + ( ( a >> 2 ) & 3 ) + ( a >> 4 ) <= kTile - 1 - MicroCells::rowShift[t] )
            out.push_back( t | ( a << 3 ) );
   return out;
}
inline std::vector< int32_t > diagonalBoxPoints()
{
   std::vector< int32_t > out;
   for ( int lz = 0; lz < kBox; ++lz )
      for ( int ly = 0; ly < kBox; ++ly )
         for ( int lx = 0; lx < kBox; ++lx )
            if ( lx + ly + lz <= 2 * ( kTile - 1 ) + 2 )
               out.push_back( lx + kBox * ( ly + kBox * lz ) );
   out.resize( kDiagBoxP, -1 );
   return out;
}

// 1/b and 1/sqrt(b) from the hardware estimates (relative error e below 2^-22)
// refined by one third-order step each (error O(e^3), below the rounding of the
// result): r (1 + e + e^2) with e = 1 - b r, and y (1 + e/2 + 3e^2/8) with
// e = 1 - b y^2
__device__ __forceinline__ double rcpNewton( double b )
{
   const double r = __builtin_amdgcn_rcp( b );
   const double e = fma( -b, r, 1.0 );
   return fma( r, fma( e, e, e ), r );
}
__device__ __forceinline__ double rsqrtNewton( double b )
{
   const double y = __builtin_amdgcn_rsq( b );
   const double e = fma( -b, y * y, 1.0 );
   return fma( y, e * fma( e, 0.375, 0.5 ), y );
}
// FP64: refined estimates; FP32: the hardware estimates (about 1 ulp)
__device__ __forceinline__ double fastRcp( double b ) { return rcpNewton( b ); }
__device__ __forceinline__ double fastRsqrt( double b ) { return rsqrtNewton( b ); }
__device__ __forceinline__ float  fastRcp( float b ) { return __builtin_amdgcn_rcpf( b ); }
__device__ __forceinline__ float  fastRsqrt( float b ) { return __builtin_amdgcn_rsqf( b ); }

// shell-map scalars at x with lam(x) = ls: q = |F(x)| / |x|, w (DF = q I + x w^T),
// beta = q + w . x
template < class R >
__device__ __forceinline__ void shellScalars( const P2CellGeomT< R >& g, const R x[3], R ls, R& qq, R w[3], R& beta )
{
   const R r2   = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
   const R irho = fastRsqrt( r2 );
   qq           = ( g.r0 + fabs( ls ) * g.dr ) * irho;
   const R s    = ( ls < 0 ? -g.dr : g.dr ) * irho; // DF = q I - (q / rho^2) x x^T + s x nl^T
   const R qr2  = qq * irho * irho;
#pragma unroll
   for ( int d = 0; d < 3; ++d )
      w[d] = s * g.nl[d] - qr2 * x[d];
   beta = qq + w[0] * x[0] + w[1] * x[1] + w[2] * x[2];
}

template < class R >
struct TileThread
{
   int  t;          // micro-cell type (uniform per wavefront)
   int  ax, ay, az; // anchor inside the tile
   bool valid;
   R    p0[3];      // computational-space vertex 0
   R    ls0;        // lam(p0) of the shell map
};

// micro-cell type t at anchor index a (ax + 4 ay + 16 az) of tile tc
template < class R >
__device__ __forceinline__ TileThread< R > tileThreadAt( const P2CellGeomT< R >& g, const TileCoord& tc, int64_t mm,
                                                         int t, int a )
{
   TileThread< R > th;
   th.t             = t;
   th.ax            = a & 3;
   th.ay            = ( a >> 2 ) & 3;
   th.az            = a >> 4;
   const int64_t gx = tc.x0 + th.ax, gy = tc.y0 + th.ay, gz = tc.z0 + th.az;
   th.valid         = gx + gy + gz <= mm - 1 - MicroCells::rowShift[th.t];
   const R vx = R( gx + cVof[th.t][0][0] ), vy = R( gy + cVof[th.t][0][1] ), vz = R( gz + cVof[th.t][0][2] );
#pragma unroll
   for ( int d = 0; d < 3; ++d )
      th.p0[d] = g.v0[d] + g.hE[3 * d] * vx + g.hE[3 * d + 1] * vy + g.hE[3 * d + 2] * vz;
   th.ls0 = g.nl[0] * th.p0[0] + g.nl[1] * th.p0[1] + g.nl[2] * th.p0[2] - g.lam0;
   return th;
}

// the tile kernels: one type per wavefront, the anchor from the lane
template < class R >
__device__ __forceinline__ TileThread< R > tileThread( const P2CellGeomT< R >& g, const TileCoord& tc, int64_t mm )
{
   return tileThreadAt( g, tc, mm, __builtin_amdgcn_readfirstlane( int( threadIdx.x ) >> 6 ), int( threadIdx.x ) & 63 );
}

// LDS box positions of the 10 P2 DoFs of the thread's micro-cell (vertices in
// the order of getMicroVerticesFromMicroCell, then the edges (0,1), (0,2),
// (0,3), (1,2), (1,3), (2,3))
template < class R >
__device__ __forceinline__ void elementBoxPositions( const TileThread< R >& th, int pos[10] )
{
#pragma unroll
   for ( int k = 0; k < 4; ++k )
      pos[k] = 2 * ( th.ax + cVof[th.t][k][0] ) +
               kBox * ( 2 * ( th.ay + cVof[th.t][k][1] ) + kBox * 2 * ( th.az + cVof[th.t][k][2] ) );
#pragma unroll
   for ( int e = 0; e < 6; ++e )
      pos[4 + e] = ( pos[cP2E[e][0]] + pos[cP2E[e][1]] ) / 2;
}

// ---- per-wavefront accumulation of element results ---------------------------
// In the 384-thread tile kernels a wavefront holds the micro-cells of one type at
// 64 distinct anchors, so for one local DoF i its lanes address 64 distinct box
// points.  Wavefront w sums its lanes' results into box copy w (kWaveBoxes
// copies of kBox3 values, one per wavefront of the tile), one local DoF after
// the other, with plain LDS reads and writes: a wavefront's LDS operations
// complete in issue order, and the accesses are volatile so that the compiler
// keeps that order.  The write-back adds the copies (waveBoxSum).  LDS atomics
// on one shared box cost about twice the whole kernel time of this scheme.
constexpr int kWaveBoxes = kTileThreads / 64;

template < class R >
__device__ __forceinline__ void waveBoxAccumulate( R* boxes, bool valid, const int pos[10], const R y[10] )
{
   using LdsR     = volatile __attribute__( ( address_space( 3 ) ) ) R;
   const int lane = int( threadIdx.x ) & 63;
   LdsR*     b    = (LdsR*) ( boxes + kBox3 * __builtin_amdgcn_readfirstlane( int( threadIdx.x ) >> 6 ) );
   for ( int p = lane; p < kBox3; p += 64 )
      b[p] = R( 0 );
   if ( valid )
   {
#pragma unroll
      for ( int i = 0; i < 10; ++i )
         b[pos[i]] = b[pos[i]] + y[i];
   }
}
// box point p summed over the wavefront copies (after a workgroup barrier)
template < class R >
__device__ __forceinline__ R waveBoxSum( const R* boxes, int p )
{
   static_assert( kWaveBoxes == 6, "waveBoxSum adds six copies" );
   return ( ( boxes[p] + boxes[kBox3 + p] ) + ( boxes[2 * kBox3 + p] + boxes[3 * kBox3 + p] ) ) +
          ( boxes[4 * kBox3 + p] + boxes[5 * kBox3 + p] );
}

// dst += v at a point shared with other tiles
__device__ __forceinline__ void sharedAdd( double* o, double v ) { atomicAdd( o, v ); }
__device__ __forceinline__ void sharedAdd( float* o, float v ) { atomicAdd( o, v ); }
__device__ __forceinline__ void sharedAdd( __half* o, float v )
{
   // packed FP16 atomic add on the aligned pair holding the value, with +0 for
   // the other half (function buffers are padded, so the pair never reaches past
   // the allocation)
   typedef _Float16 h2 __attribute__( ( ext_vector_type( 2 ) ) );
   const uintptr_t a   = reinterpret_cast< uintptr_t >( o );
   h2*             w   = reinterpret_cast< h2* >( a & ~uintptr_t( 3 ) );
   const _Float16  hv  = _Float16( v );
   const h2        inc = ( a & 2 ) ? h2{ _Float16( 0.0f ), hv } : h2{ hv, _Float16( 0.0f ) };
   __builtin_amdgcn_global_atomic_fadd_v2f16( (__attribute__( ( address_space( 1 ) ) ) h2*) w, inc );
}

// halo positions of every cell lattice set to zero
template < class T >
__global__ void kZeroHalos( BndGeom g, int64_t nc, T* __restrict__ v )
{
   const int64_t per = 4 * g.fin + 6 * g.ein + 4;
   const int64_t gid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( gid >= per * nc )
      return;
   const int64_t c = gid / per;
   int64_t       j = gid - c * per;
   int64_t       t[4];
   if ( j < 4 * g.fin )
   {
      const int     k = int( j / g.fin );
      int64_t       a, b;
      decodeTri( g.mm - 2, j - k * g.fin, a, b );
      faceToCell( g.meta[c], k, g.mm, a + 1, b + 1, t );
   }
   else if ( ( j -= 4 * g.fin ) < 6 * g.ein )
   {
      const int k = int( j / g.ein );
      edgeToCell( g.meta[c], k, g.mm, j - k * g.ein + 1, t );
   }
   else
   {
      t[0] = t[1] = t[2] = t[3] = 0;
      t[j - 6 * g.ein]          = g.mm;
   }
   v[c * g.sc + tetIndex( g.w, t[1], t[2], t[3] )] = fromAcc< T >( typename Acc< T >::type( 0 ) );
}

template < class T >
__global__ void kFillRange( int64_t n, T* __restrict__ v, double value )
{
   const int64_t i = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( i < n )
      v[i] = fromAcc< T >( typename Acc< T >::type( value ) );
}

namespace {
__global__ void kInvertNonZero( int64_t n, double* __restrict__ v )
{
   const int64_t i = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( i < n && v[i] != 0.0 )
      v[i] = 1.0 / v[i];
}
} // namespace

static const int kVofHost[6][4][3] = {
    { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } }, { { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 }, { 1, 0, 1 } },
    { { 1, 0, 0 }, { 0, 1, 0 }, { 1, 0, 1 }, { 0, 0, 1 } }, { { 1, 1, 0 }, { 1, 1, 1 }, { 0, 1, 1 }, { 1, 0, 1 } },
    { { 1, 0, 1 }, { 0, 1, 1 }, { 0, 0, 1 }, { 0, 1, 0 } }, { { 0, 1, 0 }, { 1, 1, 0 }, { 1, 0, 1 }, { 0, 1, 1 } } };

// per macro-cell data of the tiled kernels for micro-edge length 1 (hE holds the
// macro-cell's edge vectors; tileGeometryLevel scales them to a level)
static std::vector< P2CellGeom > tileGeometry( const Topology& t )
{
   std::vector< P2CellGeom > out( t.nc );
   for ( int c = 0; c < t.nc; ++c )
   {
      P2CellGeom&   g = out[c];
      const double* v[4];
      for ( int k = 0; k < 4; ++k )
         v[k] = &t.X[3 * t.cv[c][k]];
      for ( int d = 0; d < 3; ++d )
      {
         g.v0[d] = v[0][d];
         for ( int k = 0; k < 3; ++k )
            g.hE[3 * d + k] = v[k + 1][d] - v[0][d];
      }
      g.kind = t.blend.empty() ? 0 : 1;
      if ( g.kind )
      {
         const ShellCellMap& b = t.blend[c];
         double              a[3], tt[3], f[3];
         for ( int d = 0; d < 3; ++d )
         {
            a[d]    = b.refVertex[d] - b.rayVertex[d];
            tt[d]   = b.thrVertex[d] - b.rayVertex[d];
            f[d]    = b.forVertex[d] - b.rayVertex[d];
         }
         const double n[3]  = { tt[1] * f[2] - tt[2] * f[1], tt[2] * f[0] - tt[0] * f[2], tt[0] * f[1] - tt[1] * f[0] };
         const double volT  = a[0] * n[0] + a[1] * n[1] + a[2] * n[2];
         for ( int d = 0; d < 3; ++d )
            g.nl[d] = n[d] / volT;
         g.lam0 = g.nl[0] * b.rayVertex[0] + g.nl[1] * b.rayVertex[1] + g.nl[2] * b.rayVertex[2];
         g.r0   = b.radRayVertex;
         g.dr   = b.radRefVertex - b.radRayVertex;
      }
      else
      {
         for ( int d = 0; d < 3; ++d )
            g.nl[d] = 0.0;
         g.lam0 = g.r0 = g.dr = 0.0;
      }
   }
   return out;
}

// scale the level-independent cell data to micro-edge length 1/mm and fill the
// per-type Jacobians, their inverses and determinants and the offsets XQ of the
// nq quadrature points qp (barycentric) from vertex 0
static void tileGeometryLevel( std::vector< P2CellGeom >& gs, int64_t mm, const double ( *qp )[4], int nq )
{
   const double h = 1.0 / double( mm );
   for ( auto& g : gs )
   {
      for ( double& e : g.hE )
         e *= h;
      for ( int ty = 0; ty < 6; ++ty )
      {
         double* J = g.J[ty];
         for ( int r = 0; r < 3; ++r )
            for ( int k = 0; k < 3; ++k )
            {
               double s = 0.0;
               for ( int m = 0; m < 3; ++m )
                  s += g.hE[3 * r + m] * double( kVofHost[ty][k + 1][m] - kVofHost[ty][0][m] );
               J[3 * r + k] = s;
            }
         const double c00 = J[4] * J[8] - J[5] * J[7], c01 = J[5] * J[6] - J[3] * J[8], c02 = J[3] * J[7] - J[4] * J[6];
         const double det = J[0] * c00 + J[1] * c01 + J[2] * c02;
         const double id  = 1.0 / det;
         double*      D   = g.Dc[ty];
         D[0]             = c00 * id;
         D[1]             = ( J[2] * J[7] - J[1] * J[8] ) * id;
         D[2]             = ( J[1] * J[5] - J[2] * J[4] ) * id;
         D[3]             = c01 * id;
         D[4]             = ( J[0] * J[8] - J[2] * J[6] ) * id;
         D[5]             = ( J[2] * J[3] - J[0] * J[5] ) * id;
         D[6]             = c02 * id;
         D[7]             = ( J[1] * J[6] - J[0] * J[7] ) * id;
         D[8]             = ( J[0] * J[4] - J[1] * J[3] ) * id;
         g.adet[ty]       = std::fabs( det );
         for ( int q = 0; q < nq; ++q )
         {
            double* xq = g.XQ[ty][q];
            for ( int d = 0; d < 3; ++d )
               xq[d] = J[3 * d] * qp[q][1] + J[3 * d + 1] * qp[q][2] + J[3 * d + 2] * qp[q][3];
            xq[3] = g.nl[0] * xq[0] + g.nl[1] * xq[1] + g.nl[2] * xq[2];
         }
      }
   }
}

static std::vector< P2CellGeomT< float > > toFloat( const std::vector< P2CellGeom >& gs )
{
   std::vector< P2CellGeomT< float > > out( gs.size() );
   for ( size_t c = 0; c < gs.size(); ++c )
   {
      const P2CellGeom&      g = gs[c];
      P2CellGeomT< float >& f = out[c];
      auto cp = []( float* dst, const double* src, int n ) {
         for ( int i = 0; i < n; ++i )
            dst[i] = float( src[i] );
      };
      cp( f.v0, g.v0, 3 );
      cp( f.hE, g.hE, 9 );
      cp( &f.J[0][0], &g.J[0][0], 54 );
      cp( &f.Dc[0][0], &g.Dc[0][0], 54 );
      cp( f.adet, g.adet, 6 );
      cp( &f.XQ[0][0][0], &g.XQ[0][0][0], 6 * 11 * 4 );
      cp( f.nl, g.nl, 3 );
      f.lam0 = float( g.lam0 );
      f.r0   = float( g.r0 );
      f.dr   = float( g.dr );
      f.kind = g.kind;
   }
   return out;
}

} // namespace hg
