// Quadratic P2 grid transfer: P2toP2QuadraticProlongation (prolongate /
// prolongateAndAdd, P2toP2QuadraticProlongation.cpp:37-63,217-425) and
// P2toP2QuadraticRestriction (restrict, P2toP2QuadraticRestriction.cpp:131-290).
//
// A P2 function of level L lives on the lattice of level L+1, so the fine P2
// nodes of level L+1 are the points F of lattice L+2 and the coarse nodes the
// points C of lattice L+1.  Prolongation is the interpolation of the coarse
// quadratic at every fine node: F lies in the coarse micro-cell found by a
// fixed rule (anchor floor(F/4), the containing micro-cell type of least row
// shift), whose 10 nodes and basis values depend only on F mod 4.  That gives
// 64 translation-invariant pull stencils.  Restriction is the transpose: for a
// coarse node C the fine nodes whose located micro-cell has C as a node, with
// the basis value as weight, form one stencil per parity class of C (vertex
// node or one of the 7 edge orientations).  Both tables are built once on the
// host by enumerating a lattice patch and live in constant memory.
//
// Macro-boundary DoFs are handled like the reference's additive kernels:
// a fine point shared by several macro-cells contributes 1/(number of cells
// sharing its primitive) from each, and restricted values of coarse DoFs on
// macro vertices/edges/faces are the sum over the incident cells.
#include <algorithm>
#include <mutex>
#include <array>
#include <vector>

#include "abi_util.hpp"
#include "halo.hpp"
#include "hyteg_amd.h"
#include "runtime.hpp"

namespace hg {

constexpr int kPMax = 10;  // nonzero weights of one fine node
constexpr int kRMax = 128; // fine nodes contributing to one coarse node

__constant__ int8_t cPOff[64][kPMax][3];
__constant__ double cPW[64][kPMax];
__constant__ int    cPN[64];
__constant__ int8_t cROff[8][kRMax][3];
__constant__ double cRW[8][kRMax];
__constant__ int    cRN[8];
// the restriction stencils grouped by fine row: group g of class q covers the
// fine row (2y + oy, 2z + oz) and entries [first, first + count) of cRGX / cRGW
// (x offset and weight)
constexpr int kRGroups = 49;
struct alignas( 8 ) RowGroup
{
   int16_t oy, oz, first, count; // two dwords: scalar loads
};
__constant__ RowGroup cRG[8][kRGroups];
__constant__ int      cRGN[8];
__constant__ int32_t  cRGX[8][kRMax];
__constant__ double   cRGW[8][kRMax];
// the prolongation offsets packed into one word each (ox | oy << 8 | oz << 16,
// all in 0..2): the kernels read them with scalar loads, which work on whole
// dwords
__constant__ uint32_t cPOffP[64][kPMax];

namespace {

const int kVof[6][4][3] = {
    { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } }, { { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 }, { 1, 0, 1 } },
    { { 1, 0, 0 }, { 0, 1, 0 }, { 1, 0, 1 }, { 0, 0, 1 } }, { { 1, 1, 0 }, { 1, 1, 1 }, { 0, 1, 1 }, { 1, 0, 1 } },
    { { 1, 0, 1 }, { 0, 1, 1 }, { 0, 0, 1 }, { 0, 1, 0 } }, { { 0, 1, 0 }, { 1, 1, 0 }, { 1, 0, 1 }, { 0, 1, 1 } } };
const int kShift[6]   = { 0, 1, 1, 2, 1, 1 };
const int kEdge[6][2] = { { 0, 1 }, { 0, 2 }, { 0, 3 }, { 1, 2 }, { 1, 3 }, { 2, 3 } };

struct Located
{
   int    node[10][3]; // lattice L+1 coordinates of the 10 nodes
   double phi[10];     // basis values at the fine point
};

// coarse micro-cell containing fine point F (lattice L+2 units) by the fixed
// rule; returns its nodes and the basis values at F
Located locate( const int F[3] )
{
   int A[3], r[3];
   for ( int d = 0; d < 3; ++d )
   {
      A[d] = F[d] >= 0 ? F[d] / 4 : -( ( -F[d] + 3 ) / 4 );
      r[d] = F[d] - 4 * A[d];
   }
   int order[6] = { 0, 1, 2, 4, 5, 3 }; // types by increasing row shift
   for ( int oi = 0; oi < 6; ++oi )
   {
      const int t = order[oi];
      // 4 lam' = M^-1 (r - 4 vof0), M columns vof_k - vof_0 (det +-1)
      int M[3][3];
      for ( int k = 0; k < 3; ++k )
         for ( int d = 0; d < 3; ++d )
            M[d][k] = kVof[t][k + 1][d] - kVof[t][0][d];
      const int det = M[0][0] * ( M[1][1] * M[2][2] - M[1][2] * M[2][1] ) -
                      M[0][1] * ( M[1][0] * M[2][2] - M[1][2] * M[2][0] ) +
                      M[0][2] * ( M[1][0] * M[2][1] - M[1][1] * M[2][0] );
      int adj[3][3];
      adj[0][0] = M[1][1] * M[2][2] - M[1][2] * M[2][1];
      adj[0][1] = M[0][2] * M[2][1] - M[0][1] * M[2][2];
      adj[0][2] = M[0][1] * M[1][2] - M[0][2] * M[1][1];
      adj[1][0] = M[1][2] * M[2][0] - M[1][0] * M[2][2];
      adj[1][1] = M[0][0] * M[2][2] - M[0][2] * M[2][0];
      adj[1][2] = M[0][2] * M[1][0] - M[0][0] * M[1][2];
      adj[2][0] = M[1][0] * M[2][1] - M[1][1] * M[2][0];
      adj[2][1] = M[0][1] * M[2][0] - M[0][0] * M[2][1];
      adj[2][2] = M[0][0] * M[1][1] - M[0][1] * M[1][0];
      int rel[3], l4[4];
      for ( int d = 0; d < 3; ++d )
         rel[d] = r[d] - 4 * kVof[t][0][d];
      int s = 0;
      for ( int k = 0; k < 3; ++k )
      {
         l4[k + 1] = ( adj[k][0] * rel[0] + adj[k][1] * rel[1] + adj[k][2] * rel[2] ) * det;
         s += l4[k + 1];
      }
      l4[0] = 4 - s;
      if ( l4[0] < 0 || l4[1] < 0 || l4[2] < 0 || l4[3] < 0 )
         continue;
      Located out;
      double  lam[4];
      for ( int k = 0; k < 4; ++k )
      {
         lam[k] = 0.25 * l4[k];
         for ( int d = 0; d < 3; ++d )
            out.node[k][d] = 2 * ( A[d] + kVof[t][k][d] );
         out.phi[k] = lam[k] * ( 2.0 * lam[k] - 1.0 );
      }
      for ( int e = 0; e < 6; ++e )
      {
         const int a = kEdge[e][0], b = kEdge[e][1];
         for ( int d = 0; d < 3; ++d )
            out.node[4 + e][d] = ( out.node[a][d] + out.node[b][d] ) / 2;
         out.phi[4 + e] = 4.0 * lam[a] * lam[b];
      }
      return out;
   }
   throw Error( -1, "P2 transfer tables: point not located" );
}

struct Tables
{
   int8_t   pOff[64][kPMax][3]{};
   double   pW[64][kPMax]{};
   int      pN[64]{};
   int8_t   rOff[8][kRMax][3]{};
   double   rW[8][kRMax]{};
   int      rN[8]{};
   RowGroup rg[8][kRGroups]{};
   int      rgN[8]{};
   int32_t  rgX[8][kRMax]{};
   double   rgW[8][kRMax]{};
   uint32_t pOffP[64][kPMax]{};
};

Tables buildTables()
{
   Tables T;
   (void) kShift;
   // pull stencils: fine point F = 4 A + r, offsets of the coarse nodes from 2 A
   for ( int rz = 0; rz < 4; ++rz )
      for ( int ry = 0; ry < 4; ++ry )
         for ( int rx = 0; rx < 4; ++rx )
         {
            const int     cls  = rx + 4 * ry + 16 * rz;
            const int     F[3] = { 64 + rx, 64 + ry, 64 + rz };
            const Located l    = locate( F );
            for ( int i = 0; i < 10; ++i )
               if ( l.phi[i] != 0.0 )
               {
                  const int n = T.pN[cls]++;
                  for ( int d = 0; d < 3; ++d )
                     T.pOff[cls][n][d] = int8_t( l.node[i][d] - 32 );
                  T.pW[cls][n] = l.phi[i];
                  if ( T.pOff[cls][n][0] < 0 || T.pOff[cls][n][1] < 0 || T.pOff[cls][n][2] < 0 ||
                       T.pOff[cls][n][0] > 2 || T.pOff[cls][n][1] > 2 || T.pOff[cls][n][2] > 2 )
                     throw Error( -1, "P2 prolongation: coarse node offset outside 0..2" );
                  T.pOffP[cls][n] = uint32_t( T.pOff[cls][n][0] ) | ( uint32_t( T.pOff[cls][n][1] ) << 8 ) |
                                    ( uint32_t( T.pOff[cls][n][2] ) << 16 );
               }
         }
   // transpose stencils: coarse node C0 = 32 + q (lattice L+1), fine points F near 2 C0
   for ( int q = 0; q < 8; ++q )
   {
      const int C0[3] = { 32 + ( q & 1 ), 32 + ( ( q >> 1 ) & 1 ), 32 + ( ( q >> 2 ) & 1 ) };
      for ( int dz = -8; dz <= 8; ++dz )
         for ( int dy = -8; dy <= 8; ++dy )
            for ( int dx = -8; dx <= 8; ++dx )
            {
               const int     F[3] = { 2 * C0[0] + dx, 2 * C0[1] + dy, 2 * C0[2] + dz };
               const Located l    = locate( F );
               for ( int i = 0; i < 10; ++i )
                  if ( l.phi[i] != 0.0 && l.node[i][0] == C0[0] && l.node[i][1] == C0[1] && l.node[i][2] == C0[2] )
                  {
                     const int n = T.rN[q]++;
                     if ( n >= kRMax )
                        throw Error( -1, "P2 restriction stencil too large" );
                     T.rOff[q][n][0] = int8_t( dx );
                     T.rOff[q][n][1] = int8_t( dy );
                     T.rOff[q][n][2] = int8_t( dz );
                     T.rW[q][n]      = l.phi[i];
                  }
            }
   }
   // the same entries grouped by fine row (first appearance order of the rows,
   // table order inside a row)
   for ( int q = 0; q < 8; ++q )
   {
      int n = 0;
      for ( int e = 0; e < T.rN[q]; ++e )
      {
         bool seen = false;
         for ( int g = 0; g < T.rgN[q]; ++g )
            seen = seen || ( T.rg[q][g].oy == T.rOff[q][e][1] && T.rg[q][g].oz == T.rOff[q][e][2] );
         if ( seen )
            continue;
         if ( T.rgN[q] >= kRGroups )
            throw Error( -1, "P2 restriction: too many fine rows" );
         RowGroup& g = T.rg[q][T.rgN[q]++];
         g.oy        = T.rOff[q][e][1];
         g.oz        = T.rOff[q][e][2];
         g.first     = int16_t( n );
         for ( int f = e; f < T.rN[q]; ++f )
            if ( T.rOff[q][f][1] == g.oy && T.rOff[q][f][2] == g.oz )
            {
               T.rgX[q][n] = T.rOff[q][f][0];
               T.rgW[q][n] = T.rW[q][f];
               ++n;
            }
         g.count = int16_t( n - g.first );
         if ( g.count > 7 )
            throw Error( -1, "P2 restriction: more than 7 entries on one fine row" );
      }
   }
   return T;
}

// built once per process; contexts on several host threads may ask at once
// (function-local statics are initialised exactly once, the others wait)
const Tables& tables()
{
   static const Tables T = buildTables();
   return T;
}

void uploadTables()
{
   static std::once_flag once;
   std::call_once( once, [] {
      const Tables& T = tables();
      HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cPOff ), T.pOff, sizeof( T.pOff ) ) );
      HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cPW ), T.pW, sizeof( T.pW ) ) );
      HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cPN ), T.pN, sizeof( T.pN ) ) );
      HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cROff ), T.rOff, sizeof( T.rOff ) ) );
      HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cRW ), T.rW, sizeof( T.rW ) ) );
      HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cRN ), T.rN, sizeof( T.rN ) ) );
      HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cRG ), T.rg, sizeof( T.rg ) ) );
      HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cRGN ), T.rgN, sizeof( T.rgN ) ) );
      HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cRGX ), T.rgX, sizeof( T.rgX ) ) );
      HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cRGW ), T.rgW, sizeof( T.rgW ) ) );
      HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cPOffP ), T.pOffP, sizeof( T.pOffP ) ) );
   } );
}

} // namespace

// value of the coarse quadratic at fine point (x,y,z) of cell c
template < class T >
__device__ __forceinline__ typename Acc< T >::type p2Interp( const T* __restrict__ coarse, int64_t c, int64_t scc,
                                                             int64_t wc, int64_t x, int64_t y, int64_t z, int lv )
{
   using A         = typename Acc< T >::type;
   const int cls   = int( x & 3 ) + 4 * int( y & 3 ) + 16 * int( z & 3 );
   const int64_t bx = 2 * ( x >> 2 ), by = 2 * ( y >> 2 ), bz = 2 * ( z >> 2 );
   A             v  = A( 0 );
   for ( int e = 0; e < cPN[cls]; ++e )
   {
      const int64_t cx = bx + cPOff[cls][e][0], cy = by + cPOff[cls][e][1], cz = bz + cPOff[cls][e][2];
      if ( cx < 0 || cy < 0 || cz < 0 || cx + cy + cz > wc - 1 )
         continue;
      v += A( cPW[cls][e] ) * toAcc( coarse[( c * scc + tetIndex( wc, cx, cy, cz ) ) << lv] );
   }
   return v;
}

template < class T >
__global__ void kP2ProlongCell( const int32_t* __restrict__ rows, int64_t nrows, int64_t nc, int64_t scf, int64_t wf,
                                int64_t mmf, int64_t scc, int64_t wc, const T* __restrict__ coarse,
                                T* __restrict__ fine, int add, int lv )
{
   using A          = typename Acc< T >::type;
   const int64_t lb = xcdBlock( nrows * nc );
   coarse += blockIdx.y;
   fine += blockIdx.y;
   if ( lb >= nrows * nc )
      return;
   const int64_t c   = lb / nrows;
   const int64_t r   = lb - c * nrows;
   const int64_t y   = rows[2 * r];
   const int64_t z   = rows[2 * r + 1];
   const int64_t out = c * scf + tetRowStart( wf, y, z );
   const int64_t by = 2 * ( y >> 2 ), bz = 2 * ( z >> 2 );
   // the stencils of one fine row reach the coarse rows (by + 0..2, bz + 0..2):
   // their starts once per block instead of a lattice index per entry
   __shared__ int64_t rs[9];
   if ( threadIdx.x < 9 )
   {
      const int64_t cy = by + threadIdx.x % 3, cz = bz + threadIdx.x / 3;
      rs[threadIdx.x]  = cy + cz <= wc - 1 ? c * scc + tetRowStart( wc, cy, cz ) : -1;
   }
   __syncthreads();
   for ( int64_t x = 1 + threadIdx.x; x + y + z < mmf; x += blockDim.x )
   {
      const int     cls = int( x & 3 ) + 4 * int( y & 3 ) + 16 * int( z & 3 );
      const int64_t bx  = 2 * ( x >> 2 );
      A             v   = A( 0 );
      for ( int e = 0; e < cPN[cls]; ++e )
      {
         const uint32_t o  = cPOffP[cls][e]; // offsets 0..2: one word per entry
         const int      oy = int( ( o >> 8 ) & 255u ), oz = int( o >> 16 );
         const int64_t  cx = bx + int64_t( o & 255u ), cy = by + oy, cz = bz + oz;
         if ( cx + cy + cz > wc - 1 )
            continue;
         v += A( cPW[cls][e] ) * toAcc( coarse[( rs[oy + 3 * oz] + cx ) << lv] );
      }
      if ( add )
         v += toAcc( fine[( out + x ) << lv] );
      fine[( out + x ) << lv] = fromAcc< T >( v );
   }
}

// Prolongation on cell-interior fine points, one workgroup per chunk of the
// lattice's chunk list (prolongChunks): the points of wavefront r all have
// stencil class r + 4 (y mod 4) + 16 (z mod 4), so offsets and weights are
// uniform over the wavefront, and the four wavefronts of a chunk fill the same
// fine rows.
constexpr int kProlongChunk = 256;
template < class T >
__global__ __launch_bounds__( kProlongChunk ) void kP2ProlongChunks( const uint32_t* __restrict__ chunks,
                                                                     int64_t nchunks, int64_t nc, int64_t scf,
                                                                     int32_t wf, int64_t scc, int32_t wc,
                                                                     const T* __restrict__ coarse,
                                                                     T* __restrict__ fine, int add, int lv )
{
   using A          = typename Acc< T >::type;
   const int64_t lb = xcdBlock( nchunks * nc );
   if ( lb >= nchunks * nc )
      return;
   const int64_t  c   = lb / nchunks;
   const int64_t  k   = lb - c * nchunks;
   const uint32_t pt = chunks[k * kProlongChunk + threadIdx.x];
   if ( pt == 0xFFFFFFFFu )
      return;
   // the stencil class of the wavefront's points, from its first active lane
   const uint32_t p0  = uint32_t( __builtin_amdgcn_readfirstlane( int( pt ) ) );
   const int      cls = int( p0 & 3u ) + 4 * int( ( p0 >> 10 ) & 3u ) + 16 * int( ( p0 >> 20 ) & 3u );
   const int32_t  x = int32_t( pt & 1023u ), y = int32_t( ( pt >> 10 ) & 1023u ), z = int32_t( pt >> 20 );
   const int32_t bx = 2 * ( x >> 2 ), by = 2 * ( y >> 2 ), bz = 2 * ( z >> 2 );
   const T*      cc = coarse + blockIdx.y + ( ( c * scc ) << lv );
   A             v  = A( 0 );
   // the (at most kPMax) coarse values gathered first, then summed in entry
   // order; coarse row starts from the lane's three slice starts (the entry
   // offsets are uniform over the wavefront)
   // tet(m) = m (m + 1) (m + 2) / 6 in 32-bit unsigned arithmetic (m <= 1024 here)
   auto          tet32 = []( int32_t m ) { return int32_t( uint32_t( m ) * uint32_t( m + 1 ) * uint32_t( m + 2 ) / 6u ); };
   const int     n     = cPN[cls];
   const int32_t tw    = tet32( wc );
   const int32_t s0 = tw - tet32( wc - bz ), s1 = tw - tet32( wc - bz - 1 ), s2 = tw - tet32( wc - bz - 2 );
   T             cv[kPMax];
   bool          ok[kPMax];
   uint32_t      offs[kPMax]; // entries past n are unused (zero offsets)
#pragma unroll
   for ( int e = 0; e < kPMax; ++e )
      offs[e] = cPOffP[cls][e];
#pragma unroll
   for ( int e = 0; e < kPMax; ++e )
   {
      const uint32_t o  = offs[e];
      const int32_t  oz = int32_t( o >> 16 );
      const int32_t  cx = bx + int32_t( o & 255u ), cy = by + int32_t( ( o >> 8 ) & 255u ), cz = bz + oz;
      const int32_t  row = ( oz == 0 ? s0 : ( oz == 1 ? s1 : s2 ) ) + cy * ( wc - cz ) - cy * ( cy - 1 ) / 2;
      ok[e]              = e < n && cx + cy + cz <= wc - 1;
      cv[e]              = ok[e] ? cc[int64_t( row + cx ) << lv] : T( 0 );
   }
#pragma unroll
   for ( int e = 0; e < kPMax; ++e )
      if ( ok[e] ) // false past n
         v += A( cPW[cls][e] ) * toAcc( cv[e] );
   T* out = fine + blockIdx.y + ( ( c * scf + tetRowStart32( wf, y, z ) + x ) << lv );
   if ( add )
      v += toAcc( *out );
   *out = fromAcc< T >( v );
}

// the chunk list of a fine lattice (built on first use; lattices up to 1023
// micro-edges, the packing's limit)
static void prolongChunkList( Context* ctx, Lattice& l )
{
   if ( l.prolongChunks.p || l.mm < 4 || l.mm > 1023 )
      return;
   std::vector< uint32_t > out;
   for ( int g = 0; g < 16; ++g )
   {
      const int64_t yr = g & 3, zr = g >> 2;
      std::vector< uint32_t > wave[4];
      auto flush = [&] {
         if ( wave[0].empty() && wave[1].empty() && wave[2].empty() && wave[3].empty() )
            return;
         for ( auto& w : wave )
         {
            w.resize( 64, 0xFFFFFFFFu );
            out.insert( out.end(), w.begin(), w.end() );
            w.clear();
         }
      };
      for ( int64_t z = 1; z <= l.mm - 3; ++z )
      {
         if ( ( z & 3 ) != zr )
            continue;
         for ( int64_t y = 1; y + z <= l.mm - 2; ++y )
         {
            if ( ( y & 3 ) != yr )
               continue;
            size_t add[4] = { 0, 0, 0, 0 };
            for ( int64_t x = 1; x + y + z <= l.mm - 1; ++x )
               ++add[x & 3];
            for ( int r = 0; r < 4; ++r )
               if ( wave[r].size() + add[r] > 64 )
               {
                  flush();
                  break;
               }
            for ( int64_t x = 1; x + y + z <= l.mm - 1; ++x )
            {
               if ( wave[x & 3].size() == 64 ) // a row longer than a chunk
                  flush();
               wave[x & 3].push_back( uint32_t( x ) | ( uint32_t( y ) << 10 ) | ( uint32_t( z ) << 20 ) );
            }
         }
      }
      flush();
   }
   if ( out.empty() )
      return;
   l.prolongChunks.upload( ctx, out );
}

template < class T >
__global__ void kP2ProlongBnd( BndGeom gf, int64_t nbf, int64_t nc, const uint8_t* __restrict__ bndFlag, int flag,
                               int64_t scc, int64_t wc, const T* __restrict__ coarse, T* __restrict__ fine, int add,
                               int lv )
{
   using A         = typename Acc< T >::type;
   const int64_t b = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   coarse += blockIdx.y;
   fine += blockIdx.y;
   if ( b >= nbf || !( bndFlag[b] & flag ) )
      return;
   bool done = false;
   A    v    = A( 0 );
   forEachIncident( gf, b, [&]( int64_t c, int64_t x, int64_t y, int64_t z, int ) {
      if ( done )
         return;
      done = true;
      v    = p2Interp( coarse, c, scc, wc, x, y, z, lv );
   } );
   const int64_t o = ( nc * gf.sc + b ) << lv;
   if ( add )
      v += toAcc( fine[o] );
   fine[o] = fromAcc< T >( v );
}

template < class T >
__global__ void kP2RestrictCell( const int32_t* __restrict__ rows, int64_t nrows, int64_t nc, int64_t scc, int64_t wc,
                                 int64_t mmc, int64_t scf, int64_t wf, const T* __restrict__ fine,
                                 T* __restrict__ coarse, int lv )
{
   using A          = typename Acc< T >::type;
   const int64_t lb = xcdBlock( nrows * nc );
   fine += blockIdx.y;
   coarse += blockIdx.y;
   if ( lb >= nrows * nc )
      return;
   const int64_t c   = lb / nrows;
   const int64_t r   = lb - c * nrows;
   const int64_t y   = rows[2 * r];
   const int64_t z   = rows[2 * r + 1];
   const int64_t out = c * scc + tetRowStart( wc, y, z );
   const int64_t mmf = wf - 1;
   // the fine rows (2y + -3..3, 2z + -3..3) the stencils of this coarse row reach
   __shared__ int64_t rs[49];
   for ( int i = threadIdx.x; i < 49; i += blockDim.x )
   {
      const int64_t fy = 2 * y + i % 7 - 3, fz = 2 * z + i / 7 - 3;
      rs[i]            = fy >= 0 && fz >= 0 && fy + fz <= mmf ? c * scf + tetRowStart( wf, fy, fz ) : -1;
   }
   __syncthreads();
   // odd x, then even x: the stencil class (and length: 125 entries for a
   // vertex DoF, 19 or 27 otherwise) is uniform over a wavefront
   for ( int par = 1; par >= 0; --par )
      for ( int64_t x = 2 - par + 2 * int64_t( threadIdx.x ); x + y + z < mmc; x += 2 * int64_t( blockDim.x ) )
      {
         const int q = int( x & 1 ) + 2 * int( y & 1 ) + 4 * int( z & 1 );
         A         v = A( 0 );
         // eight gathers in flight, then their sum in entry order
         for ( int e0 = 0; e0 < cRN[q]; e0 += 8 )
         {
            A    buf[8];
            bool ok[8];
#pragma unroll
            for ( int j = 0; j < 8; ++j )
            {
               const int     e  = e0 + j;
               const int     oy = cROff[q][e][1], oz = cROff[q][e][2];
               const int64_t fx = 2 * x + cROff[q][e][0], fy = 2 * y + oy, fz = 2 * z + oz;
               ok[j]  = e < cRN[q] && !( fx < 0 || fy < 0 || fz < 0 || fx + fy + fz > mmf );
               buf[j] = ok[j] ? toAcc( fine[( rs[( oy + 3 ) + 7 * ( oz + 3 )] + fx ) << lv] ) : A( 0 );
            }
#pragma unroll
            for ( int j = 0; j < 8; ++j )
               if ( ok[j] )
                  v += A( cRW[q][e0 + j] ) * buf[j];
         }
         coarse[( out + x ) << lv] = fromAcc< T >( v );
      }
}

// Restriction on cell-interior coarse points, one thread per point from the
// lattice's class point list: a wavefront holds points of one parity class, so
// the stencil (row groups, offsets, weights) is uniform over it and every lane
// has work.  Per fine row the row start comes from the slice starts in LDS; the
// entries of a row (at most 7) are gathered before they are summed.
struct ClassWaves
{
   int32_t w[9];
};
constexpr int kRestrictBlock = 256;
constexpr int kMaxSlices     = 1025;
template < class T >
__global__ __launch_bounds__( kRestrictBlock ) void kP2RestrictCellList( const uint32_t* __restrict__ pts, ClassWaves cw,
                                                                        int64_t nc, int64_t scc, int32_t wc, int64_t scf,
                                                                        int32_t wf, const T* __restrict__ fine,
                                                                        T* __restrict__ coarse, int lv )
{
   using A = typename Acc< T >::type;
   __shared__ int32_t sliceStart[kMaxSlices];
   const int32_t      wavesPerCell = cw.w[8];
   const int64_t      bpc          = ( wavesPerCell + kRestrictBlock / 64 - 1 ) / ( kRestrictBlock / 64 );
   const int64_t      lb           = xcdBlock( nc * bpc );
   if ( lb >= nc * bpc )
      return;
   for ( int32_t z = threadIdx.x; z < wf; z += kRestrictBlock )
      sliceStart[z] = int32_t( tet( wf ) - tet( wf - z ) );
   __syncthreads();
   const int64_t c    = lb / bpc;
   const int32_t wave = int32_t( lb - c * bpc ) * ( kRestrictBlock / 64 ) +
                        __builtin_amdgcn_readfirstlane( int32_t( threadIdx.x >> 6 ) );
   if ( wave >= wavesPerCell )
      return;
   const uint32_t p = pts[int64_t( wave ) * 64 + ( threadIdx.x & 63 )];
   if ( p == 0xFFFFFFFFu )
      return;
   const int32_t x = int32_t( p & 1023u ), y = int32_t( ( p >> 10 ) & 1023u ), z = int32_t( p >> 20 );
   // the wavefront's parity class (uniform over its points), from its first active lane
   const int q = __builtin_amdgcn_readfirstlane( int( x & 1 ) + 2 * int( y & 1 ) + 4 * int( z & 1 ) );
   const int32_t mmf = wf - 1;
   const T*      f   = fine + blockIdx.y + ( ( c * scf ) << lv );
   A             v   = A( 0 );
   for ( int g = 0; g < cRGN[q]; ++g )
   {
      const RowGroup rg = cRG[q][g];
      const int32_t  fy = 2 * y + rg.oy, fz = 2 * z + rg.oz;
      if ( fy < 0 || fz < 0 || fy + fz > mmf )
         continue;
      const int32_t rs  = sliceStart[fz] + fy * ( wf - fz ) - fy * ( fy - 1 ) / 2;
      const int32_t lim = mmf - fy - fz;
      A             buf[7];
#pragma unroll
      for ( int j = 0; j < 7; ++j )
      {
         buf[j] = A( 0 );
         if ( j < rg.count )
         {
            const int32_t fx = 2 * x + cRGX[q][rg.first + j];
            if ( fx >= 0 && fx <= lim )
               buf[j] = toAcc( f[int64_t( rs + fx ) << lv] );
         }
      }
#pragma unroll
      for ( int j = 0; j < 7; ++j )
         if ( j < rg.count )
            v += A( cRGW[q][rg.first + j] ) * buf[j];
   }
   coarse[blockIdx.y + ( ( c * scc + tetRowStart( wc, y, z ) + x ) << lv )] = fromAcc< T >( v );
}

// the class point list of a lattice (built on first use)
static void classPointList( Context* ctx, Lattice& l )
{
   // (the restriction kernel keeps the slice starts of the fine lattice, width
   // 2 mm + 1, in LDS; wider lattices keep the row kernel)
   if ( l.classPoints.p || l.mm < 4 || 2 * l.mm + 1 > kMaxSlices )
      return;
   // slice by slice (neighbouring waves gather from the same fine slices), in a
   // slice the points of each of its four parity classes padded to whole waves
   std::vector< uint32_t > pts;
   for ( int64_t z = 1; z <= l.mm - 3; ++z )
      for ( int q = 4 * int( z & 1 ); q < 4 * int( z & 1 ) + 4; ++q )
      {
         const size_t n0 = pts.size();
         for ( int64_t y = 1; y + z <= l.mm - 2; ++y )
         {
            if ( ( y & 1 ) != ( ( q >> 1 ) & 1 ) )
               continue;
            for ( int64_t x = 1; x + y + z <= l.mm - 1; ++x )
               if ( ( x & 1 ) == ( q & 1 ) )
                  pts.push_back( uint32_t( x ) | ( uint32_t( y ) << 10 ) | ( uint32_t( z ) << 20 ) );
         }
         while ( ( pts.size() - n0 ) % 64 )
            pts.push_back( 0xFFFFFFFFu );
      }
   for ( int q = 0; q < 8; ++q )
      l.classWave[q] = 0; // (unused: a wave's class follows from its points)
   l.classWave[8] = int32_t( pts.size() / 64 );
   if ( pts.empty() )
      pts.push_back( 0xFFFFFFFFu );
   l.classPoints.upload( ctx, pts );
}

// one wavefront per owned boundary DoF: the lanes split the stencil entries of
// each incident cell (a macro-vertex DoF of the shell sees up to ~36 cells with
// 125 entries each) and the wave sums its lanes in a fixed butterfly order
constexpr int kBndWaves = 4;
template < class T >
__global__ __launch_bounds__( 64 * kBndWaves ) void kP2RestrictBnd( BndGeom gc, int64_t nbc, int64_t nc,
                                                                   const uint8_t* __restrict__ bndFlag, int flag,
                                                                   const double* __restrict__ clsInvN, int64_t scf,
                                                                   int64_t wf, int64_t mmf, const T* __restrict__ fine,
                                                                   T* __restrict__ coarse, int lv )
{
   using A         = typename Acc< T >::type;
   fine += blockIdx.y;
   coarse += blockIdx.y;
   const int     l = int( threadIdx.x ) & 63;
   const int64_t b = int64_t( blockIdx.x ) * kBndWaves + ( int( threadIdx.x ) >> 6 );
   if ( b >= nbc || !( bndFlag[b] & flag ) )
      return; // uniform over the wave
   A s = A( 0 );
   forEachIncident( gc, b, [&]( int64_t c, int64_t x, int64_t y, int64_t z, int ) {
      const int q = int( x & 1 ) + 2 * int( y & 1 ) + 4 * int( z & 1 );
      for ( int e = l; e < cRN[q]; e += 64 )
      {
         const int64_t fx = 2 * x + cROff[q][e][0], fy = 2 * y + cROff[q][e][1], fz = 2 * z + cROff[q][e][2];
         if ( fx < 0 || fy < 0 || fz < 0 || fx + fy + fz > mmf )
            continue;
         const int cl = pointClass( mmf, fx, fy, fz );
         const A   w  = A( cRW[q][e] ) * ( cl < 0 ? A( 1 ) : A( clsInvN[14 * c + cl] ) );
         s += w * toAcc( fine[( c * scf + tetIndex( wf, fx, fy, fz ) ) << lv] );
      }
   } );
#pragma unroll
   for ( int o = 32; o > 0; o >>= 1 )
      s += __shfl_xor( s, o, 64 );
   if ( l == 0 )
      coarse[( nc * gc.sc + b ) << lv] = fromAcc< T >( s );
}

// owned face DoFs (two incident cells each): one thread per DoF
template < class T >
__global__ void kP2RestrictBndFace( BndGeom gc, int64_t nbc, int64_t nc, const uint8_t* __restrict__ bndFlag, int flag,
                                    const double* __restrict__ clsInvN, int64_t scf, int64_t wf, int64_t mmf,
                                    const T* __restrict__ fine, T* __restrict__ coarse, int lv )
{
   using A         = typename Acc< T >::type;
   fine += blockIdx.y;
   coarse += blockIdx.y;
   const int64_t b = gc.foff + int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( b >= nbc || !( bndFlag[b] & flag ) )
      return;
   A s = A( 0 );
   forEachIncident( gc, b, [&]( int64_t c, int64_t x, int64_t y, int64_t z, int ) {
      const int q = int( x & 1 ) + 2 * int( y & 1 ) + 4 * int( z & 1 );
      for ( int e = 0; e < cRN[q]; ++e )
      {
         const int64_t fx = 2 * x + cROff[q][e][0], fy = 2 * y + cROff[q][e][1], fz = 2 * z + cROff[q][e][2];
         if ( fx < 0 || fy < 0 || fz < 0 || fx + fy + fz > mmf )
            continue;
         const int cl = pointClass( mmf, fx, fy, fz );
         const A   w  = A( cRW[q][e] ) * ( cl < 0 ? A( 1 ) : A( clsInvN[14 * c + cl] ) );
         s += w * toAcc( fine[( c * scf + tetIndex( wf, fx, fy, fz ) ) << lv] );
      }
   } );
   coarse[( nc * gc.sc + b ) << lv] = fromAcc< T >( s );
}

static constexpr int kRowBlockT = 64;

void p2Restrict( Context* ctx, Function& f, int fineL, int flag )
{
   if ( f.off != 1 || fineL - 1 < f.minL || fineL > f.maxL )
      throw Error( HG_ERR_ARG, "P2 restriction needs a P2 function holding levels fineLevel-1 and fineLevel" );
   uploadTables();
   Storage& st = *f.st;
   Lattice& lf = st.lattice( ctx, fineL + 1 );
   Lattice& lc = st.lattice( ctx, fineL );
   classPointList( ctx, lc );
   void*    fp = f.ptr( fineL );
   void*    cp = f.ptr( fineL - 1 );
   const int      lv = f.lv;
   const unsigned nv = 1u << lv;
   pullHalos( ctx, st, lf, fp, f.dtype, lv );
   dispatchType( f.dtype, [&]( auto tag ) {
      using T = decltype( tag );
      if ( lc.nrows > 0 && lc.classPoints.p )
      {
         ClassWaves cw;
         for ( int q = 0; q < 9; ++q )
            cw.w[q] = lc.classWave[q];
         const int64_t bpc = ( cw.w[8] + kRestrictBlock / 64 - 1 ) / ( kRestrictBlock / 64 );
         ctx->beginTiming();
         kP2RestrictCellList< T ><<< dim3( xcdGrid( bpc * lc.nc ), nv ), kRestrictBlock, 0, ctx->stream >>>(
             lc.classPoints.p, cw, lc.nc, lc.sc, int32_t( lc.w ), lf.sc, int32_t( lf.w ), (const T*) fp, (T*) cp, lv );
         HG_CHECK_HIP( hipGetLastError() );
         ctx->endTiming();
      }
      else if ( lc.nrows > 0 )
      {
         ctx->beginTiming();
         kP2RestrictCell< T ><<< dim3( xcdGrid( lc.nrows * lc.nc ), nv ), kRowBlockT, 0, ctx->stream >>>(
             lc.rowsNat.p, lc.nrows, lc.nc, lc.sc, lc.w, lc.mm, lf.sc, lf.w, (const T*) fp, (T*) cp, lv );
         HG_CHECK_HIP( hipGetLastError() );
         ctx->endTiming();
      }
      if ( lc.nb > 0 )
      {
         // macro-vertex and -edge DoFs (many incident cells): a wavefront each;
         // face DoFs: a thread each
         const BndGeom gc = BndGeom::of( st, lc );
         const int64_t nve = std::min( gc.foff, lc.nb );
         if ( nve > 0 )
            kP2RestrictBnd< T ><<< dim3( gridFor( nve, kBndWaves ), nv ), 64 * kBndWaves, 0, ctx->stream >>>(
                gc, nve, lc.nc, lc.bndFlag.p, flag, st.clsInvN.p, lf.sc, lf.w, lf.mm, (const T*) fp, (T*) cp, lv );
         if ( lc.nb > nve )
            kP2RestrictBndFace< T ><<< dim3( gridFor( lc.nb - nve ), nv ), kBlock, 0, ctx->stream >>>(
                gc, lc.nb, lc.nc, lc.bndFlag.p, flag, st.clsInvN.p, lf.sc, lf.w, lf.mm, (const T*) fp, (T*) cp, lv );
         HG_CHECK_HIP( hipGetLastError() );
      }
   } );
}

void p2Prolongate( Context* ctx, Function& f, int coarseL, int flag, int update )
{
   if ( f.off != 1 || coarseL < f.minL || coarseL + 1 > f.maxL )
      throw Error( HG_ERR_ARG, "P2 prolongation needs a P2 function holding levels sourceLevel and sourceLevel+1" );
   uploadTables();
   Storage& st = *f.st;
   Lattice& lf = st.lattice( ctx, coarseL + 2 );
   Lattice& lc = st.lattice( ctx, coarseL + 1 );
   void*    fp = f.ptr( coarseL + 1 );
   void*    cp = f.ptr( coarseL );
   const int      lv = f.lv;
   const unsigned nv = 1u << lv;
   pullHalos( ctx, st, lc, cp, f.dtype, lv );
   dispatchType( f.dtype, [&]( auto tag ) {
      using T = decltype( tag );
      if ( lf.nrows > 0 )
      {
         prolongChunkList( ctx, lf );
         ctx->beginTiming();
         if ( lf.prolongChunks.p )
         {
            const int64_t nchunks = int64_t( lf.prolongChunks.n ) / kProlongChunk;
            kP2ProlongChunks< T ><<< dim3( xcdGrid( nchunks * lf.nc ), nv ), kProlongChunk, 0, ctx->stream >>>(
                lf.prolongChunks.p, nchunks, lf.nc, lf.sc, int32_t( lf.w ), lc.sc, int32_t( lc.w ), (const T*) cp,
                (T*) fp, update, lv );
         }
         else
            kP2ProlongCell< T ><<< dim3( xcdGrid( lf.nrows * lf.nc ), nv ), kRowBlockT, 0, ctx->stream >>>(
                lf.rowsNat.p, lf.nrows, lf.nc, lf.sc, lf.w, lf.mm, lc.sc, lc.w, (const T*) cp, (T*) fp, update, lv );
         HG_CHECK_HIP( hipGetLastError() );
         ctx->endTiming();
      }
      if ( lf.nb > 0 )
      {
         kP2ProlongBnd< T ><<< dim3( gridFor( lf.nb ), nv ), kBlock, 0, ctx->stream >>>(
             BndGeom::of( st, lf ), lf.nb, lf.nc, lf.bndFlag.p, flag, lc.sc, lc.w, (const T*) cp, (T*) fp, update, lv );
         HG_CHECK_HIP( hipGetLastError() );
      }
   } );
}

} // namespace hg

using namespace hg;

extern "C" {

int hg_p2_restrict( hg_ctx* h, hg_handle fn, int sourceLevel, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) { p2Restrict( ctx, ctx->function( fn ), sourceLevel, flag ); } );
}

int hg_p2_prolongate( hg_ctx* h, hg_handle fn, int sourceLevel, int flag, int update )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( update != HG_REPLACE && update != HG_ADD )
         throw Error( HG_ERR_ARG, "invalid update type" );
      p2Prolongate( ctx, ctx->function( fn ), sourceLevel, flag, update == HG_ADD );
   } );
}

int hg_p2_transfer_stencil( int kind, int cls, int max, int32_t* offsets, double* weights, int* count )
{
   try
   {
      const Tables& T = tables();
      if ( !count || ( kind != 0 && kind != 1 ) || cls < 0 || cls >= ( kind == 0 ? 64 : 8 ) )
         return HG_ERR_ARG;
      const int n = kind == 0 ? T.pN[cls] : T.rN[cls];
      *count      = n;
      for ( int e = 0; e < n && e < max; ++e )
      {
         for ( int d = 0; d < 3; ++d )
            offsets[3 * e + d] = kind == 0 ? T.pOff[cls][e][d] : T.rOff[cls][e][d];
         weights[e] = kind == 0 ? T.pW[cls][e] : T.rW[cls][e];
      }
      return HG_OK;
   }
   catch ( ... )
   {
      return HG_ERR_ARG;
   }
}

} // extern "C"
