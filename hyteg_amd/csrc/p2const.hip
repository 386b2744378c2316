// P2 constant-stencil Laplace operator (P2ConstantLaplaceOperator,
// constant_stencil_operator/P2ConstantOperator.cpp:100-1350) with its
// Gauss-Seidel / SOR smoother: host set-up, HIP kernels and C-ABI entry points.
//
// On an affine macro-cell every level-L micro-tetrahedron of one of the six
// types has the same P2 element matrix, so the reference's four constant
// sub-stencils (vertex/edge to vertex/edge) are equivalent to the six 10x10
// element matrices per macro-cell and level kept here.  The apply runs element
// by element in the tile layout of the other P2 kernels (a 10x10 matrix-vector
// product per micro-cell instead of a quadrature); the smoother gathers the
// assembled row of a DoF from the element matrices of the micro-cells around it.
//
// Gauss-Seidel runs multicoloured: the owned macro vertex/edge/face DoFs first
// (greedy colouring of their coupling graph, assembled rows), then the cell
// interiors in 27 colours (lattice coordinates mod 3 -- the micro-cells around
// a DoF reach at most 2 lattice steps in each direction, so DoFs of one colour
// never share a row).  The reference sweeps primitive by primitive
// (vertices, edges, faces, cells) in lexicographic order; the multicoloured
// sweep has the same fixed point and comparable smoothing.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "abi_util.hpp"
#include "halo.hpp"
#include "hyteg_amd.h"
#include "p2tile.hpp"
#include "runtime.hpp"

namespace hg {

namespace {
// per parity class p = (X&1, Y&1, Z&1) of a level-(L+1) lattice point: the
// micro-cells of level L containing it, as (type, local DoF index, anchor
// offset from the point's level-L base (X>>1, Y>>1, Z>>1))
constexpr int kMaxInc = 24;
__constant__ int cIncN[8];
__constant__ int cInc[8][kMaxInc][5];
// lattice offsets of the 10 local DoFs of a type-t micro-cell from twice its anchor
__constant__ int cLoc[6][10][3];

struct IncTables
{
   int n[8];
   int e[8][kMaxInc][5];
   int loc[6][10][3];
};

IncTables buildIncTables()
{
   IncTables T{};
   const int edges[6][2] = { { 0, 1 }, { 0, 2 }, { 0, 3 }, { 1, 2 }, { 1, 3 }, { 2, 3 } };
   for ( int t = 0; t < 6; ++t )
   {
      for ( int k = 0; k < 4; ++k )
         for ( int d = 0; d < 3; ++d )
            T.loc[t][k][d] = 2 * kVofHost[t][k][d];
      for ( int e = 0; e < 6; ++e )
         for ( int d = 0; d < 3; ++d )
            T.loc[t][4 + e][d] = kVofHost[t][edges[e][0]][d] + kVofHost[t][edges[e][1]][d];
   }
   for ( int cls = 0; cls < 8; ++cls )
   {
      const int p[3] = { cls & 1, ( cls >> 1 ) & 1, ( cls >> 2 ) & 1 };
      T.n[cls]       = 0;
      for ( int t = 0; t < 6; ++t )
         for ( int i = 0; i < 10; ++i )
         {
            // pos_i(anchor a + dA) = 2 (a + dA) + loc_i must equal 2 a + p
            bool ok = true;
            int  dA[3];
            for ( int d = 0; d < 3; ++d )
            {
               const int r = p[d] - T.loc[t][i][d];
               if ( r % 2 != 0 )
                  ok = false;
               dA[d] = r / 2;
            }
            if ( !ok )
               continue;
            if ( T.n[cls] >= kMaxInc )
               throw Error( HG_ERR_STATE, "P2 incidence table overflow" );
            int* e = T.e[cls][T.n[cls]++];
            e[0]   = t;
            e[1]   = i;
            e[2]   = dA[0];
            e[3]   = dA[1];
            e[4]   = dA[2];
         }
   }
   return T;
}

const IncTables& incTables()
{
   static const IncTables T = buildIncTables();
   return T;
}

// exact P2 Laplace element matrix of the affine tetrahedron p[4]: the gradients
// of the P2 basis are linear, so the 4-point degree-2 rule integrates
// grad phi_i . grad phi_j exactly (P2Form_laplace; local order: 4 vertices, then
// the edges (0,1), (0,2), (0,3), (1,2), (1,3), (2,3))
void p2LaplaceElement( const double p[4][3], double A[10][10] )
{
   double J[9];
   for ( int r = 0; r < 3; ++r )
      for ( int k = 0; k < 3; ++k )
         J[3 * r + k] = p[k + 1][r] - p[0][r];
   const double c00 = J[4] * J[8] - J[5] * J[7], c01 = J[5] * J[6] - J[3] * J[8], c02 = J[3] * J[7] - J[4] * J[6];
   const double det = J[0] * c00 + J[1] * c01 + J[2] * c02;
   const double id  = 1.0 / det;
   // rows of J^-1 (the transposed cofactors over det) = gradients of lam_1..3
   double g[4][3];
   g[1][0] = c00 * id;
   g[1][1] = ( J[2] * J[7] - J[1] * J[8] ) * id;
   g[1][2] = ( J[1] * J[5] - J[2] * J[4] ) * id;
   g[2][0] = c01 * id;
   g[2][1] = ( J[0] * J[8] - J[2] * J[6] ) * id;
   g[2][2] = ( J[2] * J[3] - J[0] * J[5] ) * id;
   g[3][0] = c02 * id;
   g[3][1] = ( J[1] * J[6] - J[0] * J[7] ) * id;
   g[3][2] = ( J[0] * J[4] - J[1] * J[3] ) * id;
   for ( int d = 0; d < 3; ++d )
      g[0][d] = -( g[1][d] + g[2][d] + g[3][d] );
   const double a = 0.1381966011250105, b = 0.5854101966249685, w = std::fabs( det ) / 24.0;
   const int    edges[6][2] = { { 0, 1 }, { 0, 2 }, { 0, 3 }, { 1, 2 }, { 1, 3 }, { 2, 3 } };
   for ( int i = 0; i < 10; ++i )
      for ( int j = 0; j < 10; ++j )
         A[i][j] = 0.0;
   for ( int q = 0; q < 4; ++q )
   {
      double lam[4] = { a, a, a, a };
      lam[3 - q]    = b;
      double gp[10][3];
      for ( int k = 0; k < 4; ++k )
         for ( int d = 0; d < 3; ++d )
            gp[k][d] = ( 4.0 * lam[k] - 1.0 ) * g[k][d];
      for ( int e = 0; e < 6; ++e )
      {
         const int k = edges[e][0], l = edges[e][1];
         for ( int d = 0; d < 3; ++d )
            gp[4 + e][d] = 4.0 * ( lam[k] * g[l][d] + lam[l] * g[k][d] );
      }
      for ( int i = 0; i < 10; ++i )
         for ( int j = 0; j < 10; ++j )
            A[i][j] += w * ( gp[i][0] * gp[j][0] + gp[i][1] * gp[j][1] + gp[i][2] * gp[j][2] );
   }
}

// y = A x on the cell lattices of level L+1 (halo positions included), one
// 10x10 element product per micro-cell, tiled like the other P2 kernels
__global__ __launch_bounds__( kTileThreads ) void kP2ConstTileApply( const double* __restrict__ Ae, int64_t mm,
                                                                      int64_t ntiles, int64_t sc1, int64_t w1,
                                                                      const double* __restrict__ src,
                                                                      double* __restrict__ dst,
                                                                      const uint64_t* __restrict__ tiles )
{
   __shared__ double sb[kBox3];
   const int64_t     lb = xcdBlock( ntiles );
   if ( lb >= ntiles )
      return;
   int64_t         full;
   const TileCoord tc   = tiles ? tileFromList( tiles, lb, mm, full ) : tileOf( lb, mm );
   const int64_t   base = tc.c * sc1;
   const int32_t   top  = int32_t( 2 * mm );
   const double*   srcC = src + base;
   // the thread's (at most two) box points, both loads in flight before the LDS writes
   static_assert( kBox3 <= 2 * kTileThreads, "two box points per thread" );
   double bv[2];
#pragma unroll
   for ( int j = 0; j < 2; ++j )
   {
      const int     p  = int( threadIdx.x ) + j * kTileThreads;
      const int32_t lx = p % kBox, ly = ( p / kBox ) % kBox, lz = p / ( kBox * kBox );
      const int32_t X = 2 * int32_t( tc.x0 ) + lx, Y = 2 * int32_t( tc.y0 ) + ly, Z = 2 * int32_t( tc.z0 ) + lz;
      bv[j]           = ( p < kBox3 && X + Y + Z <= top ) ? srcC[tetRowStart32( int32_t( w1 ), Y, Z ) + X] : 0.0;
   }
#pragma unroll
   for ( int j = 0; j < 2; ++j )
   {
      const int p = int( threadIdx.x ) + j * kTileThreads;
      if ( p < kBox3 )
         sb[p] = bv[j];
   }
   __syncthreads();
   // the thread's type and anchor (the geometry itself is not needed)
   TileThread< double > th;
   th.t         = __builtin_amdgcn_readfirstlane( int( threadIdx.x ) >> 6 );
   const int an = int( threadIdx.x ) & 63;
   th.ax        = an & 3;
   th.ay        = ( an >> 2 ) & 3;
   th.az        = an >> 4;
   th.valid     = tc.x0 + th.ax + tc.y0 + th.ay + tc.z0 + th.az <= mm - 1 - MicroCells::rowShift[th.t];
   int    pos[10];
   double u[10];
   elementBoxPositions( th, pos );
#pragma unroll
   for ( int i = 0; i < 10; ++i )
      u[i] = th.valid ? sb[pos[i]] : 0.0;
   __syncthreads();
   for ( int p = threadIdx.x; p < kBox3; p += kTileThreads )
      sb[p] = 0.0;
   __syncthreads();
   if ( th.valid )
   {
      const double* A = Ae + ( tc.c * 6 + th.t ) * 100;
#pragma unroll
      for ( int i = 0; i < 10; ++i )
      {
         double y = 0.0;
#pragma unroll
         for ( int j = 0; j < 10; ++j )
            y += A[10 * i + j] * u[j];
         atomicAdd( &sb[pos[i]], y );
      }
   }
   __syncthreads();
   double* dstC = dst + base;
   for ( int p = threadIdx.x; p < kBox3; p += kTileThreads )
   {
      const int32_t lx = p % kBox, ly = ( p / kBox ) % kBox, lz = p / ( kBox * kBox );
      const int32_t X = 2 * int32_t( tc.x0 ) + lx, Y = 2 * int32_t( tc.y0 ) + ly, Z = 2 * int32_t( tc.z0 ) + lz;
      if ( X + Y + Z <= top )
         sharedAdd( dstC + tetRowStart32( int32_t( w1 ), Y, Z ) + X, sb[p] );
   }
}

// assembled row . x of the cell-interior point (X,Y,Z) of cell c, and its diagonal
__device__ __forceinline__ double rowDot( const double* __restrict__ Ae, int64_t c, int32_t w1, int32_t X, int32_t Y,
                                          int32_t Z, const double* __restrict__ xc, double& diag )
{
   const int cls = ( X & 1 ) | ( ( Y & 1 ) << 1 ) | ( ( Z & 1 ) << 2 );
   const int px = X & 1, py = Y & 1, pz = Z & 1;
   double    acc = 0.0;
   diag          = 0.0;
   for ( int k = 0; k < cIncN[cls]; ++k )
   {
      const int*    e   = cInc[cls][k];
      const int     t   = e[0], i = e[1];
      const double* row = Ae + ( c * 6 + t ) * 100 + 10 * i;
      const int32_t bx = X - px + 2 * e[2], by = Y - py + 2 * e[3], bz = Z - pz + 2 * e[4];
#pragma unroll
      for ( int j = 0; j < 10; ++j )
      {
         const int32_t qx = bx + cLoc[t][j][0], qy = by + cLoc[t][j][1], qz = bz + cLoc[t][j][2];
         acc += row[j] * xc[tetRowStart32( w1, qy, qz ) + qx];
      }
      diag += row[i];
   }
   return acc;
}

// one colour (X,Y,Z mod 3 = cx,cy,cz) of SOR on the cell-interior points, in place
__global__ void kP2ConstCellSOR( const double* __restrict__ Ae, int64_t mm, int64_t sc1, int64_t w1, int cx, int cy,
                                 int cz, int32_t nb3, double relax, const double* __restrict__ b, double* __restrict__ x )
{
   const int64_t c   = blockIdx.y;
   const int32_t tid = int32_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( tid >= nb3 * nb3 * nb3 )
      return;
   const int32_t X = 3 * ( tid % nb3 ) + cx, Y = 3 * ( ( tid / nb3 ) % nb3 ) + cy, Z = 3 * ( tid / ( nb3 * nb3 ) ) + cz;
   const int32_t top = int32_t( 2 * mm );
   if ( X < 1 || Y < 1 || Z < 1 || X + Y + Z > top - 1 )
      return;
   double*       xc   = x + c * sc1;
   const int64_t self = tetRowStart32( int32_t( w1 ), Y, Z ) + X;
   double        diag;
   const double  ax = rowDot( Ae, c, int32_t( w1 ), X, Y, Z, xc, diag );
   xc[self] += relax * ( b[c * sc1 + self] - ax ) / diag;
}

// inverse diagonal of the cell-interior points (halo and boundary positions untouched)
__global__ void kP2ConstCellInvDiag( const double* __restrict__ Ae, int64_t mm, int64_t sc1, int64_t w1, int32_t n3,
                                     double* __restrict__ d )
{
   const int64_t c   = blockIdx.y;
   const int32_t tid = int32_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int32_t w   = int32_t( w1 );
   if ( tid >= n3 * n3 * n3 )
      return;
   const int32_t X = tid % n3, Y = ( tid / n3 ) % n3, Z = tid / ( n3 * n3 );
   if ( X < 1 || Y < 1 || Z < 1 || X + Y + Z > int32_t( 2 * mm ) - 1 )
      return;
   const int cls = ( X & 1 ) | ( ( Y & 1 ) << 1 ) | ( ( Z & 1 ) << 2 );
   double    dg  = 0.0;
   for ( int k = 0; k < cIncN[cls]; ++k )
      dg += Ae[( c * 6 + cInc[cls][k][0] ) * 100 + 11 * cInc[cls][k][1]];
   d[c * sc1 + tetRowStart32( w, Y, Z ) + X] = 1.0 / dg;
}

// SOR on the owned boundary DoFs of one colour through their assembled rows
// (neighbour positions: owner slots of boundary DoFs, cell slots of interior ones)
__global__ void kP2ConstBndSOR( int64_t n, const int32_t* __restrict__ list, int64_t ownerBase,
                                const uint8_t* __restrict__ bndFlag, int flag, const int64_t* __restrict__ ptr,
                                const int64_t* __restrict__ pos, const double* __restrict__ val,
                                const double* __restrict__ invDiag, double relax, const double* __restrict__ b,
                                double* __restrict__ x )
{
   const int64_t k = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( k >= n )
      return;
   const int32_t r = list[k];
   if ( !( bndFlag[r] & flag ) )
      return;
   double acc = 0.0;
   for ( int64_t q = ptr[r]; q < ptr[r + 1]; ++q )
      acc += val[q] * x[pos[q]];
   const double d = 1.0 / invDiag[r];
   const double o = x[ownerBase + r];
   x[ownerBase + r] = o + relax * ( b[ownerBase + r] - acc - d * o ) * invDiag[r];
}

// inverse diagonal of the owned boundary DoFs into their owner slots
__global__ void kP2ConstBndInvDiag( int64_t nb, int64_t ownerBase, const double* __restrict__ invDiag,
                                    double* __restrict__ d )
{
   const int64_t r = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( r < nb )
      d[ownerBase + r] = invDiag[r];
}
} // namespace

struct P2ConstLevel
{
   DevBuf< double >       Ae;      // nc * 6 * 100 element matrices
   DevBuf< double >       invDiag; // owned boundary DoFs of the level-(L+1) lattice
   DevBuf< int64_t >      csrPtr, csrPos;
   DevBuf< double >       csrVal;
   std::vector< int64_t > colorStart;
   DevBuf< int32_t >      colorRows;
};

struct P2ConstOperator : OperatorBase
{
   Storage*                      st = nullptr;
   int                           minL = 0, maxL = 0;
   std::map< int, P2ConstLevel > lv;
};

namespace {
void buildLevel( Context* ctx, Storage& st, int L, P2ConstLevel& o )
{
   const Topology&   t   = st.topo;
   Lattice&          lat = st.lattice( ctx, L + 1 );
   const int64_t     mm  = int64_t( 1 ) << L;
   const IncTables&  T   = incTables();
   std::vector< double > Ae( size_t( t.nc ) * 600 );
   for ( int c = 0; c < t.nc; ++c )
   {
      const double* v[4];
      for ( int k = 0; k < 4; ++k )
         v[k] = &t.X[3 * t.cv[c][k]];
      for ( int ty = 0; ty < 6; ++ty )
      {
         double p[4][3];
         for ( int k = 0; k < 4; ++k )
            for ( int d = 0; d < 3; ++d )
               p[k][d] = v[0][d] + ( double( kVofHost[ty][k][0] ) * ( v[1][d] - v[0][d] ) +
                                     double( kVofHost[ty][k][1] ) * ( v[2][d] - v[0][d] ) +
                                     double( kVofHost[ty][k][2] ) * ( v[3][d] - v[0][d] ) ) /
                                        double( mm );
         double A[10][10];
         p2LaplaceElement( p, A );
         for ( int i = 0; i < 10; ++i )
            for ( int j = 0; j < 10; ++j )
               Ae[( size_t( c ) * 6 + ty ) * 100 + 10 * i + j] = A[i][j];
      }
   }
   o.Ae.upload( ctx, Ae );
   // assembled rows of the owned boundary DoFs from the micro-cells touching the
   // macro-cell boundary (every row of a boundary DoF only involves those)
   const int64_t ownerBase = lat.nc * lat.sc, w1 = lat.w;
   std::vector< double >                                     diag( lat.nb, 0.0 );
   std::vector< std::map< int64_t, double > >                row( lat.nb );
   std::vector< std::vector< int32_t > >                     adj( lat.nb );
   for ( int c = 0; c < t.nc; ++c )
      for ( int ty = 0; ty < 6; ++ty )
      {
         const int64_t wt = mm - MicroCells::rowShift[ty];
         for ( int64_t z = 0; z < wt; ++z )
            for ( int64_t y = 0; y + z < wt; ++y )
               for ( int64_t x = 0; x + y + z < wt; ++x )
               {
                  // only micro-cells with a vertex on the macro-cell boundary
                  if ( x > 0 && y > 0 && z > 0 && x + y + z + MicroCells::rowShift[ty] < mm - 1 )
                     continue;
                  int64_t P[10][3], id[10];
                  bool    bnd[10], any = false;
                  for ( int j = 0; j < 10; ++j )
                  {
                     P[j][0] = 2 * x + T.loc[ty][j][0];
                     P[j][1] = 2 * y + T.loc[ty][j][1];
                     P[j][2] = 2 * z + T.loc[ty][j][2];
                     id[j]   = lat.num.locate( t, c, P[j][0], P[j][1], P[j][2] );
                     bnd[j]  = id[j] < lat.nb;
                     any     = any || bnd[j];
                  }
                  if ( !any )
                     continue;
                  const double* A = &Ae[( size_t( c ) * 6 + ty ) * 100];
                  for ( int i = 0; i < 10; ++i )
                  {
                     if ( !bnd[i] )
                        continue;
                     for ( int j = 0; j < 10; ++j )
                     {
                        if ( id[j] == id[i] )
                        {
                           diag[id[i]] += A[10 * i + j];
                           continue;
                        }
                        const int64_t q = bnd[j] ? ownerBase + id[j] : c * lat.sc + tetIndex( w1, P[j][0], P[j][1], P[j][2] );
                        row[id[i]][q] += A[10 * i + j];
                        if ( bnd[j] )
                           adj[id[i]].push_back( int32_t( id[j] ) );
                     }
                  }
               }
      }
   std::vector< int64_t > ptr( lat.nb + 1, 0 ), pos;
   std::vector< double >  val, inv( lat.nb );
   for ( int64_t r = 0; r < lat.nb; ++r )
   {
      for ( const auto& e : row[r] )
      {
         pos.push_back( e.first );
         val.push_back( e.second );
      }
      ptr[r + 1] = int64_t( pos.size() );
      inv[r]     = 1.0 / diag[r];
   }
   o.csrPtr.upload( ctx, ptr );
   o.csrPos.upload( ctx, pos.empty() ? std::vector< int64_t >( 1, 0 ) : pos );
   o.csrVal.upload( ctx, val.empty() ? std::vector< double >( 1, 0.0 ) : val );
   o.invDiag.upload( ctx, inv.empty() ? std::vector< double >( 1, 0.0 ) : inv );
   // greedy colouring in canonical order
   std::vector< int >  color( lat.nb, -1 );
   int                 ncolor = 0;
   std::vector< char > used;
   for ( int64_t r = 0; r < lat.nb; ++r )
   {
      used.assign( ncolor + 1, 0 );
      for ( int32_t q : adj[r] )
         if ( color[q] >= 0 )
            used[color[q]] = 1;
      int cc = 0;
      while ( used[cc] )
         ++cc;
      color[r] = cc;
      ncolor   = std::max( ncolor, cc + 1 );
   }
   std::vector< std::vector< int32_t > > byColor( ncolor );
   for ( int64_t r = 0; r < lat.nb; ++r )
      byColor[color[r]].push_back( int32_t( r ) );
   std::vector< int32_t > rows;
   o.colorStart.assign( ncolor + 1, 0 );
   for ( int cc = 0; cc < ncolor; ++cc )
   {
      o.colorStart[cc] = int64_t( rows.size() );
      rows.insert( rows.end(), byColor[cc].begin(), byColor[cc].end() );
   }
   o.colorStart[ncolor] = int64_t( rows.size() );
   o.colorRows.upload( ctx, rows.empty() ? std::vector< int32_t >( 1, 0 ) : rows );
}

void checkFn( const P2ConstOperator& op, const Function& f, int L )
{
   if ( f.st != op.st || f.off != 1 || f.lv != 0 || f.dtype != 0 )
      throw Error( HG_ERR_ARG, "the P2 constant operator works on FP64 P2 functions of its storage" );
   if ( L < op.minL || L > op.maxL || L < f.minL || L > f.maxL )
      throw Error( HG_ERR_ARG, "level outside the operator's or function's level range" );
}

int32_t boxCount( int64_t w ) { return int32_t( ( w + 2 ) / 3 ); }
} // namespace
} // namespace hg

using namespace hg;

extern "C" {

int hg_p2_const_create( hg_ctx* h, hg_handle s, int minLevel, int maxLevel, hg_handle* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !out || minLevel < 0 || maxLevel < minLevel || maxLevel > 9 )
         throw Error( HG_ERR_ARG, "invalid P2 constant operator levels" );
      static std::once_flag tablesOnce;
      std::call_once( tablesOnce, [] {
         const IncTables& T = incTables();
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cIncN ), T.n, sizeof( T.n ) ) );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cInc ), T.e, sizeof( T.e ) ) );
         HG_CHECK_HIP( hipMemcpyToSymbol( HIP_SYMBOL( cLoc ), T.loc, sizeof( T.loc ) ) );
      } );
      Storage& st = ctx->storage( s );
      auto     op = std::make_unique< P2ConstOperator >();
      op->st      = &st;
      op->minL    = minLevel;
      op->maxL    = maxLevel;
      for ( int L = minLevel; L <= maxLevel; ++L )
      {
         st.lattice( ctx, L );
         buildLevel( ctx, st, L, op->lv[L] );
      }
      *out                  = ctx->nextHandle++;
      ctx->operators[*out] = std::move( op );
   } );
}

int hg_p2_const_apply( hg_ctx* h, hg_handle opH, hg_handle src, hg_handle dst, int level, int flag, int update )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( update != HG_REPLACE && update != HG_ADD )
         throw Error( HG_ERR_ARG, "invalid update type" );
      P2ConstOperator& op = ctx->op< P2ConstOperator >( opH );
      Function&        s  = ctx->function( src );
      Function&        d  = ctx->function( dst );
      checkFn( op, s, level );
      checkFn( op, d, level );
      if ( &s == &d )
         throw Error( HG_ERR_ARG, "apply needs distinct src and dst" );
      Storage&      st  = *op.st;
      Lattice&      lat = st.lattice( ctx, level + 1 );
      const int64_t mm  = int64_t( 1 ) << level;
      if ( lat.w > 1024 )
         throw Error( HG_ERR_UNSUPPORT, "P2 apply supports levels up to 9 (32-bit lattice offsets)" );
      double* sp = (double*) s.ptr( level );
      double* dp = (double*) d.ptr( level );
      pullHalos( ctx, st, lat, sp, 0 );
      if ( update == HG_REPLACE && ( flag & kInner ) )
         kFillRange< double ><<< gridFor( lat.nc * lat.sc ), kBlock, 0, ctx->stream >>>( lat.nc * lat.sc, dp, 0.0 );
      else if ( lat.haloPerCell() > 0 )
         kZeroHalos< double ><<< gridFor( lat.haloPerCell() * lat.nc ), kBlock, 0, ctx->stream >>>(
             BndGeom::of( st, lat ), lat.nc, dp );
      HG_CHECK_HIP( hipGetLastError() );
      const int64_t   ntiles = tilesPerCell( mm ) * lat.nc;
      const uint64_t* tiles  = tileList( ctx, lat, mm, kTilesAll );
      ctx->beginTiming();
      kP2ConstTileApply<<< xcdGrid( ntiles ), kTileThreads, 0, ctx->stream >>>( op.lv.at( level ).Ae.p, mm, ntiles,
                                                                                lat.sc, lat.w, sp, dp, tiles );
      HG_CHECK_HIP( hipGetLastError() );
      ctx->endTiming();
      pushAdditive( ctx, st, lat, dp, 0, kAll ^ flag, update == HG_REPLACE );
   } );
}

// SOR in multicolour order: forward, the owned macro vertex/edge/face DoFs
// colour by colour, then the 27 cell-interior colours k = (X%3) + 3 (Y%3) +
// 9 (Z%3); backward (smooth_sor_backwards, P2ConstantOperator.hpp:95-106) the
// same colours in reverse order, the cell interiors first.
static void p2ConstSOR( Context* ctx, P2ConstOperator& op, Function& x, Function& b, int level, int flag, double relax,
                        bool backward )
{
   checkFn( op, x, level );
   checkFn( op, b, level );
   if ( &x == &b )
      throw Error( HG_ERR_ARG, "smooth needs distinct x and b" );
   Storage&      st  = *op.st;
   Lattice&      lat = st.lattice( ctx, level + 1 );
   P2ConstLevel& o   = op.lv.at( level );
   const int64_t mm  = int64_t( 1 ) << level;
   double*       xp  = (double*) x.ptr( level );
   const double* bp  = (const double*) b.ptr( level );
   pullHalos( ctx, st, lat, xp, 0 );
   const size_t nbc = o.colorStart.empty() ? 0 : o.colorStart.size() - 1;
   auto boundaryColour = [&]( size_t cc ) {
      const int64_t n = o.colorStart[cc + 1] - o.colorStart[cc];
      if ( n == 0 )
         return;
      kP2ConstBndSOR<<< gridFor( n ), kBlock, 0, ctx->stream >>>( n, o.colorRows.p + o.colorStart[cc], lat.nc * lat.sc,
                                                                   lat.bndFlag.p, flag, o.csrPtr.p, o.csrPos.p,
                                                                   o.csrVal.p, o.invDiag.p, relax, bp, xp );
      HG_CHECK_HIP( hipGetLastError() );
   };
   const int32_t nb3 = boxCount( lat.w );
   const dim3    grid( unsigned( ( int64_t( nb3 ) * nb3 * nb3 + 255 ) / 256 ), unsigned( lat.nc ) );
   auto interiorColour = [&]( int k ) {
      kP2ConstCellSOR<<< grid, 256, 0, ctx->stream >>>( o.Ae.p, mm, lat.sc, lat.w, k % 3, ( k / 3 ) % 3, k / 9, nb3,
                                                         relax, bp, xp );
      HG_CHECK_HIP( hipGetLastError() );
   };
   if ( !backward )
   {
      for ( size_t cc = 0; cc < nbc; ++cc )
         boundaryColour( cc );
      // the cell interiors read the new boundary values through their halos
      pullHalos( ctx, st, lat, xp, 0 );
      if ( flag & kInner )
         for ( int k = 0; k < 27; ++k )
            interiorColour( k );
   }
   else
   {
      if ( flag & kInner )
         for ( int k = 26; k >= 0; --k )
            interiorColour( k );
      for ( size_t cc = nbc; cc-- > 0; )
         boundaryColour( cc );
      pullHalos( ctx, st, lat, xp, 0 );
   }
}

int hg_p2_const_smooth_sor( hg_ctx* h, hg_handle opH, hg_handle xH, hg_handle bH, int level, int flag, double relax )
{
   return hg_guard( h, [&]( Context* ctx ) {
      p2ConstSOR( ctx, ctx->op< P2ConstOperator >( opH ), ctx->function( xH ), ctx->function( bH ), level, flag, relax,
                  false );
   } );
}

int hg_p2_const_smooth_sor_backwards( hg_ctx* h, hg_handle opH, hg_handle xH, hg_handle bH, int level, int flag,
                                      double relax )
{
   return hg_guard( h, [&]( Context* ctx ) {
      p2ConstSOR( ctx, ctx->op< P2ConstOperator >( opH ), ctx->function( xH ), ctx->function( bH ), level, flag, relax,
                  true );
   } );
}

int hg_p2_const_sor_colours( hg_ctx* h, hg_handle opH, int level, int32_t* colour, int* nBoundaryColours )
{
   return hg_guard( h, [&]( Context* ctx ) {
      P2ConstOperator& op = ctx->op< P2ConstOperator >( opH );
      if ( !colour || !nBoundaryColours )
         throw Error( HG_ERR_ARG, "sor_colours: null output" );
      if ( level < op.minL || level > op.maxL )
         throw Error( HG_ERR_ARG, "level outside the operator's level range" );
      P2ConstLevel&   o   = op.lv.at( level );
      Lattice&        lat = op.st->lattice( ctx, level + 1 );
      const Topology& t   = op.st->topo;
      const int       nbc = o.colorStart.empty() ? 0 : int( o.colorStart.size() ) - 1;
      std::vector< int32_t > rows;
      if ( lat.nb > 0 )
      {
         rows.resize( size_t( lat.nb ) );
         HG_CHECK_HIP( hipMemcpy( rows.data(), o.colorRows.p, size_t( lat.nb ) * sizeof( int32_t ), hipMemcpyDeviceToHost ) );
      }
      for ( int cc = 0; cc < nbc; ++cc )
         for ( int64_t k = o.colorStart[cc]; k < o.colorStart[cc + 1]; ++k )
            colour[rows[size_t( k )]] = cc;
      const int64_t mm = lat.mm;
      for ( int c = 0; c < t.nc; ++c )
         for ( int64_t z = 1; z < mm; ++z )
            for ( int64_t y = 1; y + z < mm; ++y )
               for ( int64_t x = 1; x + y + z < mm; ++x )
                  colour[lat.num.locate( t, c, x, y, z )] = nbc + int32_t( x % 3 + 3 * ( y % 3 ) + 9 * ( z % 3 ) );
      *nBoundaryColours = nbc;
   } );
}

int hg_p2_const_inverse_diagonal( hg_ctx* h, hg_handle opH, int level, hg_handle dst )
{
   return hg_guard( h, [&]( Context* ctx ) {
      P2ConstOperator& op = ctx->op< P2ConstOperator >( opH );
      Function&        d  = ctx->function( dst );
      checkFn( op, d, level );
      Storage&      st  = *op.st;
      Lattice&      lat = st.lattice( ctx, level + 1 );
      P2ConstLevel& o   = op.lv.at( level );
      double*       dp  = (double*) d.ptr( level );
      const int32_t n3  = int32_t( lat.w );
      const dim3    grid( unsigned( ( int64_t( n3 ) * n3 * n3 + 255 ) / 256 ), unsigned( lat.nc ) );
      kP2ConstCellInvDiag<<< grid, 256, 0, ctx->stream >>>( o.Ae.p, int64_t( 1 ) << level, lat.sc, lat.w, n3, dp );
      if ( lat.nb > 0 )
         kP2ConstBndInvDiag<<< gridFor( lat.nb ), kBlock, 0, ctx->stream >>>( lat.nb, lat.nc * lat.sc, o.invDiag.p, dp );
      HG_CHECK_HIP( hipGetLastError() );
      pullHalos( ctx, st, lat, dp, 0 );
   } );
}

} // extern "C"
