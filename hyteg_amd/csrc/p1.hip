// P1 constant-stencil operator (apply, multicolour SOR), P1 linear restriction
// and prolongation: host set-up, HIP kernels and C-ABI entry points.
//
// Work split per level (HBM-bound stencil work, no MFMA):
//  * cell interiors: one structured 15-point stencil per macro-cell, swept row
//    by row ((y,z) rows, x contiguous) -- apply_3D_macrocell_vertexdof_to_vertexdof
//    (P1generatedKernels/apply_3D_macrocell_vertexdof_to_vertexdof_replace.cpp:34-79);
//  * DoFs owned by macro-vertices/edges/faces: sparse rows over owner storage
//    positions, assembled from the micro-cells of all neighbour cells, i.e. the
//    macro-face/edge/vertex stencils of P1Operator (P1ConstantOperator.cpp:239-358,
//    P1Operator.hpp:211-245, 1044-1108).
#include <algorithm>
#include <array>
#include <cmath>
#include <map>

#include "abi_util.hpp"
#include "hyteg_amd.h"
#include "runtime.hpp"

namespace hg {

__constant__ int cOff15[15][3] = { { 0, 0, 0 },  { -1, 0, 0 }, { -1, 0, 1 }, { -1, 1, -1 }, { -1, 1, 0 },
                                   { 0, -1, 0 }, { 0, -1, 1 }, { 0, 0, -1 }, { 0, 0, 1 },   { 0, 1, -1 },
                                   { 0, 1, 0 },  { 1, -1, 0 }, { 1, -1, 1 }, { 1, 0, -1 },  { 1, 0, 0 } };
// for a fine lattice point with parity pattern p = (x&1) + 2(y&1) + 4(z&1) != 0, the
// stencil direction d such that the point is the midpoint of coarse edge (q-d, q+d)
__constant__ int cMidDir[8][3] = { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 1, -1, 0 },
                                   { 0, 0, 1 }, { 1, 0, -1 }, { 0, 1, -1 }, { 1, -1, 1 } };

static const int kMidDirHost[8][3] = { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 1, -1, 0 },
                                       { 0, 0, 1 }, { 1, 0, -1 }, { 0, 1, -1 }, { 1, -1, 1 } };

// ------------------------------------------------------------------------
// host: P1 element stiffness of a tetrahedron, gradients of the barycentric
// coordinates from face normals: grad l_i = n_i / (n_i . (p_i - p_j)) where
// n_i is the normal of the face opposite vertex i.
// ------------------------------------------------------------------------
static void cross3( const double a[3], const double b[3], double o[3] )
{
   o[0] = a[1] * b[2] - a[2] * b[1];
   o[1] = a[2] * b[0] - a[0] * b[2];
   o[2] = a[0] * b[1] - a[1] * b[0];
}
static void p1ElementStiffness( const double p[4][3], double K[4][4] )
{
   double g[4][3];
   double vol6 = 0.0;
   for ( int i = 0; i < 4; ++i )
   {
      const int j = ( i + 1 ) % 4, k = ( i + 2 ) % 4, l = ( i + 3 ) % 4;
      double    e1[3], e2[3], n[3], d[3];
      for ( int a = 0; a < 3; ++a )
      {
         e1[a] = p[k][a] - p[j][a];
         e2[a] = p[l][a] - p[j][a];
         d[a]  = p[i][a] - p[j][a];
      }
      cross3( e1, e2, n );
      const double s = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
      for ( int a = 0; a < 3; ++a )
         g[i][a] = n[a] / s;
      if ( i == 0 )
         vol6 = std::fabs( s );
   }
   const double vol = vol6 / 6.0;
   for ( int i = 0; i < 4; ++i )
      for ( int j = 0; j < 4; ++j )
         K[i][j] = vol * ( g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2] );
}

// element matrices of the 6 micro-cell types of macro-cell c on level L (affine cell)
static void microTypeMatrices( const Topology& t, int c, int64_t mm, double K[6][4][4] )
{
   const double* v[4];
   for ( int k = 0; k < 4; ++k )
      v[k] = &t.X[3 * t.cv[c][k]];
   for ( int ty = 0; ty < 6; ++ty )
   {
      double p[4][3];
      for ( int k = 0; k < 4; ++k )
      {
         const int* o = MicroCells::vof[ty][k];
         for ( int a = 0; a < 3; ++a )
            p[k][a] = v[0][a] + ( o[0] / double( mm ) ) * ( v[1][a] - v[0][a] ) + ( o[1] / double( mm ) ) * ( v[2][a] - v[0][a] ) +
                      ( o[2] / double( mm ) ) * ( v[3][a] - v[0][a] );
      }
      p1ElementStiffness( p, K[ty] );
   }
}

static int slot15( int dx, int dy, int dz )
{
   for ( int s = 0; s < 15; ++s )
      if ( Stencil15::off[s][0] == dx && Stencil15::off[s][1] == dy && Stencil15::off[s][2] == dz )
         return s;
   throw Error( -4, "offset outside the 15-point stencil" );
}

static bool onCellBoundary( int64_t mm, int64_t x, int64_t y, int64_t z )
{
   return x == 0 || y == 0 || z == 0 || x + y + z == mm;
}

// sparse rows keyed by owner storage position
struct RowBuilder
{
   std::vector< std::vector< std::pair< int64_t, double > > > rows;
   explicit RowBuilder( int64_t n )
   : rows( n )
   {}
   void add( int64_t r, int64_t col, double v )
   {
      for ( auto& e : rows[r] )
         if ( e.first == col )
         {
            e.second += v;
            return;
         }
      rows[r].emplace_back( col, v );
   }
   void set( int64_t r, int64_t col, double v )
   {
      for ( auto& e : rows[r] )
         if ( e.first == col )
         {
            e.second = v;
            return;
         }
      rows[r].emplace_back( col, v );
   }
};

struct DevRows
{
   int64_t           nrows = 0;
   DevBuf< int64_t > ptr, col;
   DevBuf< double >  val;
   void              build( Context* ctx, RowBuilder& rb )
   {
      nrows = int64_t( rb.rows.size() );
      std::vector< int64_t > p( nrows + 1, 0 ), c;
      std::vector< double >  v;
      for ( int64_t r = 0; r < nrows; ++r )
      {
         std::sort( rb.rows[r].begin(), rb.rows[r].end() );
         for ( auto& e : rb.rows[r] )
         {
            c.push_back( e.first );
            v.push_back( e.second );
         }
         p[r + 1] = int64_t( c.size() );
      }
      ptr.upload( ctx, p );
      col.upload( ctx, c );
      val.upload( ctx, v );
   }
};

struct P1OpLevel
{
   std::vector< double > hStencil; // nc*15
   DevBuf< double >      stencil;
   DevRows               bnd;      // rows of owned vertex/edge/face DoFs
   DevBuf< double >      invDiag;
   std::vector< int64_t > colorStart; // rows grouped by colour
   DevBuf< int32_t >     colorRows;
};

struct P1Operator : OperatorBase
{
   Storage*                   st = nullptr;
   int                        minL = 0, maxL = 0;
   std::map< int, P1OpLevel > lv;
};

static void buildP1Level( Context* ctx, Storage& st, int L, P1OpLevel& o )
{
   const Topology& t   = st.topo;
   Lattice&        lat = st.lattice( ctx, L );
   const int64_t   mm  = lat.mm;
   o.hStencil.assign( 15 * size_t( t.nc ), 0.0 );
   RowBuilder rb( lat.nb );
   for ( int c = 0; c < t.nc; ++c )
   {
      double K[6][4][4];
      microTypeMatrices( t, c, mm, K );
      // inner stencil: the 24 micro-cells around a point (P1Elements.hpp:627-700)
      for ( int ty = 0; ty < 6; ++ty )
         for ( int i = 0; i < 4; ++i )
            for ( int j = 0; j < 4; ++j )
            {
               const int* a = MicroCells::vof[ty][i];
               const int* b = MicroCells::vof[ty][j];
               o.hStencil[15 * c + slot15( b[0] - a[0], b[1] - a[1], b[2] - a[2] )] += K[ty][i][j];
            }
      // rows of the owned DoFs on this cell's boundary: this cell's share
      for ( int64_t z = 0; z <= mm; ++z )
         for ( int64_t y = 0; y + z <= mm; ++y )
            for ( int64_t x = 0; x + y + z <= mm; ++x )
            {
               if ( !onCellBoundary( mm, x, y, z ) )
                  continue;
               const int64_t r = lat.num.locate( t, c, x, y, z );
               for ( int ty = 0; ty < 6; ++ty )
               {
                  const int64_t wt = mm - MicroCells::rowShift[ty];
                  for ( int i = 0; i < 4; ++i )
                  {
                     const int*    vi = MicroCells::vof[ty][i];
                     const int64_t ax = x - vi[0], ay = y - vi[1], az = z - vi[2];
                     if ( ax < 0 || ay < 0 || az < 0 || ax + ay + az > wt - 1 )
                        continue;
                     for ( int j = 0; j < 4; ++j )
                     {
                        const int* vj = MicroCells::vof[ty][j];
                        rb.add( r, lat.ownerPos( t, c, ax + vj[0], ay + vj[1], az + vj[2] ), K[ty][i][j] );
                     }
                  }
               }
            }
   }
   o.stencil.upload( ctx, o.hStencil );

   // inverse diagonals and a greedy colouring of the owned-boundary rows
   const int64_t           ncell = lat.nc * lat.sc;
   std::vector< double >   inv( lat.nb, 0.0 );
   std::vector< int >      color( lat.nb, -1 );
   int                     ncolor = 0;
   for ( int64_t r = 0; r < lat.nb; ++r )
   {
      std::vector< char > used( 64, 0 );
      for ( auto& e : rb.rows[r] )
      {
         if ( e.first == ncell + r )
            inv[r] = 1.0 / e.second;
         else if ( e.first >= ncell && color[e.first - ncell] >= 0 )
         {
            const int cc = color[e.first - ncell];
            if ( cc >= int( used.size() ) )
               used.resize( cc + 1, 0 );
            used[cc] = 1;
         }
      }
      int cc = 0;
      while ( cc < int( used.size() ) && used[cc] )
         ++cc;
      color[r] = cc;
      ncolor   = std::max( ncolor, cc + 1 );
   }
   o.colorStart.assign( ncolor + 1, 0 );
   std::vector< int32_t > rowsByColor;
   for ( int cc = 0; cc < ncolor; ++cc )
   {
      o.colorStart[cc] = int64_t( rowsByColor.size() );
      for ( int64_t r = 0; r < lat.nb; ++r )
         if ( color[r] == cc )
            rowsByColor.push_back( int32_t( r ) );
   }
   o.colorStart[ncolor] = int64_t( rowsByColor.size() );
   o.colorRows.upload( ctx, rowsByColor );
   o.invDiag.upload( ctx, inv );
   o.bnd.build( ctx, rb );
}

// ------------------------------------------------------------------------
// transfer set-up (per storage, per fine level)
// ------------------------------------------------------------------------
struct P1Transfer : OperatorBase
{
   DevRows restrictRows; // coarse owned rows <- fine storage
   DevRows prolongRows;  // fine owned rows   <- coarse storage
};

static P1Transfer& transferFor( Context* ctx, Storage& st, int fineL )
{
   auto it = st.transfers.find( fineL );
   if ( it != st.transfers.end() )
      return static_cast< P1Transfer& >( *it->second );
   const Topology& t  = st.topo;
   Lattice&        lf = st.lattice( ctx, fineL );
   Lattice&        lc = st.lattice( ctx, fineL - 1 );
   RowBuilder      rr( lc.nb ), pr( lf.nb );
   for ( int c = 0; c < t.nc; ++c )
   {
      // R = P^T on coarse DoFs owned by vertices/edges/faces: centre weight 1,
      // the 14 fine neighbours in the cell closure weight 1/2, every
      // (coarse, fine) pair counted once over all cells -- what the additive
      // communication of restrict_3D_macrocell_P1_pull_additive yields
      for ( int64_t z = 0; z <= lc.mm; ++z )
         for ( int64_t y = 0; y + z <= lc.mm; ++y )
            for ( int64_t x = 0; x + y + z <= lc.mm; ++x )
            {
               if ( !onCellBoundary( lc.mm, x, y, z ) )
                  continue;
               const int64_t r = lc.num.locate( t, c, x, y, z );
               rr.set( r, lf.ownerPos( t, c, 2 * x, 2 * y, 2 * z ), 1.0 );
               for ( int s = 1; s < 15; ++s )
               {
                  const int64_t qx = 2 * x + Stencil15::off[s][0], qy = 2 * y + Stencil15::off[s][1],
                                qz = 2 * z + Stencil15::off[s][2];
                  if ( qx < 0 || qy < 0 || qz < 0 || qx + qy + qz > lf.mm )
                     continue;
                  rr.set( r, lf.ownerPos( t, c, qx, qy, qz ), 0.5 );
               }
            }
      // linear interpolation onto fine DoFs owned by vertices/edges/faces
      for ( int64_t z = 0; z <= lf.mm; ++z )
         for ( int64_t y = 0; y + z <= lf.mm; ++y )
            for ( int64_t x = 0; x + y + z <= lf.mm; ++x )
            {
               if ( !onCellBoundary( lf.mm, x, y, z ) )
                  continue;
               const int64_t r = lf.num.locate( t, c, x, y, z );
               const int     p = int( x & 1 ) + 2 * int( y & 1 ) + 4 * int( z & 1 );
               if ( p == 0 )
               {
                  pr.set( r, lc.ownerPos( t, c, x / 2, y / 2, z / 2 ), 1.0 );
                  continue;
               }
               const int* d = kMidDirHost[p];
               pr.set( r, lc.ownerPos( t, c, ( x - d[0] ) / 2, ( y - d[1] ) / 2, ( z - d[2] ) / 2 ), 0.5 );
               pr.set( r, lc.ownerPos( t, c, ( x + d[0] ) / 2, ( y + d[1] ) / 2, ( z + d[2] ) / 2 ), 0.5 );
            }
   }
   auto tr = std::make_unique< P1Transfer >();
   tr->restrictRows.build( ctx, rr );
   tr->prolongRows.build( ctx, pr );
   auto& ref            = *tr;
   st.transfers[fineL] = std::move( tr );
   return ref;
}

// ------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------
// y[r] (=|+=) sum_k val[k] * x[col[k]] on rows whose owner type matches flag
template < class T >
__global__ void kRowsApply( int64_t nrows, const int64_t* __restrict__ ptr, const int64_t* __restrict__ col,
                            const double* __restrict__ val, const uint8_t* __restrict__ rowFlag, int flag,
                            const T* __restrict__ x, T* __restrict__ y, int add )
{
   using A         = typename Acc< T >::type;
   const int64_t r = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( r >= nrows || !( rowFlag[r] & flag ) )
      return;
   A s = A( 0 );
   for ( int64_t k = ptr[r]; k < ptr[r + 1]; ++k )
      s += A( val[k] ) * toAcc( x[col[k]] );
   if ( add )
      s += toAcc( y[r] );
   y[r] = fromAcc< T >( s );
}

// SOR update of the rows of one colour (owner positions ncell + r)
template < class T >
__global__ void kRowsSOR( int64_t n, const int32_t* __restrict__ rows, const int64_t* __restrict__ ptr,
                          const int64_t* __restrict__ col, const double* __restrict__ val,
                          const double* __restrict__ invDiag, const uint8_t* __restrict__ rowFlag, int flag,
                          int64_t ncell, double relax, const T* __restrict__ b, T* __restrict__ x )
{
   using A         = typename Acc< T >::type;
   const int64_t i = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( i >= n )
      return;
   const int64_t r = rows[i];
   if ( !( rowFlag[r] & flag ) )
      return;
   const int64_t self = ncell + r;
   A             s    = toAcc( b[self] );
   for ( int64_t k = ptr[r]; k < ptr[r + 1]; ++k )
      if ( col[k] != self )
         s -= A( val[k] ) * toAcc( x[col[k]] );
   const A xo = toAcc( x[self] );
   x[self]    = fromAcc< T >( ( A( 1 ) - A( relax ) ) * xo + A( relax ) * A( invDiag[r] ) * s );
}

// 15-point stencil on cell interiors: one block per (interior row, cell)
template < class T >
__global__ void kCellApply( const int32_t* __restrict__ rows, int64_t sc, int64_t w, int64_t mm,
                            const double* __restrict__ stencil, const T* __restrict__ src, T* __restrict__ dst, int add )
{
   using A            = typename Acc< T >::type;
   const int64_t y    = rows[2 * blockIdx.x];
   const int64_t z    = rows[2 * blockIdx.x + 1];
   const int64_t c    = blockIdx.y;
   const int64_t base = c * sc;
   A             st[15];
   int64_t       rb[15];
#pragma unroll
   for ( int s = 0; s < 15; ++s )
   {
      st[s] = A( stencil[15 * c + s] );
      rb[s] = base + tetRowStart( w, y + cOff15[s][1], z + cOff15[s][2] ) + cOff15[s][0];
   }
   const int64_t out = base + tetRowStart( w, y, z );
   for ( int64_t x = 1 + threadIdx.x; x + y + z < mm; x += blockDim.x )
   {
      A v = A( 0 );
#pragma unroll
      for ( int s = 0; s < 15; ++s )
         v += st[s] * toAcc( src[rb[s] + x] );
      if ( add )
         v += toAcc( dst[out + x] );
      dst[out + x] = fromAcc< T >( v );
   }
}

// SOR on the cell-interior points of parity colour (cx, row group) in place
template < class T >
__global__ void kCellSOR( const int32_t* __restrict__ rows, int64_t sc, int64_t w, int64_t mm,
                          const double* __restrict__ stencil, int cx, double relax, const T* __restrict__ b,
                          T* __restrict__ x )
{
   using A            = typename Acc< T >::type;
   const int64_t y    = rows[2 * blockIdx.x];
   const int64_t z    = rows[2 * blockIdx.x + 1];
   const int64_t c    = blockIdx.y;
   const int64_t base = c * sc;
   A             st[15];
   int64_t       rb[15];
#pragma unroll
   for ( int s = 0; s < 15; ++s )
   {
      st[s] = A( stencil[15 * c + s] );
      rb[s] = base + tetRowStart( w, y + cOff15[s][1], z + cOff15[s][2] ) + cOff15[s][0];
   }
   const A       invc = A( 1 ) / st[0];
   const int64_t x0   = cx ? 1 : 2;
   for ( int64_t xx = x0 + 2 * int64_t( threadIdx.x ); xx + y + z < mm; xx += 2 * int64_t( blockDim.x ) )
   {
      A v = toAcc( b[rb[0] + xx] );
#pragma unroll
      for ( int s = 1; s < 15; ++s )
         v -= st[s] * toAcc( x[rb[s] + xx] );
      const A xo    = toAcc( x[rb[0] + xx] );
      x[rb[0] + xx] = fromAcc< T >( ( A( 1 ) - A( relax ) ) * xo + A( relax ) * invc * v );
   }
}

// restriction to coarse cell-interior points (full weighting, R = P^T)
template < class T >
__global__ void kRestrictCell( const int32_t* __restrict__ rows, int64_t scc, int64_t wc, int64_t mmc, int64_t scf,
                               int64_t wf, const T* __restrict__ fine, T* __restrict__ coarse )
{
   using A         = typename Acc< T >::type;
   const int64_t y = rows[2 * blockIdx.x];
   const int64_t z = rows[2 * blockIdx.x + 1];
   const int64_t c = blockIdx.y;
   int64_t       rb[15];
#pragma unroll
   for ( int s = 0; s < 15; ++s )
      rb[s] = c * scf + tetRowStart( wf, 2 * y + cOff15[s][1], 2 * z + cOff15[s][2] ) + cOff15[s][0];
   const int64_t out = c * scc + tetRowStart( wc, y, z );
   for ( int64_t x = 1 + threadIdx.x; x + y + z < mmc; x += blockDim.x )
   {
      A v = A( 0 );
#pragma unroll
      for ( int s = 1; s < 15; ++s )
         v += toAcc( fine[rb[s] + 2 * x] );
      v = toAcc( fine[rb[0] + 2 * x] ) + A( 0.5 ) * v;
      coarse[out + x] = fromAcc< T >( v );
   }
}

// linear interpolation to fine cell-interior points
template < class T >
__global__ void kProlongCell( const int32_t* __restrict__ rows, int64_t scf, int64_t wf, int64_t mmf, int64_t scc,
                              int64_t wc, const T* __restrict__ coarse, T* __restrict__ fine, int add )
{
   using A           = typename Acc< T >::type;
   const int64_t y   = rows[2 * blockIdx.x];
   const int64_t z   = rows[2 * blockIdx.x + 1];
   const int64_t c   = blockIdx.y;
   const int64_t out = c * scf + tetRowStart( wf, y, z );
   for ( int64_t x = 1 + threadIdx.x; x + y + z < mmf; x += blockDim.x )
   {
      const int p = int( x & 1 ) + 2 * int( y & 1 ) + 4 * int( z & 1 );
      A         v;
      if ( p == 0 )
         v = toAcc( coarse[c * scc + tetIndex( wc, x / 2, y / 2, z / 2 )] );
      else
      {
         const int dx = cMidDir[p][0], dy = cMidDir[p][1], dz = cMidDir[p][2];
         v = A( 0.5 ) * ( toAcc( coarse[c * scc + tetIndex( wc, ( x - dx ) / 2, ( y - dy ) / 2, ( z - dz ) / 2 )] ) +
                          toAcc( coarse[c * scc + tetIndex( wc, ( x + dx ) / 2, ( y + dy ) / 2, ( z + dz ) / 2 )] ) );
      }
      if ( add )
         v += toAcc( fine[out + x] );
      fine[out + x] = fromAcc< T >( v );
   }
}

// ------------------------------------------------------------------------
// host drivers
// ------------------------------------------------------------------------
static constexpr int kRowBlock = 64;

static void rowsApply( Context* ctx, const DevRows& R, const uint8_t* rowFlag, int flag, const void* x, void* y,
                       int dtype, int add )
{
   if ( R.nrows == 0 )
      return;
   dispatchType( dtype, [&]( auto tag ) {
      using T = decltype( tag );
      kRowsApply< T ><<< gridFor( R.nrows ), kBlock, 0, ctx->stream >>>( R.nrows, R.ptr.p, R.col.p, R.val.p, rowFlag,
                                                                        flag, (const T*) x, (T*) y, add );
   } );
   HG_CHECK_HIP( hipGetLastError() );
}

static P1OpLevel& levelOf( Context* ctx, P1Operator& op, int L )
{
   if ( L < op.minL || L > op.maxL )
      throw Error( -1, "level outside the operator's level range" );
   auto it = op.lv.find( L );
   if ( it == op.lv.end() )
   {
      P1OpLevel& o = op.lv[L];
      buildP1Level( ctx, *op.st, L, o );
      return o;
   }
   return it->second;
}

void p1Apply( Context* ctx, P1Operator& op, Function& src, Function& dst, int L, int flag, int update )
{
   if ( src.st != op.st || dst.st != op.st || src.dtype != dst.dtype )
      throw Error( -1, "apply needs src/dst on the operator's storage with one value type" );
   if ( &src == &dst )
      throw Error( -1, "apply needs distinct src and dst" );
   P1OpLevel& o   = levelOf( ctx, op, L );
   Lattice&   lat = op.st->lattice( ctx, L );
   void*      s   = src.ptr( L );
   void*      d   = dst.ptr( L );
   ctx->beginTiming();
   pullHalos( ctx, *op.st, lat, s, src.dtype );
   const int64_t ncell = lat.nc * lat.sc;
   dispatchType( src.dtype, [&]( auto tag ) {
      using T = decltype( tag );
      rowsApply( ctx, o.bnd, lat.bndFlag.p, flag, s, (T*) d + ncell, src.dtype, update );
      if ( ( flag & kInner ) && lat.nrows > 0 )
      {
         dim3 grid( unsigned( lat.nrows ), unsigned( lat.nc ) );
         kCellApply< T ><<< grid, kRowBlock, 0, ctx->stream >>>( lat.rows.p, lat.sc, lat.w, lat.mm, o.stencil.p,
                                                               (const T*) s, (T*) d, update );
         HG_CHECK_HIP( hipGetLastError() );
      }
   } );
   ctx->endTiming();
}

void p1SOR( Context* ctx, P1Operator& op, Function& x, Function& b, double relax, int L, int flag )
{
   if ( x.st != op.st || b.st != op.st || x.dtype != b.dtype || &x == &b )
      throw Error( -1, "smooth_sor needs distinct x/b on the operator's storage with one value type" );
   P1OpLevel&    o     = levelOf( ctx, op, L );
   Lattice&      lat   = op.st->lattice( ctx, L );
   const int64_t ncell = lat.nc * lat.sc;
   void*         xp    = x.ptr( L );
   const void*   bp    = b.ptr( L );
   ctx->beginTiming();
   dispatchType( x.dtype, [&]( auto tag ) {
      using T = decltype( tag );
      // owned vertex/edge/face DoFs, colour by colour (rows read owner values and
      // cell interiors directly)
      for ( size_t cc = 0; cc + 1 < o.colorStart.size(); ++cc )
      {
         const int64_t n = o.colorStart[cc + 1] - o.colorStart[cc];
         if ( n == 0 )
            continue;
         kRowsSOR< T ><<< gridFor( n ), kBlock, 0, ctx->stream >>>(
             n, o.colorRows.p + o.colorStart[cc], o.bnd.ptr.p, o.bnd.col.p, o.bnd.val.p, o.invDiag.p, lat.bndFlag.p,
             flag, ncell, relax, (const T*) bp, (T*) xp );
         HG_CHECK_HIP( hipGetLastError() );
      }
      if ( !( flag & kInner ) || lat.nrows == 0 )
         return;
      pullHalos( ctx, *op.st, lat, xp, x.dtype );
      // cell interiors: 8 parity colours (every non-centre stencil offset has an
      // odd component, so points of one colour never couple)
      for ( int cz = 0; cz < 2; ++cz )
         for ( int cy = 0; cy < 2; ++cy )
            for ( int cx = 0; cx < 2; ++cx )
            {
               const int     g = cy + 2 * cz;
               const int64_t n = lat.rowGroup[g + 1] - lat.rowGroup[g];
               if ( n == 0 )
                  continue;
               dim3 grid( unsigned( n ), unsigned( lat.nc ) );
               kCellSOR< T ><<< grid, kRowBlock, 0, ctx->stream >>>( lat.rows.p + 2 * lat.rowGroup[g], lat.sc, lat.w,
                                                                   lat.mm, o.stencil.p, cx, relax, (const T*) bp,
                                                                   (T*) xp );
               HG_CHECK_HIP( hipGetLastError() );
            }
   } );
   ctx->endTiming();
}

void p1Restrict( Context* ctx, Function& f, int fineL, int flag )
{
   if ( fineL - 1 < f.minL || fineL > f.maxL )
      throw Error( -1, "restriction source level outside the function's level range" );
   Storage&    st = *f.st;
   Lattice&    lf = st.lattice( ctx, fineL );
   Lattice&    lc = st.lattice( ctx, fineL - 1 );
   P1Transfer& tr = transferFor( ctx, st, fineL );
   void*       fp = f.ptr( fineL );
   void*       cp = f.ptr( fineL - 1 );
   ctx->beginTiming();
   pullHalos( ctx, st, lf, fp, f.dtype );
   dispatchType( f.dtype, [&]( auto tag ) {
      using T = decltype( tag );
      if ( lc.nrows > 0 )
      {
         dim3 grid( unsigned( lc.nrows ), unsigned( lc.nc ) );
         kRestrictCell< T ><<< grid, kRowBlock, 0, ctx->stream >>>( lc.rows.p, lc.sc, lc.w, lc.mm, lf.sc, lf.w,
                                                                  (const T*) fp, (T*) cp );
         HG_CHECK_HIP( hipGetLastError() );
      }
      rowsApply( ctx, tr.restrictRows, lc.bndFlag.p, flag, fp, (T*) cp + lc.nc * lc.sc, f.dtype, 0 );
   } );
   ctx->endTiming();
}

void p1Prolongate( Context* ctx, Function& f, int coarseL, int flag, int update )
{
   if ( coarseL < f.minL || coarseL + 1 > f.maxL )
      throw Error( -1, "prolongation source level outside the function's level range" );
   Storage&    st = *f.st;
   Lattice&    lf = st.lattice( ctx, coarseL + 1 );
   Lattice&    lc = st.lattice( ctx, coarseL );
   P1Transfer& tr = transferFor( ctx, st, coarseL + 1 );
   void*       fp = f.ptr( coarseL + 1 );
   void*       cp = f.ptr( coarseL );
   ctx->beginTiming();
   pullHalos( ctx, st, lc, cp, f.dtype );
   dispatchType( f.dtype, [&]( auto tag ) {
      using T = decltype( tag );
      if ( lf.nrows > 0 )
      {
         dim3 grid( unsigned( lf.nrows ), unsigned( lf.nc ) );
         kProlongCell< T ><<< grid, kRowBlock, 0, ctx->stream >>>( lf.rows.p, lf.sc, lf.w, lf.mm, lc.sc, lc.w,
                                                                 (const T*) cp, (T*) fp, update );
         HG_CHECK_HIP( hipGetLastError() );
      }
      rowsApply( ctx, tr.prolongRows, lf.bndFlag.p, flag, cp, (T*) fp + lf.nc * lf.sc, f.dtype, update );
   } );
   ctx->endTiming();
}

} // namespace hg

using namespace hg;

extern "C" {

int hg_p1_laplace_create( hg_ctx* h, hg_handle s, int minLevel, int maxLevel, hg_handle* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !out || minLevel < 0 || maxLevel < minLevel || maxLevel > 12 )
         throw Error( HG_ERR_ARG, "invalid operator parameters" );
      Storage& st = ctx->storage( s );
      if ( !st.topo.blend.empty() )
         throw Error( HG_ERR_UNSUPPORT, "P1ConstantLaplaceOperator ignores blending; use an affine storage" );
      auto op  = std::make_unique< P1Operator >();
      op->st   = &st;
      op->minL = minLevel;
      op->maxL = maxLevel;
      for ( int L = minLevel; L <= maxLevel; ++L )
         buildP1Level( ctx, st, L, op->lv[L] );
      *out                  = ctx->nextHandle++;
      ctx->operators[*out] = std::move( op );
   } );
}

int hg_operator_destroy( hg_ctx* h, hg_handle op )
{
   return hg_guard( h, [&]( Context* ctx ) {
      HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
      if ( ctx->operators.erase( op ) == 0 )
         throw Error( HG_ERR_HANDLE, "unknown operator handle" );
   } );
}

int hg_p1_apply( hg_ctx* h, hg_handle op, hg_handle src, hg_handle dst, int level, int flag, int update )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( update != HG_REPLACE && update != HG_ADD )
         throw Error( HG_ERR_ARG, "invalid update type" );
      p1Apply( ctx, ctx->op< P1Operator >( op ), ctx->function( src ), ctx->function( dst ), level, flag, update );
   } );
}

int hg_p1_smooth_sor( hg_ctx* h, hg_handle op, hg_handle x, hg_handle b, double relax, int level, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      p1SOR( ctx, ctx->op< P1Operator >( op ), ctx->function( x ), ctx->function( b ), relax, level, flag );
   } );
}

int hg_p1_cell_stencil( hg_ctx* h, hg_handle op, int level, int cell, double* out15 )
{
   return hg_guard( h, [&]( Context* ctx ) {
      P1Operator& o = ctx->op< P1Operator >( op );
      if ( !out15 || cell < 0 || cell >= o.st->topo.nc )
         throw Error( HG_ERR_ARG, "invalid cell" );
      P1OpLevel& l = levelOf( ctx, o, level );
      for ( int s = 0; s < 15; ++s )
         out15[s] = l.hStencil[15 * size_t( cell ) + s];
   } );
}

int hg_p1_restrict( hg_ctx* h, hg_handle fn, int sourceLevel, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) { p1Restrict( ctx, ctx->function( fn ), sourceLevel, flag ); } );
}

int hg_p1_prolongate( hg_ctx* h, hg_handle fn, int sourceLevel, int flag, int update )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( update != HG_REPLACE && update != HG_ADD )
         throw Error( HG_ERR_ARG, "invalid update type" );
      p1Prolongate( ctx, ctx->function( fn ), sourceLevel, flag, update );
   } );
}

} // extern "C"
