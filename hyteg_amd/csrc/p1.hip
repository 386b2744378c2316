// P1 constant-stencil operator (apply, multicolour SOR), P1 linear restriction
// and prolongation: host set-up, HIP kernels and C-ABI entry points.
//
// Work split per level (HBM-bound stencil work, no MFMA):
//  * cell interiors: one structured 15-point stencil per macro-cell, swept row
//    by row ((y,z) rows, x contiguous) -- apply_3D_macrocell_vertexdof_to_vertexdof
//    (P1generatedKernels/apply_3D_macrocell_vertexdof_to_vertexdof_replace.cpp:34-79);
//  * DoFs owned by macro-vertices/edges/faces: the sum over incident cells of
//    each cell's share of the stencil (one partial stencil per boundary class
//    of the cell), read from the cell lattices' halo copies -- the macro-face/
//    edge/vertex stencils of P1Operator (P1ConstantOperator.cpp:239-358,
//    P1Operator.hpp:211-245, 1044-1108) evaluated cell by cell.
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <thread>

#include "abi_util.hpp"
#include "halo.hpp"
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


// Representative lattice points of the 14 boundary classes of a cell (local
// faces 0..3, edges 4..9, vertices 10..13) in a virtual lattice of 8 micro-edges
// per macro-edge: which micro-cells of a cell contain a boundary point depends
// on the point's class only.
static constexpr int kVirt            = 8;
static const int     kClassRep[14][3] = { { 2, 2, 0 }, { 2, 0, 2 }, { 0, 2, 2 }, { 2, 2, 4 }, { 4, 0, 0 },
                                          { 0, 4, 0 }, { 4, 4, 0 }, { 0, 0, 4 }, { 4, 0, 4 }, { 0, 4, 4 },
                                          { 0, 0, 0 }, { 8, 0, 0 }, { 0, 8, 0 }, { 0, 0, 8 } };

static int hostPointClass( int64_t mm, int64_t x, int64_t y, int64_t z )
{
   if ( x < 0 || y < 0 || z < 0 || x + y + z > mm )
      return -2;
   const int64_t t[4] = { mm - x - y - z, x, y, z };
   int           zmask = 0, nz = 0;
   for ( int k = 0; k < 4; ++k )
      if ( t[k] == 0 )
      {
         zmask |= 1 << k;
         ++nz;
      }
   if ( nz == 0 )
      return -1;
   const int keep = ( ~zmask ) & 15;
   if ( nz == 1 )
      for ( int k = 0; k < 4; ++k )
         if ( zmask == ( 1 << ( 3 - k ) ) )
            return k;
   if ( nz == 2 )
   {
      const int pairs[6] = { 3, 5, 6, 9, 10, 12 };
      for ( int k = 0; k < 6; ++k )
         if ( keep == pairs[k] )
            return 4 + k;
   }
   for ( int k = 0; k < 4; ++k )
      if ( keep == ( 1 << k ) )
         return 10 + k;
   return -1;
}

// Sliced ELL rows: slot i (a row, or -1 for padding) lies in slice i / 64 at
// lane i % 64; the slice's entries are column-major (entry j of the lane at
// off[slice] + 64 j + lane) and its width is that of its longest row rounded up
// to kSellUnroll, so the 64 lanes of a wavefront read each entry column with one
// coalesced load and run the same number of iterations.  Padding entries point
// at the row's own slot with weight 0.
constexpr int kSellRows   = 64;
constexpr int kSellUnroll = 8;

struct SellRows
{
   DevBuf< int64_t > off; // nslot / 64 + 1 entry offsets
   DevBuf< int32_t > row; // owned boundary DoF of each slot, -1: padding
   DevBuf< int32_t > pos; // storage positions
   DevBuf< double >  val;
   int64_t           nslot = 0;
};

// slots: the row of each slot (a multiple of 64 of them after padding);
// withDiag puts the row's own slot with weight 1 / inv[r] first
static void buildSell( Context* ctx, SellRows& out, std::vector< int32_t > slots,
                       const std::vector< std::vector< std::pair< int64_t, double > > >& rows,
                       const std::vector< double >& inv, int64_t ownerBase, bool withDiag )
{
   slots.resize( ( slots.size() + kSellRows - 1 ) / kSellRows * kSellRows, -1 );
   const size_t           nsl = slots.size() / kSellRows;
   std::vector< int64_t > off( nsl + 1, 0 );
   for ( size_t s = 0; s < nsl; ++s )
   {
      size_t w = 1;
      for ( int l = 0; l < kSellRows; ++l )
      {
         const int32_t r = slots[s * kSellRows + l];
         if ( r >= 0 )
            w = std::max( w, rows[size_t( r )].size() + ( withDiag ? 1 : 0 ) );
      }
      w          = ( w + kSellUnroll - 1 ) / kSellUnroll * kSellUnroll;
      off[s + 1] = off[s] + int64_t( w ) * kSellRows;
   }
   std::vector< int32_t > pos( size_t( off[nsl] ), 0 );
   std::vector< double >  val( size_t( off[nsl] ), 0.0 );
   for ( size_t s = 0; s < nsl; ++s )
   {
      const int64_t w = ( off[s + 1] - off[s] ) / kSellRows;
      for ( int l = 0; l < kSellRows; ++l )
      {
         const int32_t r = slots[s * kSellRows + l];
         if ( r < 0 )
            continue;
         const int64_t own = ownerBase + r;
         int64_t       j   = 0;
         auto          put = [&]( int64_t p, double v ) {
            pos[size_t( off[s] + j * kSellRows + l )] = int32_t( p );
            val[size_t( off[s] + j * kSellRows + l )] = v;
            ++j;
         };
         if ( withDiag )
            put( own, 1.0 / inv[size_t( r )] );
         for ( const auto& e : rows[size_t( r )] )
            put( e.first, e.second );
         while ( j < w )
            put( own, 0.0 );
      }
   }
   out.nslot = int64_t( slots.size() );
   out.off.upload( ctx, off );
   out.row.upload( ctx, slots );
   out.pos.upload( ctx, pos );
   out.val.upload( ctx, val );
}

struct P1OpLevel
{
   std::vector< double >  hStencil; // nc*15, cell interiors
   std::vector< double >  hCls;     // nc*14*15: the cell's share of the stencil at its boundary classes
   DevBuf< double >       stencil, clsStencil;
   DevBuf< double >       invDiag;    // per owned vertex/edge/face DoF
   std::vector< int64_t > colorStart; // owned boundary DoFs grouped by colour
   DevBuf< int32_t >      colorRows;
   // assembled rows of the owned boundary DoFs (built when they fit, see
   // kCsrRowBudget), as sliced ELL: storage positions of the neighbours -- the
   // owner slot of boundary neighbours, the cell-lattice slot of cell-interior
   // ones -- and weights.  applyRows holds every row in DoF order with its
   // diagonal first; sorRows holds the off-diagonal part in colour order, each
   // colour starting on a slice (sorColorStart: slot offsets).
   bool                   csr = false;
   SellRows               applyRows, sorRows;
   std::vector< int64_t > sorColorStart;
   // inverse diagonal of every storage position (cell lattices, halo copies
   // included, then the owned DoFs), built on first use
   DevBuf< double > invDiagFull;
};

// boundary rows are assembled up to this many owned boundary DoFs (two copies
// of about 40 entries of 12 bytes each: below 4.1 GB) when every storage
// position fits 31 bits; beyond it the kernels sum the per-cell stencil shares
// in place
static constexpr int64_t kCsrRowBudget = int64_t( 1 ) << 22;

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
   o.hCls.assign( 14 * 15 * size_t( t.nc ), 0.0 );
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
      // boundary classes: only the micro-cells inside this cell (the per-cell
      // parts of the macro-face/edge/vertex stencils, P1ConstantOperator.cpp:680-730)
      for ( int cls = 0; cls < 14; ++cls )
      {
         const int* p = kClassRep[cls];
         for ( int ty = 0; ty < 6; ++ty )
         {
            const int wt = kVirt - MicroCells::rowShift[ty];
            for ( int i = 0; i < 4; ++i )
            {
               const int* vi = MicroCells::vof[ty][i];
               const int  ax = p[0] - vi[0], ay = p[1] - vi[1], az = p[2] - vi[2];
               if ( ax < 0 || ay < 0 || az < 0 || ax + ay + az > wt - 1 )
                  continue;
               for ( int j = 0; j < 4; ++j )
               {
                  const int* vj = MicroCells::vof[ty][j];
                  o.hCls[( 14 * size_t( c ) + cls ) * 15 + slot15( vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2] )] +=
                      K[ty][i][j];
               }
            }
         }
      }
   }
   o.stencil.upload( ctx, o.hStencil );
   o.clsStencil.upload( ctx, o.hCls );

   // diagonal and coupling graph of the owned boundary DoFs.  Phase 1 walks the
   // boundary points of contiguous cell ranges in parallel and records every
   // contribution in a bucket of its row; phase 2 applies the buckets in
   // parallel, each in the order of the cell ranges, so every row receives its
   // additions in the order of a serial walk over the cells (the assembled
   // values do not depend on the thread count).
   std::vector< double >                 diag( lat.nb, 0.0 );
   std::vector< std::vector< int32_t > > adj( lat.nb );
   o.csr = lat.nb > 0 && lat.nb <= kCsrRowBudget && lat.nc * lat.sc + lat.nb <= int64_t( INT32_MAX );
   std::vector< std::vector< std::pair< int64_t, double > > > row( o.csr ? lat.nb : 0 );
   const int64_t                                             ownerBase = lat.nc * lat.sc;
   struct Contribution
   {
      int64_t pos;  // storage position of the neighbour, -1: the diagonal
      double  w;
      int32_t r;    // row (owned boundary DoF)
      int32_t adjq; // boundary neighbour for the colouring graph, -1: none
   };
   const int nthreads = int( std::max( 1u, std::min( 16u, std::thread::hardware_concurrency() ) ) );
   const int nbucket  = nthreads;
   std::vector< std::vector< std::vector< Contribution > > > rec{
       size_t( nthreads ), std::vector< std::vector< Contribution > >( static_cast< size_t >( nbucket ) ) };
   std::vector< char > outside( size_t( nthreads ), 0 );
   auto bucketOf = [&]( int64_t r ) { return int( ( r * nbucket ) / std::max< int64_t >( lat.nb, 1 ) ); };
   auto phase1   = [&]( int th ) {
      const int c0 = int( ( int64_t( t.nc ) * th ) / nthreads ), c1 = int( ( int64_t( t.nc ) * ( th + 1 ) ) / nthreads );
      auto&     out = rec[size_t( th )];
      for ( int c = c0; c < c1; ++c )
         for ( int64_t z = 0; z <= mm; ++z )
            for ( int64_t y = 0; y + z <= mm; ++y )
               for ( int64_t x = 0; x + y + z <= mm; ++x )
               {
                  const int cls = hostPointClass( mm, x, y, z );
                  if ( cls < 0 )
                     continue;
                  const int64_t r  = lat.num.locate( t, c, x, y, z );
                  auto&         bk = out[size_t( bucketOf( r ) )];
                  const double* W  = &o.hCls[( 14 * size_t( c ) + cls ) * 15];
                  bk.push_back( { -1, W[0], int32_t( r ), -1 } );
                  for ( int s = 1; s < 15; ++s )
                  {
                     const int64_t qx = x + Stencil15::off[s][0], qy = y + Stencil15::off[s][1],
                                   qz = z + Stencil15::off[s][2];
                     const int qc = hostPointClass( mm, qx, qy, qz );
                     if ( qc == -2 )
                     {
                        if ( W[s] != 0.0 )
                           outside[size_t( th )] = 1;
                        continue;
                     }
                     if ( W[s] == 0.0 )
                        continue;
                     const int64_t qb  = qc < 0 ? -1 : lat.num.locate( t, c, qx, qy, qz );
                     const int64_t pos = qc < 0 ? int64_t( c ) * lat.sc + tetIndex( lat.w, qx, qy, qz ) : ownerBase + qb;
                     bk.push_back( { pos, W[s], int32_t( r ), int32_t( qb ) } );
                  }
               }
   };
   auto phase2 = [&]( int b ) {
      for ( int th = 0; th < nthreads; ++th )
         for ( const Contribution& e : rec[size_t( th )][size_t( b )] )
         {
            if ( e.pos < 0 )
            {
               diag[size_t( e.r )] += e.w;
               continue;
            }
            if ( o.csr )
            {
               auto& rr = row[size_t( e.r )];
               auto  it = std::find_if( rr.begin(), rr.end(), [&]( const auto& v ) { return v.first == e.pos; } );
               if ( it == rr.end() )
                  rr.emplace_back( e.pos, e.w );
               else
                  it->second += e.w;
            }
            if ( e.adjq >= 0 )
               adj[size_t( e.r )].push_back( e.adjq );
         }
   };
   {
      std::vector< std::thread > pool;
      for ( int th = 1; th < nthreads; ++th )
         pool.emplace_back( phase1, th );
      phase1( 0 );
      for ( auto& p : pool )
         p.join();
   }
   for ( char bad : outside )
      if ( bad )
         throw Error( -4, "boundary stencil reaches outside its cell" );
   {
      std::vector< std::thread > pool;
      for ( int b = 1; b < nbucket; ++b )
         pool.emplace_back( phase2, b );
      phase2( 0 );
      for ( auto& p : pool )
         p.join();
   }
   rec.clear();
   for ( auto& rr : row )
      std::sort( rr.begin(), rr.end() );
   std::vector< double > inv( lat.nb );
   for ( int64_t r = 0; r < lat.nb; ++r )
      inv[r] = 1.0 / diag[r];
   o.invDiag.upload( ctx, inv );
   // greedy colouring in canonical order (vertices, edges, faces)
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
   o.colorStart.assign( ncolor + 1, 0 );
   std::vector< std::vector< int32_t > > byColor( ncolor );
   for ( int64_t r = 0; r < lat.nb; ++r )
      byColor[color[r]].push_back( int32_t( r ) );
   std::vector< int32_t > rows;
   for ( int cc = 0; cc < ncolor; ++cc )
   {
      o.colorStart[cc] = int64_t( rows.size() );
      rows.insert( rows.end(), byColor[cc].begin(), byColor[cc].end() );
   }
   o.colorStart[ncolor] = int64_t( rows.size() );
   o.colorRows.upload( ctx, rows );
   if ( o.csr )
   {
      std::vector< int32_t > natural( size_t( lat.nb ) ), sorOrder;
      for ( int64_t r = 0; r < lat.nb; ++r )
         natural[size_t( r )] = int32_t( r );
      o.sorColorStart.assign( size_t( ncolor ) + 1, 0 );
      for ( int cc = 0; cc < ncolor; ++cc )
      {
         o.sorColorStart[size_t( cc )] = int64_t( sorOrder.size() );
         sorOrder.insert( sorOrder.end(), byColor[cc].begin(), byColor[cc].end() );
         sorOrder.resize( ( sorOrder.size() + kSellRows - 1 ) / kSellRows * kSellRows, -1 );
      }
      o.sorColorStart[size_t( ncolor )] = int64_t( sorOrder.size() );
      buildSell( ctx, o.applyRows, natural, row, inv, ownerBase, true );
      buildSell( ctx, o.sorRows, sorOrder, row, inv, ownerBase, false );
   }
}

// ------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------
// Block b of a 1D launch of 8*ceil(n/8) blocks -> logical block such that each
// XCD (blocks are dispatched round-robin over the 8 XCDs) sweeps one contiguous
// range of rows and keeps the neighbour rows it shares in its own L2.
// Cell-interior kernels (apply, SOR): a block of 256 threads holds 4 row
// workers of 64 lanes, each sweeping one interior row (y, z) of a cell with x
// (and, for interleaved functions, the vector index k) across the lanes; short
// rows (small levels) are packed 2^rpbLog per worker.  The 15-point stencil
// reaches 7 rows -- (y, z) and its neighbours (y, z+1), (y+1, z-1), (y+1, z),
// (y-1, z), (y-1, z+1), (y, z-1), 2 or 3 consecutive points each -- whose
// starts are computed once per row in 32-bit cell-local offsets.
// Functions may interleave 2^lv vectors per lattice point (value k of point p
// at (p << lv) + k).
constexpr int kCellWorkers = 4;

struct StencilRows
{
   int32_t r00, r01, r1m, r10, rm0, rm1, r0m;
   __device__ __forceinline__ StencilRows( int32_t w, int32_t y, int32_t z )
   : r00( tetRowStart32( w, y, z ) )
   , r01( tetRowStart32( w, y, z + 1 ) )
   , r1m( tetRowStart32( w, y + 1, z - 1 ) )
   , r10( tetRowStart32( w, y + 1, z ) )
   , rm0( tetRowStart32( w, y - 1, z ) )
   , rm1( tetRowStart32( w, y - 1, z + 1 ) )
   , r0m( tetRowStart32( w, y, z - 1 ) )
   {}
   // cell-local offset of stencil entry s (the order of cOff15) at x
   __device__ __forceinline__ int32_t at( int s, int32_t x ) const
   {
      switch ( s )
      {
      case 0: return r00 + x;
      case 1: return r00 + x - 1;
      case 2: return r01 + x - 1;
      case 3: return r1m + x - 1;
      case 4: return r10 + x - 1;
      case 5: return rm0 + x;
      case 6: return rm1 + x;
      case 7: return r0m + x;
      case 8: return r01 + x;
      case 9: return r1m + x;
      case 10: return r10 + x;
      case 11: return rm0 + x + 1;
      case 12: return rm1 + x + 1;
      case 13: return r0m + x + 1;
      default: return r00 + x + 1;
      }
   }
};

// the row worker of this thread: logical row index (cell-major over the rows
// list) and lane inside the worker
struct RowWorker
{
   int64_t row;
   int     lane, lanes;
};
__device__ __forceinline__ RowWorker rowWorker( int64_t nrc, int rpbLog )
{
   const int64_t per = int64_t( kCellWorkers ) << rpbLog;
   RowWorker     rw;
   rw.row   = xcdBlock( ( nrc + per - 1 ) / per ) * per + ( threadIdx.x >> ( 6 - rpbLog ) );
   rw.lanes = 64 >> rpbLog;
   rw.lane  = int( threadIdx.x ) & ( rw.lanes - 1 );
   return rw;
}

// 15-point stencil on cell interiors (apply_3D_macrocell_vertexdof_to_vertexdof,
// P1generatedKernels/apply_3D_macrocell_vertexdof_to_vertexdof_replace.cpp:34-79):
// dst (=|+=) alpha sum_s st_s src[p + o_s], summed in the order of the offsets
template < class T >
__global__ __launch_bounds__( 64 * kCellWorkers ) void kCellApply( const int32_t* __restrict__ rows, int64_t nrows,
                                                                   int64_t nc, int64_t sc, int64_t w, int64_t mm,
                                                                   const double* __restrict__ stencil,
                                                                   const T* __restrict__ src, T* __restrict__ dst,
                                                                   int add, int lv, int rpbLog, double alpha )
{
   using A              = typename Acc< T >::type;
   const int64_t   nrc  = nrows * nc;
   const RowWorker rw   = rowWorker( nrc, rpbLog );
   if ( rw.row >= nrc )
      return;
   const int64_t     c = rw.row / nrows;
   const int64_t     r = rw.row - c * nrows;
   const int32_t     y = rows[2 * r], z = rows[2 * r + 1];
   const StencilRows sr( int32_t( w ), y, z );
   const T*          s = src + ( ( c * sc ) << lv );
   T*                d = dst + ( ( c * sc ) << lv );
   A                 st[15];
#pragma unroll
   for ( int q = 0; q < 15; ++q )
      st[q] = A( stencil[15 * c + q] );
   const int32_t msk = ( 1 << lv ) - 1, top = int32_t( mm ) - y - z;
   for ( int32_t t = rw.lane;; t += rw.lanes )
   {
      const int32_t x = 1 + ( t >> lv ), k = t & msk;
      if ( x >= top )
         break;
      // all 15 loads in flight before the sum (same order of the sum)
      T val[15];
#pragma unroll
      for ( int q = 0; q < 15; ++q )
         val[q] = s[( sr.at( q, x ) << lv ) + k];
      __builtin_amdgcn_sched_barrier( 0 ); // keeps the scheduler from moving the loads down to their uses
      A v = A( 0 );
#pragma unroll
      for ( int q = 0; q < 15; ++q )
         v += st[q] * toAcc( val[q] );
      v               = A( alpha ) * v;
      const int32_t o = ( sr.at( 0, x ) << lv ) + k;
      if ( add )
         v += toAcc( d[o] );
      d[o] = fromAcc< T >( v );
   }
}

// The same apply over bricks of 64 x 8 x 8 cell-interior points (x, y rows, z
// slices) for single-vector functions on lattices of at least 64 micro-edges:
// a workgroup (8 wavefronts, one per z slice) stages the brick with its
// one-point halo (66 x 10 x 10 values) in LDS with coalesced row loads, then
// every lane sums its eight points' 15 entries from LDS, in the order of
// kCellApply (same results bit for bit).  (Bricks of 64 x 4 x 4 points took
// 0.63 ms at cube level 9, 64 x 8 x 8 points 0.57 ms.)
// The row kernel issues 15 global loads per point and is bound by their
// issue, not by HBM traffic.
constexpr int kBrX = 64, kBrY = 8, kBrZ = 8;
constexpr int kBrLX = kBrX + 2, kBrLY = kBrY + 2, kBrL = kBrLX * kBrLY * ( kBrZ + 2 );
constexpr int kBrThreads = kBrX * kBrZ;
// LDS offsets of the 15 stencil entries (the order of StencilRows::at)
__device__ __forceinline__ constexpr int brickOffset( int q )
{
   constexpr int dx[15] = { 0, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1 };
   constexpr int dy[15] = { 0, 0, 0, 1, 1, -1, -1, 0, 0, 1, 1, -1, -1, 0, 0 };
   constexpr int dz[15] = { 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 0 };
   return dx[q] + kBrLX * ( dy[q] + kBrLY * dz[q] );
}

template < class T >
__global__ __launch_bounds__( kBrThreads ) void kCellApplyBricks( const uint64_t* __restrict__ bricks, int64_t nbricks,
                                                                  int64_t sc, int32_t w, int32_t mm,
                                                                  const double* __restrict__ stencil,
                                                                  const T* __restrict__ src, T* __restrict__ dst,
                                                                  int add, double alpha )
{
   using A = typename Acc< T >::type;
   __shared__ A  lds[kBrL];
   const int64_t bi = xcdBlock( nbricks );
   if ( bi >= nbricks )
      return;
   using ConstU64   = const __attribute__( ( address_space( 4 ) ) ) uint64_t;
   const uint64_t e = ( (ConstU64*) bricks )[bi];
   const int64_t  c = int64_t( e & 0xFFFFFFu );
   const int32_t  x0 = kBrX * int32_t( ( e >> 24 ) & 255u ), y0 = kBrY * int32_t( ( e >> 32 ) & 0xFFFFu ),
                 z0 = kBrZ * int32_t( e >> 48 );
   const T* s = src + c * sc;
   T*       d = dst + c * sc;
   // the brick and its halo: all loads issued before the first LDS write
   constexpr int kPer = ( kBrL + kBrThreads - 1 ) / kBrThreads;
   T             v[kPer];
#pragma unroll
   for ( int j = 0; j < kPer; ++j )
   {
      const int     i  = int( threadIdx.x ) + j * kBrThreads;
      const int     lx = i % kBrLX, ly = ( i / kBrLX ) % kBrLY, lz = i / ( kBrLX * kBrLY );
      const int32_t X = x0 - 1 + lx, Y = y0 - 1 + ly, Z = z0 - 1 + lz;
      v[j] = ( i < kBrL && X >= 0 && Y >= 0 && Z >= 0 && X + Y + Z <= mm ) ? s[tetRowStart32( w, Y, Z ) + X] : T( 0 );
   }
   __builtin_amdgcn_sched_barrier( 0 );
#pragma unroll
   for ( int j = 0; j < kPer; ++j )
   {
      const int i = int( threadIdx.x ) + j * kBrThreads;
      if ( i < kBrL )
         lds[i] = A( toAcc( v[j] ) );
   }
   A st[15];
#pragma unroll
   for ( int q = 0; q < 15; ++q )
      st[q] = A( stencil[15 * c + q] );
   __syncthreads();
   const int     lane = int( threadIdx.x ) & 63, wv = int( threadIdx.x ) >> 6;
   const int32_t X = x0 + lane, Z = z0 + wv;
#pragma unroll
   for ( int yy = 0; yy < kBrY; ++yy )
   {
      const int32_t Y = y0 + yy;
      if ( X < 1 || Y < 1 || Z < 1 || X + Y + Z > mm - 1 )
         continue;
      const int li = ( lane + 1 ) + kBrLX * ( ( yy + 1 ) + kBrLY * ( wv + 1 ) );
      A         r  = A( 0 );
#pragma unroll
      for ( int q = 0; q < 15; ++q )
         r += st[q] * lds[li + brickOffset( q )];
      r               = A( alpha ) * r;
      const int32_t o = tetRowStart32( w, Y, Z ) + X;
      if ( add )
         r += toAcc( d[o] );
      d[o] = fromAcc< T >( r );
   }
}

// the lattice's brick list (entry c | bx << 24 | by << 32 | bz << 48), built on
// first use; null where the row kernel runs
static const uint64_t* p1BrickList( Context* ctx, Lattice& l )
{
   if ( l.mm < kBrX || l.mm > 1023 || l.nc >= ( int64_t( 1 ) << 24 ) )
      return nullptr;
   if ( !l.p1Bricks.p )
   {
      std::vector< uint64_t > e;
      const int64_t           mm = l.mm;
      for ( int64_t c = 0; c < l.nc; ++c )
         for ( int64_t bz = 0; kBrZ * bz <= mm; ++bz )
            for ( int64_t by = 0; kBrY * by <= mm; ++by )
               for ( int64_t bx = 0; kBrX * bx <= mm; ++bx )
                  if ( std::max< int64_t >( kBrX * bx, 1 ) + std::max< int64_t >( kBrY * by, 1 ) +
                           std::max< int64_t >( kBrZ * bz, 1 ) <=
                       mm - 1 )
                     e.push_back( uint64_t( c ) | ( uint64_t( bx ) << 24 ) | ( uint64_t( by ) << 32 ) |
                                  ( uint64_t( bz ) << 48 ) );
      if ( e.empty() )
         return nullptr;
      l.p1Bricks.upload( ctx, e );
   }
   return l.p1Bricks.p;
}

// SOR on the cell-interior points of the two parity colours (cx, cy, cz) and
// (1 - cx, cy, cz) of one row group, in place: colour cxFirst first, then the
// other.  The only stencil neighbours of a point that lie in rows of the same
// group are x - 1 and x + 1 of its own row, so the second colour depends on the
// first through the row alone; a row's lanes are in one wavefront, and a
// workgroup-scope fence between the two passes makes the first pass's stores
// visible to the second (the result is that of two colour sweeps in turn).
template < class T >
__global__ __launch_bounds__( 64 * kCellWorkers ) void kCellSOR( const int32_t* __restrict__ rows, int64_t nrows,
                                                                 int64_t nc, int64_t sc, int64_t w, int64_t mm,
                                                                 const double* __restrict__ stencil, int cxFirst,
                                                                 double relax, const T* __restrict__ b,
                                                                 T* __restrict__ x, int lv, int rpbLog )
{
   using A              = typename Acc< T >::type;
   const int64_t   nrc  = nrows * nc;
   const RowWorker rw   = rowWorker( nrc, rpbLog );
   if ( rw.row >= nrc )
      return;
   const int64_t     c = rw.row / nrows;
   const int64_t     r = rw.row - c * nrows;
   const int32_t     y = rows[2 * r], z = rows[2 * r + 1];
   const StencilRows sr( int32_t( w ), y, z );
   const T*          bc = b + ( ( c * sc ) << lv );
   T*                xc = x + ( ( c * sc ) << lv );
   A                 st[15];
#pragma unroll
   for ( int q = 0; q < 15; ++q )
      st[q] = A( stencil[15 * c + q] );
   const A       invc = A( 1 ) / st[0];
   const int32_t msk = ( 1 << lv ) - 1, top = int32_t( mm ) - y - z;
   for ( int ph = 0; ph < 2; ++ph )
   {
      const int32_t x0 = ( ph ? 1 - cxFirst : cxFirst ) ? 1 : 2;
      for ( int32_t t = rw.lane;; t += rw.lanes )
      {
         const int32_t xx = x0 + 2 * ( t >> lv ), k = t & msk;
         if ( xx >= top )
            break;
         const int32_t o = ( sr.at( 0, xx ) << lv ) + k;
         // all loads in flight before the sum (same order of the sum)
         T val[15];
#pragma unroll
         for ( int q = 0; q < 15; ++q )
            val[q] = xc[( sr.at( q, xx ) << lv ) + k];
         A v = toAcc( bc[o] );
         __builtin_amdgcn_sched_barrier( 0 ); // keeps the scheduler from moving the loads down to their uses
#pragma unroll
         for ( int q = 1; q < 15; ++q )
            v -= st[q] * toAcc( val[q] );
         const A xo = toAcc( val[0] );
         xc[o]      = fromAcc< T >( ( A( 1 ) - A( relax ) ) * xo + A( relax ) * invc * v );
      }
      if ( ph == 0 )
         __threadfence_block();
   }
}

// partial stencil sum of this DoF's share in cell c; skip = 1 leaves out the centre
template < class T >
__device__ __forceinline__ typename Acc< T >::type bndStencilSum( const BndGeom& g, const double* __restrict__ cls,
                                                                  const T* __restrict__ v, int64_t c, int64_t x,
                                                                  int64_t y, int64_t z, int k, int skip, int lv,
                                                                  int64_t vec )
{
   using A           = typename Acc< T >::type;
   const double* W   = cls + ( 14 * c + k ) * 15;
   A             s   = A( 0 );
   const int64_t base = c * g.sc;
   for ( int q = skip; q < 15; ++q )
   {
      const double wq = W[q];
      if ( wq == 0.0 )
         continue;
      const int64_t nx = x + cOff15[q][0], ny = y + cOff15[q][1], nz = z + cOff15[q][2];
      if ( nx < 0 || ny < 0 || nz < 0 || nx + ny + nz > g.mm )
         continue;
      s += A( wq ) * toAcc( v[( ( base + tetIndex( g.w, nx, ny, nz ) ) << lv ) + vec] );
   }
   return s;
}

// apply on DoFs owned by macro-vertices/edges/faces: sum of every incident
// cell's share, read from the cell lattices (halos refreshed before)
template < class T >
__global__ void kBndApply( BndGeom g, int64_t nb, int64_t nc, const uint8_t* __restrict__ bndFlag, int flag,
                           const double* __restrict__ cls, const T* __restrict__ src, T* __restrict__ dst, int add,
                           int lv, double alpha )
{
   using A           = typename Acc< T >::type;
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t b   = tid >> lv, vec = tid & ( ( int64_t( 1 ) << lv ) - 1 );
   if ( b >= nb || !( bndFlag[b] & flag ) )
      return;
   A s = A( 0 );
   forEachIncident( g, b, [&]( int64_t c, int64_t x, int64_t y, int64_t z, int k ) {
      s += bndStencilSum( g, cls, src, c, x, y, z, k, 0, lv, vec );
   } );
   s               = A( alpha ) * s;
   const int64_t o = ( ( nc * g.sc + b ) << lv ) + vec;
   if ( add )
      s += toAcc( dst[o] );
   dst[o] = fromAcc< T >( s );
}

// SOR on the owned boundary DoFs of one colour; the new value is written to
// the owner and to every halo copy, so later colours read it
template < class T >
__global__ void kBndSOR( BndGeom g, int64_t n, const int32_t* __restrict__ list, int64_t nc,
                         const uint8_t* __restrict__ bndFlag, int flag, const double* __restrict__ cls,
                         const double* __restrict__ invDiag, double relax, const T* __restrict__ rhs,
                         T* __restrict__ x, int lv )
{
   using A           = typename Acc< T >::type;
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t i   = tid >> lv, vec = tid & ( ( int64_t( 1 ) << lv ) - 1 );
   if ( i >= n )
      return;
   const int64_t b = list[i];
   if ( !( bndFlag[b] & flag ) )
      return;
   const int64_t o = ( ( nc * g.sc + b ) << lv ) + vec;
   A             s = toAcc( rhs[o] );
   forEachIncident( g, b, [&]( int64_t c, int64_t xx, int64_t yy, int64_t zz, int k ) {
      s -= bndStencilSum( g, cls, x, c, xx, yy, zz, k, 1, lv, vec );
   } );
   const T nv = fromAcc< T >( ( A( 1 ) - A( relax ) ) * toAcc( x[o] ) + A( relax ) * A( invDiag[b] ) * s );
   x[o]       = nv;
   forEachIncident( g, b, [&]( int64_t c, int64_t xx, int64_t yy, int64_t zz, int ) {
      x[( ( c * g.sc + tetIndex( g.w, xx, yy, zz ) ) << lv ) + vec] = nv;
   } );
}

// apply / SOR on owned boundary DoFs from their assembled rows (sliced ELL):
// a thread takes one slot (and one vector of a multi-vector function), issues
// the position and weight loads of kSellUnroll entries, then their gathers,
// then sums them in entry order.  Logical blocks are spread over the XCDs in
// contiguous ranges, so neighbouring rows share the gathered lines in one L2.
template < class T >
__global__ void __launch_bounds__( kBlock ) kBndApplySell( int64_t nblk, int64_t nslot, const int32_t* __restrict__ rowOf,
                                                          const uint8_t* __restrict__ bndFlag, int flag,
                                                          const int64_t* __restrict__ off, const int32_t* __restrict__ pos,
                                                          const double* __restrict__ val, int64_t ownerBase,
                                                          const T* __restrict__ src, T* __restrict__ dst, int add, int lv,
                                                          double alpha )
{
   using A          = typename Acc< T >::type;
   const int64_t lb = xcdBlock( nblk );
   if ( lb >= nblk )
      return;
   const int64_t tid = lb * kBlock + threadIdx.x;
   const int64_t i = tid >> lv, vec = tid & ( ( int64_t( 1 ) << lv ) - 1 );
   if ( i >= nslot )
      return;
   const int32_t b = rowOf[i];
   if ( b < 0 || !( bndFlag[b] & flag ) )
      return;
   const int64_t sl = i / kSellRows, e0 = off[sl] + ( i % kSellRows ), w = ( off[sl + 1] - off[sl] ) / kSellRows;
   A             s  = A( 0 );
   for ( int64_t j = 0; j < w; j += kSellUnroll )
   {
      int32_t p[kSellUnroll];
      double  v[kSellUnroll];
      A       x[kSellUnroll];
#pragma unroll
      for ( int u = 0; u < kSellUnroll; ++u )
      {
         p[u] = pos[e0 + ( j + u ) * kSellRows];
         v[u] = val[e0 + ( j + u ) * kSellRows];
      }
#pragma unroll
      for ( int u = 0; u < kSellUnroll; ++u )
         x[u] = toAcc( src[( int64_t( p[u] ) << lv ) + vec] );
#pragma unroll
      for ( int u = 0; u < kSellUnroll; ++u )
         s += A( v[u] ) * x[u];
   }
   s               = A( alpha ) * s;
   const int64_t o = ( ( ownerBase + b ) << lv ) + vec;
   if ( add )
      s += toAcc( dst[o] );
   dst[o] = fromAcc< T >( s );
}

template < class T >
__global__ void __launch_bounds__( kBlock ) kBndSORSell( int64_t nblk, int64_t slot0, int64_t nslot,
                                                        const int32_t* __restrict__ rowOf,
                                                        const uint8_t* __restrict__ bndFlag, int flag,
                                                        const int64_t* __restrict__ off, const int32_t* __restrict__ pos,
                                                        const double* __restrict__ val, int64_t ownerBase,
                                                        const double* __restrict__ invDiag, double relax,
                                                        const T* __restrict__ rhs, T* __restrict__ x, int lv )
{
   using A          = typename Acc< T >::type;
   const int64_t lb = xcdBlock( nblk );
   if ( lb >= nblk )
      return;
   const int64_t tid = lb * kBlock + threadIdx.x;
   const int64_t li = tid >> lv, vec = tid & ( ( int64_t( 1 ) << lv ) - 1 );
   if ( li >= nslot )
      return;
   const int64_t i = slot0 + li;
   const int32_t b = rowOf[i];
   if ( b < 0 || !( bndFlag[b] & flag ) )
      return;
   const int64_t sl = i / kSellRows, e0 = off[sl] + ( i % kSellRows ), w = ( off[sl + 1] - off[sl] ) / kSellRows;
   const int64_t o  = ( ( ownerBase + b ) << lv ) + vec;
   A             s  = toAcc( rhs[o] );
   for ( int64_t j = 0; j < w; j += kSellUnroll )
   {
      int32_t p[kSellUnroll];
      double  v[kSellUnroll];
      A       xv[kSellUnroll];
#pragma unroll
      for ( int u = 0; u < kSellUnroll; ++u )
      {
         p[u] = pos[e0 + ( j + u ) * kSellRows];
         v[u] = val[e0 + ( j + u ) * kSellRows];
      }
#pragma unroll
      for ( int u = 0; u < kSellUnroll; ++u )
         xv[u] = toAcc( x[( int64_t( p[u] ) << lv ) + vec] );
#pragma unroll
      for ( int u = 0; u < kSellUnroll; ++u )
         s -= A( v[u] ) * xv[u];
   }
   x[o] = fromAcc< T >( ( A( 1 ) - A( relax ) ) * toAcc( x[o] ) + A( relax ) * A( invDiag[b] ) * s );
}

// restriction to coarse cell-interior points (full weighting, R = P^T)
template < class T >
__global__ void kRestrictCell( const int32_t* __restrict__ rows, int64_t nrows, int64_t nc, int64_t scc, int64_t wc,
                               int64_t mmc, int64_t scf, int64_t wf, const T* __restrict__ fine, T* __restrict__ coarse,
                               int lv )
{
   using A          = typename Acc< T >::type;
   const int64_t lb = xcdBlock( nrows * nc );
   if ( lb >= nrows * nc )
      return;
   const int64_t c = lb / nrows;
   const int64_t r = lb - c * nrows;
   const int64_t y = rows[2 * r];
   const int64_t z = rows[2 * r + 1];
   int64_t       rb[15];
#pragma unroll
   for ( int s = 0; s < 15; ++s )
      rb[s] = c * scf + tetRowStart( wf, 2 * y + cOff15[s][1], 2 * z + cOff15[s][2] ) + cOff15[s][0];
   const int64_t out = c * scc + tetRowStart( wc, y, z );
   const int64_t msk = ( int64_t( 1 ) << lv ) - 1;
   for ( int64_t t = threadIdx.x;; t += blockDim.x )
   {
      const int64_t x = 1 + ( t >> lv ), k = t & msk;
      if ( x + y + z >= mmc )
         break;
      A v = A( 0 );
#pragma unroll
      for ( int s = 1; s < 15; ++s )
         v += toAcc( fine[( ( rb[s] + 2 * x ) << lv ) + k] );
      v                              = toAcc( fine[( ( rb[0] + 2 * x ) << lv ) + k] ) + A( 0.5 ) * v;
      coarse[( ( out + x ) << lv ) + k] = fromAcc< T >( v );
   }
}

// restriction to coarse DoFs owned by vertices/edges/faces: every incident
// cell adds its fine points, each scaled by 1/(cells sharing the fine point's
// primitive), i.e. restrict_3D_macrocell_P1_pull_additive + additive communication
template < class T >
__global__ void kRestrictBnd( BndGeom gc, int64_t nbc, int64_t nc, const uint8_t* __restrict__ bndFlag, int flag,
                              const double* __restrict__ clsInvN, int64_t scf, int64_t wf, int64_t mmf,
                              const T* __restrict__ fine, T* __restrict__ coarse, int lv )
{
   using A           = typename Acc< T >::type;
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t b   = tid >> lv, vec = tid & ( ( int64_t( 1 ) << lv ) - 1 );
   if ( b >= nbc || !( bndFlag[b] & flag ) )
      return;
   A s = A( 0 );
   forEachIncident( gc, b, [&]( int64_t c, int64_t x, int64_t y, int64_t z, int ) {
      for ( int q = 0; q < 15; ++q )
      {
         const int64_t fx = 2 * x + cOff15[q][0], fy = 2 * y + cOff15[q][1], fz = 2 * z + cOff15[q][2];
         if ( fx < 0 || fy < 0 || fz < 0 || fx + fy + fz > mmf )
            continue;
         const int cl  = pointClass( mmf, fx, fy, fz );
         const A   wgt = ( q == 0 ? A( 1 ) : A( 0.5 ) ) * ( cl < 0 ? A( 1 ) : A( clsInvN[14 * c + cl] ) );
         s += wgt * toAcc( fine[( ( c * scf + tetIndex( wf, fx, fy, fz ) ) << lv ) + vec] );
      }
   } );
   coarse[( ( nc * gc.sc + b ) << lv ) + vec] = fromAcc< T >( s );
}

// linear interpolation to fine cell-interior points
template < class T >
__global__ void kProlongCell( const int32_t* __restrict__ rows, int64_t nrows, int64_t nc, int64_t scf, int64_t wf,
                              int64_t mmf, int64_t scc, int64_t wc, const T* __restrict__ coarse, T* __restrict__ fine,
                              int add, int lv )
{
   using A          = typename Acc< T >::type;
   const int64_t lb = xcdBlock( nrows * nc );
   if ( lb >= nrows * nc )
      return;
   const int64_t c   = lb / nrows;
   const int64_t r   = lb - c * nrows;
   const int64_t y   = rows[2 * r];
   const int64_t z   = rows[2 * r + 1];
   const int64_t out = c * scf + tetRowStart( wf, y, z );
   const int64_t msk = ( int64_t( 1 ) << lv ) - 1;
   for ( int64_t t = threadIdx.x;; t += blockDim.x )
   {
      const int64_t x = 1 + ( t >> lv ), k = t & msk;
      if ( x + y + z >= mmf )
         break;
      const int p = int( x & 1 ) + 2 * int( y & 1 ) + 4 * int( z & 1 );
      A         v;
      if ( p == 0 )
         v = toAcc( coarse[( ( c * scc + tetIndex( wc, x / 2, y / 2, z / 2 ) ) << lv ) + k] );
      else
      {
         const int dx = cMidDir[p][0], dy = cMidDir[p][1], dz = cMidDir[p][2];
         v = A( 0.5 ) *
             ( toAcc( coarse[( ( c * scc + tetIndex( wc, ( x - dx ) / 2, ( y - dy ) / 2, ( z - dz ) / 2 ) ) << lv ) + k] ) +
               toAcc( coarse[( ( c * scc + tetIndex( wc, ( x + dx ) / 2, ( y + dy ) / 2, ( z + dz ) / 2 ) ) << lv ) + k] ) );
      }
      const int64_t o = ( ( out + x ) << lv ) + k;
      if ( add )
         v += toAcc( fine[o] );
      fine[o] = fromAcc< T >( v );
   }
}

// linear interpolation to fine DoFs owned by vertices/edges/faces, read from
// the coarse lattice of the first incident cell (halos refreshed before)
template < class T >
__global__ void kProlongBnd( BndGeom gf, int64_t nbf, int64_t nc, const uint8_t* __restrict__ bndFlag, int flag,
                             int64_t scc, int64_t wc, const T* __restrict__ coarse, T* __restrict__ fine, int add,
                             int lv )
{
   using A           = typename Acc< T >::type;
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t b   = tid >> lv, vec = tid & ( ( int64_t( 1 ) << lv ) - 1 );
   if ( b >= nbf || !( bndFlag[b] & flag ) )
      return;
   bool done = false;
   A    v    = A( 0 );
   forEachIncident( gf, b, [&]( int64_t c, int64_t x, int64_t y, int64_t z, int ) {
      if ( done )
         return;
      done        = true;
      const int p = int( x & 1 ) + 2 * int( y & 1 ) + 4 * int( z & 1 );
      if ( p == 0 )
         v = toAcc( coarse[( ( c * scc + tetIndex( wc, x / 2, y / 2, z / 2 ) ) << lv ) + vec] );
      else
      {
         const int dx = cMidDir[p][0], dy = cMidDir[p][1], dz = cMidDir[p][2];
         v = A( 0.5 ) *
             ( toAcc( coarse[( ( c * scc + tetIndex( wc, ( x - dx ) / 2, ( y - dy ) / 2, ( z - dz ) / 2 ) ) << lv ) + vec] ) +
               toAcc( coarse[( ( c * scc + tetIndex( wc, ( x + dx ) / 2, ( y + dy ) / 2, ( z + dz ) / 2 ) ) << lv ) + vec] ) );
      }
   } );
   const int64_t o = ( ( nc * gf.sc + b ) << lv ) + vec;
   if ( add )
      v += toAcc( fine[o] );
   fine[o] = fromAcc< T >( v );
}

// inverse diagonal of every storage position: 1 / stencil centre on the cell
// lattices, the assembled boundary rows' inverse diagonal on the owned DoFs
// (halo positions are refreshed from the owners afterwards)
__global__ void kP1InvDiagFull( int64_t nc, int64_t sc, int64_t nb, const double* __restrict__ stencil,
                                const double* __restrict__ invBnd, double* __restrict__ out )
{
   const int64_t i = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( i < nc * sc )
      out[i] = 1.0 / stencil[15 * ( i / sc )];
   else if ( i < nc * sc + nb )
      out[i] = invBnd[i - nc * sc];
}

// ------------------------------------------------------------------------
// host drivers
// ------------------------------------------------------------------------
static constexpr int kRowBlock = 64; // the transfer kernels' row blocks

// log2 of the rows one 64-lane row worker of a cell kernel takes: enough lanes
// for `lanes` values per row, at most 64 rows
static int rowsPerBlockLog( int64_t lanes )
{
   int r = 0;
   while ( r < 6 && ( int64_t( 64 ) >> ( r + 1 ) ) >= lanes )
      ++r;
   return r;
}

static unsigned cellGrid( int64_t rowsTimesCells, int rpbLog )
{
   const int64_t per = int64_t( kCellWorkers ) << rpbLog;
   return xcdGrid( ( rowsTimesCells + per - 1 ) / per );
}

// the cell kernels index a cell's values (2^lv per lattice point) in 32 bits
static void checkCellIndexing( const Lattice& lat, int lv )
{
   if ( ( lat.sc << lv ) >= ( int64_t( 1 ) << 31 ) )
      throw Error( HG_ERR_UNSUPPORT, "a macro-cell lattice of this level holds more than 2^31 values" );
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

// dst = alpha A src (update Replace) or dst + alpha A src (update Add) on the
// DoFs of type flag
static const double* p1InverseDiagonal( Context* ctx, P1Operator& op, int L )
{
   P1OpLevel& o = levelOf( ctx, op, L );
   if ( o.invDiagFull.p )
      return o.invDiagFull.p;
   Lattice&      lat = op.st->lattice( ctx, L );
   const int64_t n   = lat.nc * lat.sc + lat.nb;
   o.invDiagFull.alloc( ctx, size_t( n ) );
   kP1InvDiagFull<<< gridFor( n ), kBlock, 0, ctx->stream >>>( lat.nc, lat.sc, lat.nb, o.stencil.p, o.invDiag.p,
                                                              o.invDiagFull.p );
   HG_CHECK_HIP( hipGetLastError() );
   pullHalos( ctx, *op.st, lat, o.invDiagFull.p, 0 );
   return o.invDiagFull.p;
}

void p1Apply( Context* ctx, P1Operator& op, Function& src, Function& dst, int L, int flag, int update,
              double alpha = 1.0 )
{
   if ( src.st != op.st || dst.st != op.st || src.dtype != dst.dtype || src.off != 0 || dst.off != 0 ||
        src.lv != dst.lv )
      throw Error( -1, "apply needs P1 src/dst on the operator's storage with one value type and vector count" );
   const int lv = src.lv;
   if ( &src == &dst )
      throw Error( -1, "apply needs distinct src and dst" );
   P1OpLevel& o   = levelOf( ctx, op, L );
   Lattice&   lat = op.st->lattice( ctx, L );
   void*      s   = src.ptr( L );
   void*      d   = dst.ptr( L );
   pullHalos( ctx, *op.st, lat, s, src.dtype, lv );
   dispatchType( src.dtype, [&]( auto tag ) {
      using T = decltype( tag );
      if ( lat.nb > 0 && o.csr )
      {
         const SellRows& sr   = o.applyRows;
         const int64_t   nblk = ( ( sr.nslot << lv ) + kBlock - 1 ) / kBlock;
         kBndApplySell< T ><<< xcdGrid( nblk ), kBlock, 0, ctx->stream >>>(
             nblk, sr.nslot, sr.row.p, lat.bndFlag.p, flag, sr.off.p, sr.pos.p, sr.val.p, lat.nc * lat.sc,
             (const T*) s, (T*) d, update, lv, alpha );
         HG_CHECK_HIP( hipGetLastError() );
      }
      else if ( lat.nb > 0 )
      {
         kBndApply< T ><<< gridFor( lat.nb << lv ), kBlock, 0, ctx->stream >>>( BndGeom::of( *op.st, lat ), lat.nb,
                                                                               lat.nc, lat.bndFlag.p, flag,
                                                                               o.clsStencil.p, (const T*) s, (T*) d,
                                                                               update, lv, alpha );
         HG_CHECK_HIP( hipGetLastError() );
      }
      if ( ( flag & kInner ) && lat.nrows > 0 )
      {
         checkCellIndexing( lat, lv );
         const uint64_t* bricks = lv == 0 ? p1BrickList( ctx, lat ) : nullptr;
         ctx->beginTiming();
         if ( bricks )
         {
            const int64_t nb = int64_t( lat.p1Bricks.n );
            kCellApplyBricks< T ><<< xcdGrid( nb ), kBrThreads, 0, ctx->stream >>>(
                bricks, nb, lat.sc, int32_t( lat.w ), int32_t( lat.mm ), o.stencil.p, (const T*) s, (T*) d, update,
                alpha );
         }
         else
         {
            const int rpb = rowsPerBlockLog( ( lat.mm - 3 ) << lv );
            kCellApply< T ><<< cellGrid( lat.nrows * lat.nc, rpb ), 64 * kCellWorkers, 0, ctx->stream >>>(
                lat.rowsNat.p, lat.nrows, lat.nc, lat.sc, lat.w, lat.mm, o.stencil.p, (const T*) s, (T*) d, update,
                lv, rpb, alpha );
         }
         HG_CHECK_HIP( hipGetLastError() );
         ctx->endTiming();
      }
   } );
}

// One SOR sweep in multicolour order.  Forward: the owned vertex/edge/face DoFs
// colour by colour (greedy colours of their coupling graph), then the 8 parity
// colours k = (x&1) + 2(y&1) + 4(z&1) of the cell interiors (every non-centre
// offset of the 15-point stencil has an odd component, so points of one parity
// never couple); backward: the same colours in reverse order
// (smooth_sor_backwards, P1Operator.hpp:331-336,381-396).
void p1SOR( Context* ctx, P1Operator& op, Function& x, Function& b, double relax, int L, int flag,
            bool backward = false )
{
   if ( x.st != op.st || b.st != op.st || x.dtype != b.dtype || &x == &b || x.off != 0 || b.off != 0 || x.lv != b.lv )
      throw Error( -1, "smooth_sor needs distinct x/b on the operator's storage with one value type and vector count" );
   const int lv = x.lv;
   P1OpLevel& o   = levelOf( ctx, op, L );
   Lattice&   lat = op.st->lattice( ctx, L );
   void*      xp  = x.ptr( L );
   const void* bp = b.ptr( L );
   ctx->beginTiming();
   pullHalos( ctx, *op.st, lat, xp, x.dtype, lv );
   dispatchType( x.dtype, [&]( auto tag ) {
      using T = decltype( tag );
      const BndGeom g = BndGeom::of( *op.st, lat );
      // boundary colour cc; each update of the non-CSR kernel also refreshes the
      // DoF's halo copies, the CSR kernel updates the owner only
      auto boundaryColour = [&]( size_t cc ) {
         const int64_t n = o.colorStart[cc + 1] - o.colorStart[cc];
         if ( n == 0 )
            return;
         if ( o.csr )
         {
            const SellRows& sr    = o.sorRows;
            const int64_t   slot0 = o.sorColorStart[cc], ns = o.sorColorStart[cc + 1] - slot0;
            const int64_t   nblk  = ( ( ns << lv ) + kBlock - 1 ) / kBlock;
            kBndSORSell< T ><<< xcdGrid( nblk ), kBlock, 0, ctx->stream >>>(
                nblk, slot0, ns, sr.row.p, lat.bndFlag.p, flag, sr.off.p, sr.pos.p, sr.val.p, lat.nc * lat.sc,
                o.invDiag.p, relax, (const T*) bp, (T*) xp, lv );
         }
         else
            kBndSOR< T ><<< gridFor( n << lv ), kBlock, 0, ctx->stream >>>( g, n, o.colorRows.p + o.colorStart[cc],
                                                                           lat.nc, lat.bndFlag.p, flag, o.clsStencil.p,
                                                                           o.invDiag.p, relax, (const T*) bp, (T*) xp,
                                                                           lv );
         HG_CHECK_HIP( hipGetLastError() );
      };
      // cell-interior parity colours k = cx + 2 cy + 4 cz, two at a time: the
      // row group grp = cy + 2 cz with cx = cxFirst, then 1 - cxFirst
      auto interiorGroup = [&]( int grp, int cxFirst ) {
         const int64_t n = lat.rowGroup[grp + 1] - lat.rowGroup[grp];
         if ( n == 0 )
            return;
         const int rpb = rowsPerBlockLog( ( ( lat.mm - 2 ) / 2 ) << lv );
         kCellSOR< T ><<< cellGrid( n * lat.nc, rpb ), 64 * kCellWorkers, 0, ctx->stream >>>(
             lat.rows.p + 2 * lat.rowGroup[grp], n, lat.nc, lat.sc, lat.w, lat.mm, o.stencil.p, cxFirst, relax,
             (const T*) bp, (T*) xp, lv, rpb );
         HG_CHECK_HIP( hipGetLastError() );
      };
      const size_t nbc      = o.colorStart.empty() ? 0 : o.colorStart.size() - 1;
      const bool   interior = ( flag & kInner ) && lat.nrows > 0;
      if ( interior )
         checkCellIndexing( lat, lv );
      if ( !backward )
      {
         for ( size_t cc = 0; cc < nbc; ++cc )
            boundaryColour( cc );
         // the CSR sweep updated the owners only: refresh the cell copies the
         // interior sweep reads
         if ( o.csr && nbc > 0 )
            pullHalos( ctx, *op.st, lat, xp, x.dtype, lv );
         if ( interior )
            for ( int grp = 0; grp < 4; ++grp )
               interiorGroup( grp, 0 );
      }
      else
      {
         if ( interior )
            for ( int grp = 3; grp >= 0; --grp )
               interiorGroup( grp, 1 );
         for ( size_t cc = nbc; cc-- > 0; )
            boundaryColour( cc );
         if ( o.csr && nbc > 0 )
            pullHalos( ctx, *op.st, lat, xp, x.dtype, lv );
      }
   } );
   ctx->endTiming();
}

void p1Restrict( Context* ctx, Function& f, int fineL, int flag )
{
   if ( fineL - 1 < f.minL || fineL > f.maxL || f.off != 0 )
      throw Error( -1, "P1 restriction source level outside the function's level range" );
   Storage& st = *f.st;
   Lattice& lf = st.lattice( ctx, fineL );
   Lattice& lc = st.lattice( ctx, fineL - 1 );
   void*    fp = f.ptr( fineL );
   void*    cp = f.ptr( fineL - 1 );
   ctx->beginTiming();
   pullHalos( ctx, st, lf, fp, f.dtype, f.lv );
   dispatchType( f.dtype, [&]( auto tag ) {
      using T = decltype( tag );
      if ( lc.nrows > 0 )
      {
         kRestrictCell< T ><<< xcdGrid( lc.nrows * lc.nc ), kRowBlock, 0, ctx->stream >>>(
             lc.rowsNat.p, lc.nrows, lc.nc, lc.sc, lc.w, lc.mm, lf.sc, lf.w, (const T*) fp, (T*) cp, f.lv );
         HG_CHECK_HIP( hipGetLastError() );
      }
      if ( lc.nb > 0 )
      {
         kRestrictBnd< T ><<< gridFor( lc.nb << f.lv ), kBlock, 0, ctx->stream >>>(
             BndGeom::of( st, lc ), lc.nb, lc.nc, lc.bndFlag.p, flag, st.clsInvN.p, lf.sc, lf.w, lf.mm, (const T*) fp,
             (T*) cp, f.lv );
         HG_CHECK_HIP( hipGetLastError() );
      }
   } );
   ctx->endTiming();
}

void p1Prolongate( Context* ctx, Function& f, int coarseL, int flag, int update )
{
   if ( coarseL < f.minL || coarseL + 1 > f.maxL || f.off != 0 )
      throw Error( -1, "P1 prolongation source level outside the function's level range" );
   Storage& st = *f.st;
   Lattice& lf = st.lattice( ctx, coarseL + 1 );
   Lattice& lc = st.lattice( ctx, coarseL );
   void*    fp = f.ptr( coarseL + 1 );
   void*    cp = f.ptr( coarseL );
   ctx->beginTiming();
   pullHalos( ctx, st, lc, cp, f.dtype, f.lv );
   dispatchType( f.dtype, [&]( auto tag ) {
      using T = decltype( tag );
      if ( lf.nrows > 0 )
      {
         kProlongCell< T ><<< xcdGrid( lf.nrows * lf.nc ), kRowBlock, 0, ctx->stream >>>(
             lf.rowsNat.p, lf.nrows, lf.nc, lf.sc, lf.w, lf.mm, lc.sc, lc.w, (const T*) cp, (T*) fp, update, f.lv );
         HG_CHECK_HIP( hipGetLastError() );
      }
      if ( lf.nb > 0 )
      {
         kProlongBnd< T ><<< gridFor( lf.nb << f.lv ), kBlock, 0, ctx->stream >>>(
             BndGeom::of( st, lf ), lf.nb, lf.nc, lf.bndFlag.p, flag, lc.sc, lc.w, (const T*) cp, (T*) fp, update, f.lv );
         HG_CHECK_HIP( hipGetLastError() );
      }
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

int hg_p1_smooth_sor_backwards( hg_ctx* h, hg_handle op, hg_handle x, hg_handle b, double relax, int level, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      p1SOR( ctx, ctx->op< P1Operator >( op ), ctx->function( x ), ctx->function( b ), relax, level, flag, true );
   } );
}

int hg_p1_sor_colours( hg_ctx* h, hg_handle opH, int level, int32_t* colour, int* nBoundaryColours )
{
   return hg_guard( h, [&]( Context* ctx ) {
      P1Operator& op = ctx->op< P1Operator >( opH );
      if ( !colour || !nBoundaryColours )
         throw Error( HG_ERR_ARG, "sor_colours: null output" );
      P1OpLevel&      o   = levelOf( ctx, op, level );
      Lattice&        lat = op.st->lattice( ctx, level );
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
                  colour[lat.num.locate( t, c, x, y, z )] = nbc + int32_t( ( x & 1 ) + 2 * ( y & 1 ) + 4 * ( z & 1 ) );
      *nBoundaryColours = nbc;
   } );
}

int hg_p1_inverse_diagonal( hg_ctx* h, hg_handle opH, int level, hg_handle dst )
{
   return hg_guard( h, [&]( Context* ctx ) {
      P1Operator& op = ctx->op< P1Operator >( opH );
      Function&   d  = ctx->function( dst );
      if ( d.st != op.st || d.off != 0 || d.dtype != 0 || d.lv != 0 || level < d.minL || level > d.maxL )
         throw Error( HG_ERR_ARG, "inverse diagonal needs a single-vector FP64 P1 function on the operator's storage" );
      const double* inv = p1InverseDiagonal( ctx, op, level );
      Lattice&      lat = op.st->lattice( ctx, level );
      HG_CHECK_HIP( hipMemcpyAsync( d.ptr( level ), inv, size_t( lat.nc * lat.sc + lat.nb ) * sizeof( double ),
                                    hipMemcpyDeviceToDevice, ctx->stream ) );
   } );
}

int hg_p1_smooth_jac( hg_ctx* h, hg_handle opH, hg_handle dst, hg_handle rhs, hg_handle src, double relax, int level,
                      int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      P1Operator& op = ctx->op< P1Operator >( opH );
      Function&   d  = ctx->function( dst );
      Function&   r  = ctx->function( rhs );
      Function&   s  = ctx->function( src );
      if ( r.st != d.st || r.dtype != d.dtype || r.off != d.off || r.lv != d.lv || &r == &d || &s == &d )
         throw Error( HG_ERR_ARG, "smooth_jac needs dst, rhs, src of one storage, value type and vector count, dst distinct" );
      const double* inv = p1InverseDiagonal( ctx, op, level );
      p1Apply( ctx, op, s, d, level, flag, HG_REPLACE );
      jacobiUpdate( ctx, op.st->lattice( ctx, level ), d.dtype, d.lv, flag, relax, r.ptr( level ), s.ptr( level ), inv,
                    d.ptr( level ) );
   } );
}

int hg_p1_gemv( hg_ctx* h, hg_handle opH, double alpha, hg_handle src, double beta, hg_handle dst, int level, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      P1Operator& op = ctx->op< P1Operator >( opH );
      Function&   s  = ctx->function( src );
      Function&   d  = ctx->function( dst );
      if ( beta == 0.0 )
      {
         p1Apply( ctx, op, s, d, level, flag, HG_REPLACE, alpha );
         return;
      }
      if ( beta != 1.0 )
         scaleFlagged( ctx, d, level, flag, beta );
      p1Apply( ctx, op, s, d, level, flag, HG_ADD, alpha );
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
