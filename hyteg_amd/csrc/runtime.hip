// Context, storage, per-level lattice layout, halo exchange, functions and
// BLAS-1 kernels of the hyteg_amd library, plus their C-ABI entry points.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include "abi_util.hpp"
#include "halo.hpp"
#include "hyteg_amd.h"
#include "runtime.hpp"

namespace hg {

// ------------------------------------------------------------------------
// device buffers
// ------------------------------------------------------------------------
// ------------------------------------------------------------------------
// storage / lattice set-up
// ------------------------------------------------------------------------
void Storage::init( Context* ctx )
{
   const Topology&           t = topo;
   std::vector< CellMeta >   m( t.nc );
   for ( int c = 0; c < t.nc; ++c )
   {
      auto localOf = [&]( int g ) {
         for ( int k = 0; k < 4; ++k )
            if ( t.cv[c][k] == g )
               return int8_t( k );
         throw Error( -4, "inconsistent topology" );
      };
      for ( int k = 0; k < 4; ++k )
      {
         m[c].face[k] = t.cf[c][k];
         for ( int j = 0; j < 3; ++j )
            m[c].fperm[k][j] = localOf( t.fv[t.cf[c][k]][j] );
         m[c].vert[k] = t.cv[c][k];
      }
      for ( int k = 0; k < 6; ++k )
      {
         m[c].edge[k] = t.ce[c][k];
         for ( int j = 0; j < 2; ++j )
            m[c].eperm[k][j] = localOf( t.ev[t.ce[c][k]][j] );
      }
   }
   meta.upload( ctx, m );

   auto csr = [&]( int nprim, auto primOfLocal, int nlocal, DevBuf< int32_t >& ptr, DevBuf< int32_t >& ent ) {
      std::vector< std::vector< int32_t > > lists( nprim );
      for ( int c = 0; c < t.nc; ++c )
         for ( int k = 0; k < nlocal; ++k )
         {
            lists[primOfLocal( c, k )].push_back( c );
            lists[primOfLocal( c, k )].push_back( k );
         }
      std::vector< int32_t > p( nprim + 1, 0 ), e;
      for ( int i = 0; i < nprim; ++i )
      {
         p[i + 1] = p[i] + int32_t( lists[i].size() / 2 );
         e.insert( e.end(), lists[i].begin(), lists[i].end() );
      }
      ptr.upload( ctx, p );
      ent.upload( ctx, e );
   };
   std::vector< double > inv( 14 * size_t( t.nc ) );
   for ( int c = 0; c < t.nc; ++c )
   {
      for ( int k = 0; k < 4; ++k )
         inv[14 * c + k] = 1.0 / t.fnc[t.cf[c][k]];
      for ( int k = 0; k < 6; ++k )
         inv[14 * c + 4 + k] = 1.0 / t.enc[t.ce[c][k]];
      for ( int k = 0; k < 4; ++k )
         inv[14 * c + 10 + k] = 1.0 / t.vnc[t.cv[c][k]];
   }
   clsInvN.upload( ctx, inv );
   csr( t.nf, [&]( int c, int k ) { return t.cf[c][k]; }, 4, faceCellPtr, faceCells );
   csr( t.ne, [&]( int c, int k ) { return t.ce[c][k]; }, 6, edgeCellPtr, edgeCells );
   csr( t.nv, [&]( int c, int k ) { return t.cv[c][k]; }, 4, vertCellPtr, vertCells );
}

Lattice& Storage::lattice( Context* ctx, int L )
{
   if ( L < 0 || L > 12 )
      throw Error( -1, "refinement level out of range" );
   auto it = lattices.find( L );
   if ( it != lattices.end() )
      return *it->second;
   auto lp     = std::make_unique< Lattice >();
   Lattice& l  = *lp;
   l.L         = L;
   l.num       = LevelNumbering( topo, L );
   l.mm        = l.num.mm;
   l.w         = l.mm + 1;
   l.sc        = tet( l.w );
   l.nc        = topo.nc;
   l.nb        = l.num.coff;
   l.nstore    = l.nc * l.sc + l.nb;

   l.hBndFlag.resize( l.nb );
   for ( int64_t i = 0; i < l.nb; ++i )
      l.hBndFlag[i] = uint8_t( l.num.flagOf( topo, i ) );
   l.bndFlag.upload( ctx, l.hBndFlag );

   // interior bit mask of one cell lattice, replicated over cells
   const int64_t          ncellbits = l.nc * l.sc;
   std::vector< uint32_t > mask( ( ncellbits + 31 ) / 32, 0u );
   for ( int64_t c = 0; c < l.nc; ++c )
      for ( int64_t z = 1; z < l.mm; ++z )
         for ( int64_t y = 1; y + z < l.mm; ++y )
            for ( int64_t x = 1; x + y + z < l.mm; ++x )
            {
               const int64_t pos = c * l.sc + tetIndex( l.w, x, y, z );
               mask[pos >> 5] |= 1u << ( pos & 31 );
            }
   l.cellMask.upload( ctx, mask );

   std::vector< int32_t > rows;
   for ( int g = 0; g < 4; ++g )
   {
      l.rowGroup[g] = int64_t( rows.size() / 2 );
      for ( int64_t z = 1; z <= l.mm - 3; ++z )
         for ( int64_t y = 1; y + z <= l.mm - 2; ++y )
         {
            if ( ( y & 1 ) + 2 * ( z & 1 ) != g )
               continue;
            rows.push_back( int32_t( y ) );
            rows.push_back( int32_t( z ) );
         }
   }
   l.rowGroup[4] = int64_t( rows.size() / 2 );
   l.nrows       = l.rowGroup[4];
   l.rows.upload( ctx, rows );
   rows.clear();
   for ( int64_t z = 1; z <= l.mm - 3; ++z )
      for ( int64_t y = 1; y + z <= l.mm - 2; ++y )
      {
         rows.push_back( int32_t( y ) );
         rows.push_back( int32_t( z ) );
      }
   l.rowsNat.upload( ctx, rows );

   lattices[L] = std::move( lp );
   return l;
}

template < class T >
__global__ void kPull( const CellMeta* __restrict__ meta, int64_t nc, int64_t sc, int64_t w, int64_t mm, int64_t fin,
                       int64_t ein, int64_t eoff, int64_t foff, int lv, T* __restrict__ v )
{
   const int64_t per = 4 * fin + 6 * ein + 4;
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t gid = tid >> lv;
   const int64_t vec = tid & ( ( int64_t( 1 ) << lv ) - 1 );
   if ( gid >= per * nc )
      return;
   const int64_t   c = gid / per;
   int64_t         j = gid - c * per;
   const CellMeta& m = meta[c];
   int64_t         t[4];
   int64_t         src;
   const int64_t   base = nc * sc;
   if ( j < 4 * fin )
   {
      const int     k = int( j / fin );
      const int64_t i = j - k * fin;
      int64_t       a, b;
      decodeTri( mm - 2, i, a, b );
      faceToCell( m, k, mm, a + 1, b + 1, t );
      src = base + foff + int64_t( m.face[k] ) * fin + i;
   }
   else if ( ( j -= 4 * fin ) < 6 * ein )
   {
      const int     k = int( j / ein );
      const int64_t s = j - k * ein + 1;
      edgeToCell( m, k, mm, s, t );
      src = base + eoff + int64_t( m.edge[k] ) * ein + ( s - 1 );
   }
   else
   {
      const int k = int( j - 6 * ein );
      t[0] = t[1] = t[2] = t[3] = 0;
      t[k]                      = mm;
      src                       = base + m.vert[k];
   }
   v[( ( c * sc + tetIndex( w, t[1], t[2], t[3] ) ) << lv ) + vec] = v[( src << lv ) + vec];
}

template < class T >
__global__ void kPushAdditive( BndGeom g, const uint8_t* __restrict__ bndFlag, int64_t nb, int64_t nc,
                               int excludeFlag, int zeroOut, int lv, T* __restrict__ v )
{
   using A           = typename Acc< T >::type;
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t b   = tid >> lv;
   const int64_t vec = tid & ( ( int64_t( 1 ) << lv ) - 1 );
   if ( b >= nb || ( bndFlag[b] & excludeFlag ) )
      return;
   const int64_t base = nc * g.sc;
   A             s    = zeroOut ? A( 0 ) : toAcc( v[( ( base + b ) << lv ) + vec] );
   // the halo copies in the order of forEachIncident, eight loads in flight (a
   // macro-vertex of the shell has up to ~36 incident cells)
   int32_t e0, e1;
   incidentRange( g, b, e0, e1 );
   for ( int32_t e = e0; e < e1; e += 8 )
   {
      A buf[8];
#pragma unroll
      for ( int j = 0; j < 8; ++j )
      {
         buf[j] = A( 0 );
         if ( e + j < e1 )
         {
            int64_t       x, y, z;
            const int64_t c = incidentAt( g, b, e + j, x, y, z );
            buf[j]          = toAcc( v[( ( c * g.sc + tetIndex( g.w, x, y, z ) ) << lv ) + vec] );
         }
      }
#pragma unroll
      for ( int j = 0; j < 8; ++j )
         if ( e + j < e1 )
            s += buf[j];
   }
   v[( ( base + b ) << lv ) + vec] = fromAcc< T >( s );
}

void pullHalos( Context* ctx, Storage& st, Lattice& lat, void* data, int dtype, int lv )
{
   const int64_t n = lat.haloPerCell() * lat.nc;
   if ( n == 0 )
      return;
   dispatchType( dtype, [&]( auto tag ) {
      using T = decltype( tag );
      kPull< T ><<< gridFor( n << lv ), kBlock, 0, ctx->stream >>>( st.meta.p, lat.nc, lat.sc, lat.w, lat.mm,
                                                                   lat.num.fin, lat.num.ein, lat.num.eoff,
                                                                   lat.num.foff, lv, (T*) data );
   } );
   HG_CHECK_HIP( hipGetLastError() );
}

void pushAdditive( Context* ctx, Storage& st, Lattice& lat, void* data, int dtype, int excludeFlag, bool zeroOut, int lv )
{
   if ( lat.nb == 0 )
      return;
   dispatchType( dtype, [&]( auto tag ) {
      using T = decltype( tag );
      kPushAdditive< T ><<< gridFor( lat.nb << lv ), kBlock, 0, ctx->stream >>>(
          BndGeom::of( st, lat ), lat.bndFlag.p, lat.nb, lat.nc, excludeFlag, zeroOut ? 1 : 0, lv, (T*) data );
   } );
   HG_CHECK_HIP( hipGetLastError() );
}

// ------------------------------------------------------------------------
// BLAS-1 kernels over owned DoFs.  The cell part is processed flat (halo
// copies included: they are refreshed before any read anyway); the owned
// vertex/edge/face part is filtered by DoFType.
// ------------------------------------------------------------------------
// (2^lv interleaved vectors: thread i handles value i & (2^lv - 1) of point i >> lv)
template < class T >
__global__ void kLinComb( int64_t ncell, int64_t nb, const uint8_t* __restrict__ bndFlag, int flag, int nsrc,
                          double s0, double s1, double s2, const T* __restrict__ f0, const T* __restrict__ f1,
                          const T* __restrict__ f2, int add, int lv, T* __restrict__ dst )
{
   using A           = typename Acc< T >::type;
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t i   = tid >> lv;
   const int64_t n   = ( flag & kInner ? ncell : 0 ) + nb;
   if ( i >= n )
      return;
   const int64_t pt = ( flag & kInner ) ? i : ncell + i;
   if ( pt >= ncell && !( bndFlag[pt - ncell] & flag ) )
      return;
   const int64_t p = ( pt << lv ) + ( tid & ( ( int64_t( 1 ) << lv ) - 1 ) );
   A v = A( 0 );
   if ( nsrc > 0 )
      v += A( s0 ) * toAcc( f0[p] );
   if ( nsrc > 1 )
      v += A( s1 ) * toAcc( f1[p] );
   if ( nsrc > 2 )
      v += A( s2 ) * toAcc( f2[p] );
   if ( add )
      v += toAcc( dst[p] );
   dst[p] = fromAcc< T >( v );
}

// dst (+)= sum_i c_i f_i with c_i = sign_i * (slots[num_i] / slots[den_i]); a
// negative slot index stands for 1.  The quotient is formed first and then
// signed, which is exactly the host's alpha = num / den followed by -alpha.  A
// zero denominator gives c_i = 0 (a Krylov step on an exactly solved system
// leaves the iterate alone instead of turning it into NaN).
struct SlotCoef
{
   double sign;
   int    num, den;
   int    gate; // slot index: the coefficient is 0 while that slot is nonzero (-1: none)
};
__device__ __forceinline__ double slotCoef( const SlotCoef& c, const double* __restrict__ slots )
{
   if ( c.gate >= 0 && slots[c.gate] != 0.0 )
      return 0.0;
   const double n = c.num >= 0 ? slots[c.num] : 1.0;
   const double d = c.den >= 0 ? slots[c.den] : 1.0;
   return d == 0.0 ? 0.0 : c.sign * ( n / d );
}
template < class T >
__global__ void kLinCombSlots( int64_t ncell, int64_t nb, const uint8_t* __restrict__ bndFlag, int flag, int nsrc,
                               SlotCoef c0, SlotCoef c1, SlotCoef c2, const double* __restrict__ slots,
                               const T* __restrict__ f0, const T* __restrict__ f1, const T* __restrict__ f2, int add,
                               int lv, T* __restrict__ dst )
{
   using A           = typename Acc< T >::type;
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t i   = tid >> lv;
   const int64_t n   = ( flag & kInner ? ncell : 0 ) + nb;
   if ( i >= n )
      return;
   const int64_t pt = ( flag & kInner ) ? i : ncell + i;
   if ( pt >= ncell && !( bndFlag[pt - ncell] & flag ) )
      return;
   const int64_t p = ( pt << lv ) + ( tid & ( ( int64_t( 1 ) << lv ) - 1 ) );
   A v = A( 0 );
   if ( nsrc > 0 )
      v += A( slotCoef( c0, slots ) ) * toAcc( f0[p] );
   if ( nsrc > 1 )
      v += A( slotCoef( c1, slots ) ) * toAcc( f1[p] );
   if ( nsrc > 2 )
      v += A( slotCoef( c2, slots ) ) * toAcc( f2[p] );
   if ( add )
      v += toAcc( dst[p] );
   dst[p] = fromAcc< T >( v );
}

// the stopping test of CGSolver (CGSolver.hpp:180-186) on device scalars:
// init: done = (|r0| < absTol); otherwise, while done == 0,
// done = (|r| / |r0| < relTol or |r| < absTol), with |r| = sqrt(slot rr)
__global__ void kSlotsCGStop( double* __restrict__ slots, int rr, int rr0, double relTol, double absTol, int done,
                              int init )
{
   if ( threadIdx.x != 0 || blockIdx.x != 0 )
      return;
   const double r0 = sqrt( slots[rr0] );
   if ( init )
   {
      slots[done] = r0 < absTol ? 1.0 : 0.0;
      return;
   }
   if ( slots[done] != 0.0 )
      return;
   const double r = sqrt( slots[rr] );
   slots[done]    = ( r / r0 < relTol || r < absTol ) ? 1.0 : 0.0;
}

template < class T >
__global__ void kFill( int64_t ncell, int64_t nb, const uint8_t* __restrict__ bndFlag, int flag, double value, int lv,
                       T* __restrict__ dst )
{
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t i   = tid >> lv;
   const int64_t n   = ( flag & kInner ? ncell : 0 ) + nb;
   if ( i >= n )
      return;
   const int64_t pt = ( flag & kInner ) ? i : ncell + i;
   if ( pt >= ncell && !( bndFlag[pt - ncell] & flag ) )
      return;
   dst[( pt << lv ) + ( tid & ( ( int64_t( 1 ) << lv ) - 1 ) )] = fromAcc< T >( typename Acc< T >::type( value ) );
}

// per-vector scaling v[p][k] *= scale[k] on flagged DoFs
template < class T >
__global__ void kScaleVectors( int64_t ncell, int64_t nb, const uint8_t* __restrict__ bndFlag, int flag,
                               const double* __restrict__ scale, int lv, T* __restrict__ v )
{
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t i   = tid >> lv;
   const int64_t n   = ( flag & kInner ? ncell : 0 ) + nb;
   if ( i >= n )
      return;
   const int64_t pt = ( flag & kInner ) ? i : ncell + i;
   if ( pt >= ncell && !( bndFlag[pt - ncell] & flag ) )
      return;
   const int64_t k = tid & ( ( int64_t( 1 ) << lv ) - 1 );
   const int64_t p = ( pt << lv ) + k;
   v[p]            = fromAcc< T >( typename Acc< T >::type( scale[k] ) * toAcc( v[p] ) );
}

// SplitMix64 of (seed, canonical DoF id) -> uniform [lo, hi): independent of
// the storage layout and of the traversal order
__device__ __forceinline__ double canonicalUniform( uint64_t seed, uint64_t id, double lo, double hi )
{
   uint64_t z = seed + ( id + 1 ) * 0x9E3779B97F4A7C15ull;
   z          = ( z ^ ( z >> 30 ) ) * 0xBF58476D1CE4E5B9ull;
   z          = ( z ^ ( z >> 27 ) ) * 0x94D049BB133111EBull;
   z ^= z >> 31;
   return lo + ( hi - lo ) * ( double( z >> 11 ) * ( 1.0 / 9007199254740992.0 ) );
}

// cell-interior points: one block per (cell, interior row)
// (vector k of an interleaved function uses seed + k)
template < class T >
__global__ void kRandomCells( const int32_t* __restrict__ rows, int64_t nrows, int64_t sc, int64_t w, int64_t mm,
                              int64_t coff, int64_t cin, uint64_t seed, double lo, double hi, int lv,
                              T* __restrict__ v )
{
   const int64_t c   = blockIdx.y;
   const int64_t r   = blockIdx.x;
   const int64_t y   = rows[2 * r], z = rows[2 * r + 1];
   const int64_t out = c * sc + tetRowStart( w, y, z );
   const int64_t cid = coff + c * cin + tetIndex( mm - 3, 0, y - 1, z - 1 );
   const int64_t msk = ( int64_t( 1 ) << lv ) - 1;
   for ( int64_t t = threadIdx.x;; t += blockDim.x )
   {
      const int64_t x = 1 + ( t >> lv ), k = t & msk;
      if ( x + y + z >= mm )
         break;
      v[( ( out + x ) << lv ) + k] = fromAcc< T >(
          typename Acc< T >::type( canonicalUniform( seed + uint64_t( k ), uint64_t( cid + x - 1 ), lo, hi ) ) );
   }
}

template < class T >
__global__ void kRandomOwned( int64_t nb, int64_t base, const uint8_t* __restrict__ bndFlag, int flag, uint64_t seed,
                              double lo, double hi, int lv, T* __restrict__ v )
{
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t b   = tid >> lv, k = tid & ( ( int64_t( 1 ) << lv ) - 1 );
   if ( b < nb && ( bndFlag[b] & flag ) )
      v[( ( base + b ) << lv ) + k] =
          fromAcc< T >( typename Acc< T >::type( canonicalUniform( seed + uint64_t( k ), uint64_t( b ), lo, hi ) ) );
}

template < class TD, class TS >
__global__ void kCast( int64_t n, const TS* __restrict__ src, TD* __restrict__ dst )
{
   const int64_t i = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( i < n )
   {
      double v;
      if constexpr ( std::is_same< TS, __half >::value )
         v = double( __half2float( src[i] ) );
      else
         v = double( src[i] );
      if constexpr ( std::is_same< TD, __half >::value )
         dst[i] = __float2half( float( v ) );
      else
         dst[i] = TD( v );
   }
}

// dst vector kd (of 2^ld interleaved) = src vector ks (of 2^ls), n points
template < class T >
__global__ void kCopyVector( int64_t n, const T* __restrict__ src, int ls, int ks, T* __restrict__ dst, int ld, int kd )
{
   const int64_t i = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   if ( i < n )
      dst[( i << ld ) + kd] = src[( i << ls ) + ks];
}

template < class T >
__global__ void kJacobiUpdate( int64_t ncell, int64_t nb, const uint8_t* __restrict__ bndFlag, int flag, double relax,
                               const T* __restrict__ rhs, const T* __restrict__ src,
                               const double* __restrict__ invDiag, int lv, T* __restrict__ dst )
{
   // the statements (and so the fused multiply-adds) of kLinComb and kMult
   using A           = typename Acc< T >::type;
   const int64_t tid = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
   const int64_t i   = tid >> lv;
   if ( i >= ncell + nb )
      return;
   const bool on = i < ncell ? ( flag & kInner ) != 0 : ( bndFlag[i - ncell] & flag ) != 0;
   if ( !on )
      return;
   A v = A( 0 );
   v += A( 1.0 ) * toAcc( rhs[tid] );
   v += A( -1.0 ) * toAcc( dst[tid] );
   const T r = fromAcc< T >( v );
   const T t = fromAcc< T >( toAcc( fromAcc< T >( A( invDiag[i] ) ) ) * toAcc( r ) );
   A       u = A( 0 );
   u += A( 1.0 ) * toAcc( src[tid] );
   u += A( relax ) * toAcc( t );
   dst[tid] = fromAcc< T >( u );
}

void jacobiUpdate( Context* ctx, Lattice& lat, int dtype, int lv, int flag, double relax, const void* rhs,
                   const void* src, const double* invDiag, void* dst )
{
   const int64_t n = lat.nc * lat.sc + lat.nb;
   if ( n == 0 )
      return;
   dispatchType( dtype, [&]( auto tag ) {
      using T = decltype( tag );
      kJacobiUpdate< T ><<< gridFor( n << lv ), kBlock, 0, ctx->stream >>>(
          lat.nc * lat.sc, lat.nb, lat.bndFlag.p, flag, relax, (const T*) rhs, (const T*) src, invDiag, lv, (T*) dst );
   } );
   HG_CHECK_HIP( hipGetLastError() );
}

void castCopy( Context* ctx, const void* src, int sdtype, void* dst, int ddtype, int64_t n )
{
   dispatchType( ddtype, [&]( auto td ) {
      dispatchType( sdtype, [&]( auto ts ) {
         using TD = decltype( td );
         using TS = decltype( ts );
         kCast< TD, TS ><<< gridFor( n ), kBlock, 0, ctx->stream >>>( n, (const TS*) src, (TD*) dst );
      } );
   } );
   HG_CHECK_HIP( hipGetLastError() );
}

// per-block partial sums; with 2^lv interleaved vectors thread t always sees
// vector t & (2^lv - 1), and partials[block * 2^lv + k] holds vector k's sum
template < class T >
__global__ void kDotPartial( int64_t ncell, int64_t nb, const uint32_t* __restrict__ cellMask,
                             const uint8_t* __restrict__ bndFlag, int flag, const T* __restrict__ a,
                             const T* __restrict__ b, int lv, double* __restrict__ partials )
{
   __shared__ double red[kBlock];
   double            s      = 0.0;
   const int64_t     k      = threadIdx.x & ( ( 1 << lv ) - 1 );
   const int64_t     stride = ( int64_t( gridDim.x ) * blockDim.x ) >> lv;
   const int64_t     first  = ( int64_t( blockIdx.x ) * blockDim.x + threadIdx.x ) >> lv;
   if ( flag & kInner )
      for ( int64_t p = first; p < ncell; p += stride )
         if ( ( cellMask[p >> 5] >> ( p & 31 ) ) & 1u )
            s += double( toAcc( a[( p << lv ) + k] ) ) * double( toAcc( b[( p << lv ) + k] ) );
   for ( int64_t i = first; i < nb; i += stride )
      if ( bndFlag[i] & flag )
         s += double( toAcc( a[( ( ncell + i ) << lv ) + k] ) ) * double( toAcc( b[( ( ncell + i ) << lv ) + k] ) );
   red[threadIdx.x] = s;
   __syncthreads();
   for ( int o = kBlock / 2; o >= ( 1 << lv ); o >>= 1 )
   {
      if ( threadIdx.x < o )
         red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
   }
   if ( threadIdx.x < ( 1 << lv ) )
      partials[( int64_t( blockIdx.x ) << lv ) + threadIdx.x] = red[threadIdx.x];
}

// out[k] = sum over blocks of partials[block * 2^lv + k], one block per vector
__global__ void kDotFinal( const double* __restrict__ partials, int n, int lv, double* __restrict__ out )
{
   __shared__ double red[kBlock];
   double            s = 0.0;
   for ( int i = threadIdx.x; i < n; i += blockDim.x )
      s += partials[( int64_t( i ) << lv ) + blockIdx.x];
   red[threadIdx.x] = s;
   __syncthreads();
   for ( int o = kBlock / 2; o > 0; o >>= 1 )
   {
      if ( threadIdx.x < o )
         red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
   }
   if ( threadIdx.x == 0 )
      out[blockIdx.x] = red[0];
}

// launches the reduction of the 2^lv dot products into dev[0 .. 2^lv) on the
// context's stream; the grid covers the flagged entries (one block at least)
static int dotBlocks( const Lattice& lat, int flag, int lv )
{
   const int64_t n = ( ( flag & kInner ? lat.nc * lat.sc : 0 ) + lat.nb ) << lv;
   return int( std::min< int64_t >( kDotBlocks, std::max< int64_t >( 1, ( n + kBlock - 1 ) / kBlock ) ) );
}

// partials: (blocks << lv) doubles of scratch
static void dotLaunchWith( Context* ctx, Lattice& lat, const void* a, const void* b, int dtype, int flag, int lv,
                           double* partials, double* dev )
{
   const int nv     = 1 << lv;
   const int blocks = dotBlocks( lat, flag, lv );
   dispatchType( dtype, [&]( auto tag ) {
      using T = decltype( tag );
      kDotPartial< T ><<< blocks, kBlock, 0, ctx->stream >>>( lat.nc * lat.sc, lat.nb, lat.cellMask.p, lat.bndFlag.p,
                                                             flag, (const T*) a, (const T*) b, lv, partials );
   } );
   kDotFinal<<< nv, kBlock, 0, ctx->stream >>>( partials, blocks, lv, dev );
   HG_CHECK_HIP( hipGetLastError() );
}

static void dotLaunch( Context* ctx, Lattice& lat, const void* a, const void* b, int dtype, int flag, int lv,
                       double* dev )
{
   dotLaunchWith( ctx, lat, a, b, dtype, flag, lv, ctx->partials.p, dev );
}

// ---- dot products without allocation ------------------------------------------
// Partial sums and results of hg_function_dot_batched_noalloc live in this
// array: one copy per device, held until the process ends.  The name is a
// host-side stand-in, not a device address: host code reaches each device's
// copy only through the address recorded by recordScratch (scratchOf).
constexpr size_t kScratchDoubles = ( size_t( kDotBlocks ) << kMaxLv ) + ( size_t( 1 ) << kMaxLv );
static __device__ __align__( 256 ) double scratch[kScratchDoubles];

static std::mutex& scratchMutex()
{
   static std::mutex m;
   return m;
}
static std::map< int, double* >& scratchTable()
{
   static std::map< int, double* > t;
   return t;
}
// once per device, with that device current (context creation)
static void recordScratch( int device )
{
   std::lock_guard< std::mutex > lk( scratchMutex() );
   if ( scratchTable().count( device ) )
      return;
   double* ptr = nullptr;
   HG_CHECK_HIP( hipGetSymbolAddress( (void**) &ptr, HIP_SYMBOL( scratch ) ) );
   scratchTable()[device] = ptr;
}
static double* scratchOf( int device )
{
   std::lock_guard< std::mutex > lk( scratchMutex() );
   auto                          it = scratchTable().find( device );
   if ( it == scratchTable().end() )
      throw Error( HG_ERR_STATE, "no reduction scratch recorded for this device" );
   return it->second;
}

void dotDeviceBatched( Context* ctx, Lattice& lat, const void* a, const void* b, int dtype, int flag, int lv, double* out )
{
   const int nv  = 1 << lv;
   double*   res = ctx->partials.p + ( int64_t( kDotBlocks ) << kMaxLv );
   dotLaunch( ctx, lat, a, b, dtype, flag, lv, res );
   HG_CHECK_HIP( hipMemcpyAsync( ctx->hostResult, res, sizeof( double ) * nv, hipMemcpyDeviceToHost, ctx->stream ) );
   HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
   for ( int k = 0; k < nv; ++k )
      out[k] = ctx->hostResult[k];
}

double dotDevice( Context* ctx, Lattice& lat, const void* a, const void* b, int dtype, int flag )
{
   double r;
   dotDeviceBatched( ctx, lat, a, b, dtype, flag, 0, &r );
   return r;
}

// ------------------------------------------------------------------------
// context helpers
// ------------------------------------------------------------------------
Storage& Context::storage( int64_t h )
{
   auto it = storages.find( h );
   if ( it == storages.end() )
      throw Error( -3, "unknown storage handle" );
   return *it->second;
}
Function& Context::function( int64_t h )
{
   auto it = functions.find( h );
   if ( it == functions.end() )
      throw Error( -3, "unknown function handle" );
   return *it->second;
}
__global__ void kZeroBytes( uint8_t* __restrict__ p, size_t bytes )
{
   const size_t i  = ( size_t( blockIdx.x ) * blockDim.x + threadIdx.x ) * 16;
   if ( i + 16 <= bytes )
      *reinterpret_cast< uint4* >( p + i ) = make_uint4( 0u, 0u, 0u, 0u );
   else
      for ( size_t k = i; k < bytes; ++k )
         p[k] = 0;
}

void zeroDevice( Context* c, void* p, size_t bytes )
{
   if ( bytes == 0 )
      return;
   // device allocations are 256-byte aligned: 16-byte stores, a byte tail
   kZeroBytes<<< gridFor( int64_t( ( bytes + 15 ) / 16 ) ), kBlock, 0, c->stream >>>( static_cast< uint8_t* >( p ),
                                                                                     bytes );
   HG_CHECK_HIP( hipGetLastError() );
}

void* Context::deviceAlloc( size_t bytes )
{
   if ( capturing )
      throw Error( HG_ERR_STATE, "device allocation while a graph capture is open" );
   auto it = pool.find( bytes );
   if ( it != pool.end() )
   {
      void* p = it->second;
      pool.erase( it );
      poolBytes -= bytes;
      return p;
   }
   void*          p = nullptr;
   DeviceGateLock gate;
   HG_CHECK_HIP( hipMalloc( &p, bytes ) );
   return p;
}
void Context::deviceFree( void* p, size_t bytes )
{
   if ( capturing )
   {
      heldByCapture.push_back( p );
      return;
   }
   if ( bytes <= kPoolBlock && poolBytes + bytes <= kPoolBytes )
   {
      pool.emplace( bytes, p );
      poolBytes += bytes;
      return;
   }
   DeviceGateLock gate;
   (void) hipFree( p );
}
void Context::drainPool()
{
   DeviceGateLock gate;
   for ( auto& e : pool )
      (void) hipFree( e.second );
   pool.clear();
   poolBytes = 0;
}
void Context::beginTiming()
{
   if ( timing )
      HG_CHECK_HIP( hipEventRecord( ev0, stream ) );
}
void Context::endTiming()
{
   if ( timing )
   {
      HG_CHECK_HIP( hipEventRecord( ev1, stream ) );
      HG_CHECK_HIP( hipEventSynchronize( ev1 ) );
      float ms = 0.f;
      HG_CHECK_HIP( hipEventElapsedTime( &ms, ev0, ev1 ) );
      lastMs = ms;
   }
}

static void checkLevel( const Function& f, int L )
{
   if ( L < f.minL || L > f.maxL )
      throw Error( -1, "level " + std::to_string( L ) + " outside the function's level range" );
}

// host conversion helpers for canonical FP64 transfer
template < class T >
static T hostFrom( double v )
{
   if constexpr ( std::is_same< T, __half >::value )
      return __float2half( float( v ) );
   else
      return T( v );
}
template < class T >
static double hostTo( T v )
{
   if constexpr ( std::is_same< T, __half >::value )
      return double( __half2float( v ) );
   else
      return double( v );
}

template < class F >
static void forEachCanonical( const Lattice& l, F&& f ) // f(canonical id, storage pos)
{
   for ( int64_t b = 0; b < l.nb; ++b )
      f( b, l.nc * l.sc + b );
   int64_t id = l.num.coff;
   for ( int64_t c = 0; c < l.nc; ++c )
      for ( int64_t z = 1; z < l.mm; ++z )
         for ( int64_t y = 1; y + z < l.mm; ++y )
            for ( int64_t x = 1; x + y + z < l.mm; ++x )
               f( id++, c * l.sc + tetIndex( l.w, x, y, z ) );
}

} // namespace hg

using namespace hg;

// ==========================================================================
// C ABI
// ==========================================================================
namespace hg {
void setContextError( hg_ctx* h, const char* msg )
{
   if ( h )
      reinterpret_cast< Context* >( h )->err = msg ? msg : "";
}
std::mutex& deviceGate()
{
   static std::mutex gate;
   return gate;
}

// ends the capture on ctx's stream and releases the gate it holds
static hipError_t closeCapture( Context* ctx, hipGraph_t* g )
{
   const hipError_t e = hipStreamEndCapture( ctx->stream, g );
   if ( ctx->capturing )
   {
      ctx->capturing = false;
      --tlsCaptures;
      deviceGate().unlock();
   }
   return e;
}

void scaleFlagged( Context* ctx, Function& f, int level, int flag, double beta )
{
   checkLevel( f, level );
   Lattice&      l = f.lat( ctx, level );
   const int64_t n = ( flag & kInner ? l.nc * l.sc : 0 ) + l.nb;
   if ( n == 0 )
      return;
   dispatchType( f.dtype, [&]( auto tag ) {
      using T = decltype( tag );
      T* d    = (T*) f.ptr( level );
      kLinComb< T ><<< gridFor( n << f.lv ), kBlock, 0, ctx->stream >>>( l.nc * l.sc, l.nb, l.bndFlag.p, flag, 1, beta,
                                                                         0.0, 0.0, d, nullptr, nullptr, 0, f.lv, d );
   } );
   HG_CHECK_HIP( hipGetLastError() );
}

} // namespace hg

extern "C" {

int hg_ctx_create( int device, hg_ctx** out )
{
   if ( !out )
      return HG_ERR_ARG;
   *out = nullptr;
   auto* ctx = new Context();
   try
   {
      ctx->device = device;
      HG_CHECK_HIP( hipSetDevice( device ) );
      {
         DeviceGateLock gate;
         HG_CHECK_HIP( hipStreamCreateWithFlags( &ctx->stream, hipStreamNonBlocking ) );
         HG_CHECK_HIP(
             hipHostMalloc( (void**) &ctx->hostResult, sizeof( double ) * ( 1 << kMaxLv ), hipHostMallocDefault ) );
         HG_CHECK_HIP( hipEventCreate( &ctx->ev0 ) );
         HG_CHECK_HIP( hipEventCreate( &ctx->ev1 ) );
      }
      ctx->partials.alloc( ctx, ( size_t( kDotBlocks ) << kMaxLv ) + ( 1 << kMaxLv ) );
      recordScratch( device );
      ctx->slots.alloc( ctx, kSlots );
      ctx->slotUsed.assign( kSlots, 0 );
      HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
   } catch ( const std::exception& e )
   {
      delete ctx;
      return HG_ERR_HIP;
   }
   *out = reinterpret_cast< hg_ctx* >( ctx );
   return HG_OK;
}

static void treeRelease( Context* ctx );

int hg_ctx_destroy( hg_ctx* h )
{
   if ( !h )
      return HG_ERR_ARG;
   auto* ctx = reinterpret_cast< Context* >( h );
   (void) hipSetDevice( ctx->device );
   if ( ctx->capturing )
   {
      hipGraph_t g = nullptr;
      (void) closeCapture( ctx, &g );
      if ( g )
         (void) hipGraphDestroy( g );
   }
   if ( ctx->stream )
      (void) hipStreamSynchronize( ctx->stream );
   {
      DeviceGateLock gate;
      for ( auto& g : ctx->graphs )
         (void) hipGraphExecDestroy( g.second );
   }
   ctx->graphs.clear();
   ctx->operators.clear();
   ctx->functions.clear();
   ctx->storages.clear();
   treeRelease( ctx );
   ctx->partials.release();
   ctx->slots.release();
   ctx->drainPool();
   {
      DeviceGateLock gate;
      for ( void* p : ctx->heldByCapture )
         (void) hipFree( p );
      ctx->heldByCapture.clear();
      if ( ctx->hostResult )
         (void) hipHostFree( ctx->hostResult );
      if ( ctx->ev0 )
         (void) hipEventDestroy( ctx->ev0 );
      if ( ctx->ev1 )
         (void) hipEventDestroy( ctx->ev1 );
      if ( ctx->stream )
         (void) hipStreamDestroy( ctx->stream );
   }
   delete ctx;
   return HG_OK;
}

// ---- timing tree ------------------------------------------------------------------
static hipEvent_t treeEvent( Context* ctx )
{
   TimingTree& t = ctx->tree;
   if ( !t.spare.empty() )
   {
      hipEvent_t e = t.spare.back();
      t.spare.pop_back();
      return e;
   }
   hipEvent_t e = nullptr;
   HG_CHECK_HIP( hipEventCreate( &e ) );
   return e;
}

// reads the elapsed times of the closed regions (waits for the stream)
static void treeCollect( Context* ctx )
{
   TimingTree& t = ctx->tree;
   if ( t.closed.empty() )
      return;
   HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
   for ( auto& c : t.closed )
   {
      float ms = 0.f;
      HG_CHECK_HIP( hipEventElapsedTime( &ms, c.a, c.b ) );
      auto& s = t.sums[c.path];
      s.first += 1;
      s.second += double( ms );
      t.spare.push_back( c.a );
      t.spare.push_back( c.b );
   }
   t.closed.clear();
}

static void treeRelease( Context* ctx )
{
   TimingTree& t = ctx->tree;
   for ( auto& o : t.open )
      t.spare.push_back( o.ev );
   for ( auto& c : t.closed )
   {
      t.spare.push_back( c.a );
      t.spare.push_back( c.b );
   }
   t.open.clear();
   t.closed.clear();
   t.stack.clear();
   for ( hipEvent_t e : t.spare )
      (void) hipEventDestroy( e );
   t.spare.clear();
}

const char* hg_last_error( const hg_ctx* h )
{
   if ( !h )
      return "null context";
   return reinterpret_cast< const Context* >( h )->err.c_str();
}

int hg_graph_begin_capture( hg_ctx* h )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( ctx->capturing )
         throw Error( HG_ERR_STATE, "a capture is already open on this context" );
      deviceGate().lock();
      ctx->capturing = true;
      ++tlsCaptures;
      const hipError_t e = hipStreamBeginCapture( ctx->stream, hipStreamCaptureModeThreadLocal );
      if ( e != hipSuccess )
      {
         ctx->capturing = false;
         --tlsCaptures;
         deviceGate().unlock();
         HG_CHECK_HIP( e );
      }
   } );
}

int hg_graph_end_capture( hg_ctx* h, hg_handle* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !ctx->capturing )
         throw Error( HG_ERR_STATE, "no capture is open on this context" );
      hipGraph_t       g  = nullptr;
      const hipError_t ec = closeCapture( ctx, &g );
      if ( !out )
      {
         if ( g )
            (void) hipGraphDestroy( g );
         throw Error( HG_ERR_ARG, "null output" );
      }
      HG_CHECK_HIP( ec );
      DeviceGateLock   gate;
      hipGraphExec_t   exec = nullptr;
      const hipError_t e    = hipGraphInstantiate( &exec, g, nullptr, nullptr, 0 );
      (void) hipGraphDestroy( g );
      HG_CHECK_HIP( e );
      *out               = ctx->nextHandle++;
      ctx->graphs[*out] = exec;
   } );
}

int hg_graph_launch( hg_ctx* h, hg_handle graph )
{
   return hg_guard( h, [&]( Context* ctx ) {
      auto it = ctx->graphs.find( graph );
      if ( it == ctx->graphs.end() )
         throw Error( HG_ERR_HANDLE, "unknown graph handle" );
      HG_CHECK_HIP( hipGraphLaunch( it->second, ctx->stream ) );
   } );
}

int hg_graph_destroy( hg_ctx* h, hg_handle graph )
{
   return hg_guard( h, [&]( Context* ctx ) {
      auto it = ctx->graphs.find( graph );
      if ( it == ctx->graphs.end() )
         throw Error( HG_ERR_HANDLE, "unknown graph handle" );
      HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
      DeviceGateLock gate;
      HG_CHECK_HIP( hipGraphExecDestroy( it->second ) );
      ctx->graphs.erase( it );
   } );
}

int hg_ctx_set_graphs( hg_ctx* h, int enable )
{
   return hg_guard( h, [&]( Context* ctx ) { ctx->useGraphs = enable != 0; } );
}

int hg_ctx_graphs_enabled( hg_ctx* h, int* enabled )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !enabled )
         throw Error( HG_ERR_ARG, "null output" );
      *enabled = ctx->useGraphs ? 1 : 0;
   } );
}

int hg_ctx_synchronize( hg_ctx* h )
{
   return hg_guard( h, [&]( Context* ctx ) { HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) ); } );
}

int hg_ctx_memory( hg_ctx* h, int64_t* bytes )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !bytes )
         throw Error( HG_ERR_ARG, "null output" );
      *bytes = ctx->bytes;
   } );
}

int hg_ctx_event_timing( hg_ctx* h, int enable )
{
   return hg_guard( h, [&]( Context* ctx ) { ctx->timing = enable != 0; } );
}

int hg_ctx_last_kernel_ms( hg_ctx* h, double* ms )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !ms )
         throw Error( HG_ERR_ARG, "null output" );
      *ms = ctx->lastMs;
   } );
}

static int64_t addStorage( Context* ctx, Topology&& t )
{
   auto st  = std::make_unique< Storage >();
   st->topo = std::move( t );
   st->init( ctx );
   const int64_t h    = ctx->nextHandle++;
   ctx->storages[h] = std::move( st );
   return h;
}

int hg_storage_cuboid( hg_ctx* h, const double* lo, const double* hi, int nx, int ny, int nz, hg_handle* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !lo || !hi || !out )
         throw Error( HG_ERR_ARG, "null argument" );
      *out = addStorage( ctx, Topology::cuboid( lo, hi, nx, ny, nz ) );
   } );
}

int hg_storage_create( hg_ctx* h, int nv, const double* coords, int nc, const int64_t* cells, hg_handle* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !coords || !cells || !out || nv < 4 || nc < 1 )
         throw Error( HG_ERR_ARG, "invalid mesh arrays" );
      *out = addStorage( ctx, Topology::fromCells( nv, coords, nc, cells ) );
   } );
}

int hg_storage_counts( hg_ctx* h, hg_handle s, int64_t* out4 )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !out4 )
         throw Error( HG_ERR_ARG, "null output" );
      const Topology& t = ctx->storage( s ).topo;
      out4[0]           = t.nv;
      out4[1]           = t.ne;
      out4[2]           = t.nf;
      out4[3]           = t.nc;
   } );
}

int hg_p1_ndofs( hg_ctx* h, hg_handle s, int level, int64_t* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !out || level < 0 || level > 12 )
         throw Error( HG_ERR_ARG, "invalid argument" );
      *out = LevelNumbering( ctx->storage( s ).topo, level ).n;
   } );
}

int hg_p1_dof_coords( hg_ctx* h, hg_handle s, int level, double* xyz )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !xyz || level < 0 || level > 12 )
         throw Error( HG_ERR_ARG, "invalid argument" );
      Storage&        st = ctx->storage( s );
      const Topology& t  = st.topo;
      LevelNumbering  nm( t, level );
      const double    mm = double( nm.mm );
      // every canonical DoF from the first cell that contains it
      std::vector< char > done( nm.n, 0 );
      for ( int c = 0; c < t.nc; ++c )
      {
         const double* v[4];
         for ( int k = 0; k < 4; ++k )
            v[k] = &t.X[3 * t.cv[c][k]];
         for ( int64_t z = 0; z <= nm.mm; ++z )
            for ( int64_t y = 0; y + z <= nm.mm; ++y )
               for ( int64_t x = 0; x + y + z <= nm.mm; ++x )
               {
                  const int64_t id = nm.locate( t, c, x, y, z );
                  if ( done[id] )
                     continue;
                  done[id] = 1;
                  for ( int d = 0; d < 3; ++d )
                     xyz[3 * id + d] = v[0][d] + ( x / mm ) * ( v[1][d] - v[0][d] ) + ( y / mm ) * ( v[2][d] - v[0][d] ) +
                                       ( z / mm ) * ( v[3][d] - v[0][d] );
               }
      }
   } );
}

int hg_p1_dof_flags( hg_ctx* h, hg_handle s, int level, int32_t* flags )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !flags || level < 0 || level > 12 )
         throw Error( HG_ERR_ARG, "invalid argument" );
      const Topology& t = ctx->storage( s ).topo;
      LevelNumbering  nm( t, level );
      for ( int64_t i = 0; i < nm.n; ++i )
         flags[i] = nm.flagOf( t, i );
   } );
}

static int functionCreate( hg_ctx* h, hg_handle s, int minLevel, int maxLevel, int dtype, int off, int nvec,
                           hg_handle* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !out || minLevel < 0 || maxLevel < minLevel || maxLevel + off > 12 || dtype < 0 || dtype > 2 )
         throw Error( HG_ERR_ARG, "invalid function parameters" );
      int lv = 0;
      while ( ( 1 << lv ) < nvec )
         ++lv;
      if ( nvec < 1 || ( 1 << lv ) != nvec || lv > kMaxLv )
         throw Error( HG_ERR_ARG, "the number of interleaved vectors must be a power of two up to 64" );
      Storage& st = ctx->storage( s );
      auto     f  = std::make_unique< Function >();
      f->st       = &st;
      f->dtype    = dtype;
      f->off      = off;
      f->lv       = lv;
      f->minL     = minLevel;
      f->maxL     = maxLevel;
      for ( int L = minLevel; L <= maxLevel; ++L )
      {
         Lattice& l = st.lattice( ctx, L + off );
         // two spare values: a 16-bit value is updated atomically through its
         // aligned 32-bit word, which for the last value may extend past nstore
         f->data[L + off].alloc( ctx, size_t( l.nstore + 2 ) * f->elemSize() << lv );
      }
      HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
      *out                  = ctx->nextHandle++;
      ctx->functions[*out] = std::move( f );
   } );
}

int hg_p1_function_create( hg_ctx* h, hg_handle s, int minLevel, int maxLevel, int dtype, hg_handle* out )
{
   return functionCreate( h, s, minLevel, maxLevel, dtype, 0, 1, out );
}

int hg_p2_function_create( hg_ctx* h, hg_handle s, int minLevel, int maxLevel, int dtype, hg_handle* out )
{
   return functionCreate( h, s, minLevel, maxLevel, dtype, 1, 1, out );
}

int hg_function_destroy( hg_ctx* h, hg_handle fn )
{
   return hg_guard( h, [&]( Context* ctx ) {
      HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
      if ( ctx->functions.erase( fn ) == 0 )
         throw Error( HG_ERR_HANDLE, "unknown function handle" );
   } );
}

int hg_function_storage( hg_ctx* h, hg_handle fn, hg_handle* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !out )
         throw Error( HG_ERR_ARG, "null output" );
      Function& f = ctx->function( fn );
      for ( auto& [k, v] : ctx->storages )
         if ( v.get() == f.st )
         {
            *out = k;
            return;
         }
      throw Error( HG_ERR_STATE, "function without storage" );
   } );
}

int hg_function_upload( hg_ctx* h, hg_handle fn, int level, const double* canonical )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& f = ctx->function( fn );
      checkLevel( f, level );
      if ( !canonical )
         throw Error( HG_ERR_ARG, "null data" );
      Lattice&  l  = f.lat( ctx, level );
      const int nv = f.nvec();
      dispatchType( f.dtype, [&]( auto tag ) {
         using T = decltype( tag );
         std::vector< T > hbuf( size_t( l.nstore ) * nv, hostFrom< T >( 0.0 ) );
         forEachCanonical( l, [&]( int64_t id, int64_t pos ) {
            for ( int k = 0; k < nv; ++k )
               hbuf[pos * nv + k] = hostFrom< T >( canonical[id * nv + k] );
         } );
         HG_CHECK_HIP( hipMemcpyAsync( f.ptr( level ), hbuf.data(), hbuf.size() * sizeof( T ), hipMemcpyHostToDevice,
                                       ctx->stream ) );
         pullHalos( ctx, *f.st, l, f.ptr( level ), f.dtype, f.lv );
         HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
      } );
   } );
}

int hg_function_download( hg_ctx* h, hg_handle fn, int level, double* canonical )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& f = ctx->function( fn );
      checkLevel( f, level );
      if ( !canonical )
         throw Error( HG_ERR_ARG, "null data" );
      Lattice&  l  = f.lat( ctx, level );
      const int nv = f.nvec();
      dispatchType( f.dtype, [&]( auto tag ) {
         using T = decltype( tag );
         std::vector< T > hbuf( size_t( l.nstore ) * nv );
         HG_CHECK_HIP( hipMemcpyAsync( hbuf.data(), f.ptr( level ), hbuf.size() * sizeof( T ), hipMemcpyDeviceToHost,
                                       ctx->stream ) );
         HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
         forEachCanonical( l, [&]( int64_t id, int64_t pos ) {
            for ( int k = 0; k < nv; ++k )
               canonical[id * nv + k] = hostTo< T >( hbuf[pos * nv + k] );
         } );
      } );
   } );
}

static int linComb( hg_ctx* h, hg_handle dst, int nsrc, const double* scalars, const hg_handle* srcs, int level, int flag,
                    int add )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( nsrc < 0 || nsrc > 3 || ( nsrc > 0 && ( !scalars || !srcs ) ) )
         throw Error( HG_ERR_ARG, "assign/add take 0..3 source functions" );
      Function& d = ctx->function( dst );
      checkLevel( d, level );
      void* src[3] = { nullptr, nullptr, nullptr };
      for ( int i = 0; i < nsrc; ++i )
      {
         Function& s = ctx->function( srcs[i] );
         checkLevel( s, level );
         if ( s.st != d.st || s.dtype != d.dtype || s.off != d.off || s.lv != d.lv )
            throw Error( HG_ERR_ARG, "assign/add need functions of one storage, value type and vector count" );
         src[i] = s.ptr( level );
      }
      Lattice&      l = d.lat( ctx, level );
      const int64_t n = ( flag & kInner ? l.nc * l.sc : 0 ) + l.nb;
      if ( n == 0 )
         return;
      double s[3] = { 0, 0, 0 };
      for ( int i = 0; i < nsrc; ++i )
         s[i] = scalars[i];
      dispatchType( d.dtype, [&]( auto tag ) {
         using T = decltype( tag );
         kLinComb< T ><<< gridFor( n << d.lv ), kBlock, 0, ctx->stream >>>(
             l.nc * l.sc, l.nb, l.bndFlag.p, flag, nsrc, s[0], s[1], s[2], (const T*) src[0], (const T*) src[1],
             (const T*) src[2], add, d.lv, (T*) d.ptr( level ) );
      } );
      HG_CHECK_HIP( hipGetLastError() );
   } );
}

int hg_function_assign( hg_ctx* h, hg_handle dst, int nsrc, const double* scalars, const hg_handle* srcs, int level, int flag )
{
   return linComb( h, dst, nsrc, scalars, srcs, level, flag, 0 );
}

int hg_function_add( hg_ctx* h, hg_handle dst, int nsrc, const double* scalars, const hg_handle* srcs, int level, int flag )
{
   return linComb( h, dst, nsrc, scalars, srcs, level, flag, 1 );
}

int hg_function_interpolate_const( hg_ctx* h, hg_handle fn, double value, int level, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& f = ctx->function( fn );
      checkLevel( f, level );
      Lattice&      l = f.lat( ctx, level );
      const int64_t n = ( flag & kInner ? l.nc * l.sc : 0 ) + l.nb;
      if ( n == 0 )
         return;
      dispatchType( f.dtype, [&]( auto tag ) {
         using T = decltype( tag );
         kFill< T ><<< gridFor( n << f.lv ), kBlock, 0, ctx->stream >>>( l.nc * l.sc, l.nb, l.bndFlag.p, flag, value,
                                                                         f.lv, (T*) f.ptr( level ) );
      } );
      HG_CHECK_HIP( hipGetLastError() );
   } );
}

int hg_function_interpolate_random( hg_ctx* h, hg_handle fn, uint64_t seed, double lo, double hi, int level, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& f = ctx->function( fn );
      checkLevel( f, level );
      Lattice& l = f.lat( ctx, level );
      dispatchType( f.dtype, [&]( auto tag ) {
         using T = decltype( tag );
         if ( ( flag & kInner ) && l.nrows > 0 )
         {
            dim3 grid( unsigned( l.nrows ), unsigned( l.nc ) );
            kRandomCells< T ><<< grid, 64, 0, ctx->stream >>>( l.rowsNat.p, l.nrows, l.sc, l.w, l.mm, l.num.coff,
                                                              l.num.cin, seed, lo, hi, f.lv, (T*) f.ptr( level ) );
         }
         if ( l.nb > 0 )
            kRandomOwned< T ><<< gridFor( l.nb << f.lv ), kBlock, 0, ctx->stream >>>(
                l.nb, l.nc * l.sc, l.bndFlag.p, flag, seed, lo, hi, f.lv, (T*) f.ptr( level ) );
      } );
      HG_CHECK_HIP( hipGetLastError() );
   } );
}

int hg_slots_alloc( hg_ctx* h, int n, int* first )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( n < 1 || n > kSlots || !first )
         throw Error( HG_ERR_ARG, "slots: invalid count" );
      for ( int f = 0; f + n <= kSlots; ++f )
      {
         int k = 0;
         while ( k < n && !ctx->slotUsed[size_t( f + k )] )
            ++k;
         if ( k == n )
         {
            for ( k = 0; k < n; ++k )
               ctx->slotUsed[size_t( f + k )] = 1;
            *first = f;
            return;
         }
         f += k;
      }
      throw Error( HG_ERR_STATE, "slots: all " + std::to_string( kSlots ) + " device scalar slots are in use" );
   } );
}

int hg_slots_free( hg_ctx* h, int first, int n )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( first < 0 || n < 1 || first + n > kSlots )
         throw Error( HG_ERR_ARG, "slots: invalid range" );
      for ( int k = 0; k < n; ++k )
         ctx->slotUsed[size_t( first + k )] = 0;
   } );
}

static void checkSlots( const Context* ctx, int first, int n )
{
   if ( first < 0 || n < 1 || first + n > kSlots )
      throw Error( HG_ERR_ARG, "slot index out of range" );
   for ( int k = 0; k < n; ++k )
      if ( !ctx->slotUsed[size_t( first + k )] )
         throw Error( HG_ERR_ARG, "slot " + std::to_string( first + k ) + " is not allocated" );
}

int hg_slots_read( hg_ctx* h, int first, int n, double* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      checkSlots( ctx, first, n );
      if ( !out || n > ( 1 << kMaxLv ) )
         throw Error( HG_ERR_ARG, "slots: invalid read" );
      HG_CHECK_HIP( hipMemcpyAsync( ctx->hostResult, ctx->slots.p + first, sizeof( double ) * size_t( n ),
                                    hipMemcpyDeviceToHost, ctx->stream ) );
      HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
      for ( int k = 0; k < n; ++k )
         out[k] = ctx->hostResult[k];
   } );
}

int hg_slots_write( hg_ctx* h, int first, int n, const double* values )
{
   return hg_guard( h, [&]( Context* ctx ) {
      checkSlots( ctx, first, n );
      if ( !values )
         throw Error( HG_ERR_ARG, "slots: invalid write" );
      // a blocking copy, which a stream capture would record: refused there
      if ( tlsCaptures > 0 )
         throw Error( HG_ERR_STATE, "slots: write during a stream capture" );
      DeviceGateLock gate;
      HG_CHECK_HIP( hipMemcpyAsync( ctx->slots.p + first, values, sizeof( double ) * size_t( n ), hipMemcpyHostToDevice,
                                    ctx->stream ) );
      HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
   } );
}

int hg_function_dot_slot( hg_ctx* h, hg_handle a, hg_handle b, int level, int flag, int slot )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& fa = ctx->function( a );
      Function& fb = ctx->function( b );
      checkLevel( fa, level );
      checkLevel( fb, level );
      if ( fa.st != fb.st || fa.dtype != fb.dtype || fa.off != fb.off || fa.lv != fb.lv )
         throw Error( HG_ERR_ARG, "dot needs functions of one storage, value type and vector count" );
      checkSlots( ctx, slot, fa.nvec() );
      dotLaunch( ctx, fa.lat( ctx, level ), fa.ptr( level ), fb.ptr( level ), fa.dtype, flag, fa.lv,
                 ctx->slots.p + slot );
   } );
}

int hg_function_lincomb_slots( hg_ctx* h, hg_handle dst, int nsrc, const double* signs, const int* num, const int* den,
                               const hg_handle* srcs, int level, int flag, int add )
{
   return hg_function_lincomb_slots_gated( h, dst, nsrc, signs, num, den, srcs, level, flag, add, -1 );
}

int hg_function_lincomb_slots_gated( hg_ctx* h, hg_handle dst, int nsrc, const double* signs, const int* num,
                                     const int* den, const hg_handle* srcs, int level, int flag, int add, int gate )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( gate >= 0 )
         checkSlots( ctx, gate, 1 );
      if ( nsrc < 0 || nsrc > 3 || ( nsrc > 0 && ( !signs || !num || !den || !srcs ) ) )
         throw Error( HG_ERR_ARG, "lincomb_slots takes 0..3 source functions" );
      Function& d = ctx->function( dst );
      checkLevel( d, level );
      void*    src[3] = { nullptr, nullptr, nullptr };
      SlotCoef c[3]   = { { 0.0, -1, -1, -1 }, { 0.0, -1, -1, -1 }, { 0.0, -1, -1, -1 } };
      for ( int i = 0; i < nsrc; ++i )
      {
         Function& s = ctx->function( srcs[i] );
         checkLevel( s, level );
         if ( s.st != d.st || s.dtype != d.dtype || s.off != d.off || s.lv != d.lv )
            throw Error( HG_ERR_ARG, "lincomb_slots needs functions of one storage, value type and vector count" );
         if ( num[i] >= 0 )
            checkSlots( ctx, num[i], 1 );
         if ( den[i] >= 0 )
            checkSlots( ctx, den[i], 1 );
         src[i] = s.ptr( level );
         c[i]   = SlotCoef{ signs[i], num[i], den[i], gate };
      }
      Lattice&      l = d.lat( ctx, level );
      const int64_t n = ( flag & kInner ? l.nc * l.sc : 0 ) + l.nb;
      if ( n == 0 )
         return;
      dispatchType( d.dtype, [&]( auto tag ) {
         using T = decltype( tag );
         kLinCombSlots< T ><<< gridFor( n << d.lv ), kBlock, 0, ctx->stream >>>(
             l.nc * l.sc, l.nb, l.bndFlag.p, flag, nsrc, c[0], c[1], c[2], ctx->slots.p, (const T*) src[0],
             (const T*) src[1], (const T*) src[2], add != 0, d.lv, (T*) d.ptr( level ) );
      } );
      HG_CHECK_HIP( hipGetLastError() );
   } );
}

int hg_slots_cg_stop( hg_ctx* h, int rr, int rr0, double relTol, double absTol, int done, int init )
{
   return hg_guard( h, [&]( Context* ctx ) {
      checkSlots( ctx, rr, 1 );
      checkSlots( ctx, rr0, 1 );
      checkSlots( ctx, done, 1 );
      kSlotsCGStop<<< 1, 64, 0, ctx->stream >>>( ctx->slots.p, rr, rr0, relTol, absTol, done, init );
      HG_CHECK_HIP( hipGetLastError() );
   } );
}

int hg_function_dot( hg_ctx* h, hg_handle a, hg_handle b, int level, int flag, double* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !out )
         throw Error( HG_ERR_ARG, "null output" );
      Function& fa = ctx->function( a );
      Function& fb = ctx->function( b );
      checkLevel( fa, level );
      checkLevel( fb, level );
      if ( fa.st != fb.st || fa.dtype != fb.dtype || fa.off != fb.off || fa.lv != fb.lv )
         throw Error( HG_ERR_ARG, "dot needs functions of one storage, value type and vector count" );
      if ( fa.lv != 0 )
         throw Error( HG_ERR_ARG, "dot of interleaved vectors: use hg_function_dot_batched" );
      *out = dotDevice( ctx, fa.lat( ctx, level ), fa.ptr( level ), fb.ptr( level ), fa.dtype, flag );
   } );
}

int hg_function_dot_batched( hg_ctx* h, hg_handle a, hg_handle b, int level, int flag, double* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !out )
         throw Error( HG_ERR_ARG, "null output" );
      Function& fa = ctx->function( a );
      Function& fb = ctx->function( b );
      checkLevel( fa, level );
      checkLevel( fb, level );
      if ( fa.st != fb.st || fa.dtype != fb.dtype || fa.off != fb.off || fa.lv != fb.lv )
         throw Error( HG_ERR_ARG, "dot needs functions of one storage, value type and vector count" );
      dotDeviceBatched( ctx, fa.lat( ctx, level ), fa.ptr( level ), fb.ptr( level ), fa.dtype, flag, fa.lv, out );
   } );
}

int hg_function_dot_batched_noalloc( hg_ctx* h, hg_handle a, hg_handle b, int level, int flag, double* out )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !out )
         throw Error( HG_ERR_ARG, "null output" );
      Function& fa = ctx->function( a );
      Function& fb = ctx->function( b );
      checkLevel( fa, level );
      checkLevel( fb, level );
      if ( fa.st != fb.st || fa.dtype != fb.dtype || fa.off != fb.off || fa.lv != fb.lv )
         throw Error( HG_ERR_ARG, "dot needs functions of one storage, value type and vector count" );
      // the lattice of a function's level exists once the function does; looked
      // up, never built here (building one allocates)
      auto li = fa.st->lattices.find( level + fa.off );
      if ( li == fa.st->lattices.end() )
         throw Error( HG_ERR_STATE, "dot_noalloc: the level's lattice is not set up" );
      Lattice&     lat    = *li->second;
      const int    nv     = 1 << fa.lv;
      const size_t needed = ( size_t( dotBlocks( lat, flag, fa.lv ) ) << fa.lv ) + size_t( nv );
      if ( needed > kScratchDoubles )
         throw Error( HG_ERR_ARG, "dot_noalloc: " + std::to_string( needed ) + " partial sums exceed the scratch array of " +
                                      std::to_string( kScratchDoubles ) + " doubles" );
      double* ptr = scratchOf( ctx->device );
      double* res = ptr + ( size_t( kDotBlocks ) << kMaxLv );
      // every partial that kDotFinal reads is written by kDotPartial first, so
      // the array needs no clearing between calls
      dotLaunchWith( ctx, lat, fa.ptr( level ), fb.ptr( level ), fa.dtype, flag, fa.lv, ptr, res );
      HG_CHECK_HIP( hipMemcpyAsync( ctx->hostResult, res, sizeof( double ) * size_t( nv ), hipMemcpyDeviceToHost,
                                    ctx->stream ) );
      HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
      for ( int k = 0; k < nv; ++k )
         out[k] = ctx->hostResult[k];
   } );
}

int hg_timing_enable( hg_ctx* h, int enable )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !enable && !ctx->tree.open.empty() )
         throw Error( HG_ERR_STATE, "timing: regions are still open" );
      ctx->tree.enabled = enable != 0;
   } );
}

int hg_timing_start( hg_ctx* h, const char* name )
{
   return hg_guard( h, [&]( Context* ctx ) {
      TimingTree& t = ctx->tree;
      if ( !t.enabled || ctx->capturing )
         return;
      if ( !name || !*name || std::strchr( name, '/' ) )
         throw Error( HG_ERR_ARG, "timing: region names are non-empty and contain no '/'" );
      std::string path = t.stack.empty() ? std::string( name ) : t.open.back().path + "/" + name;
      hipEvent_t  e    = treeEvent( ctx );
      HG_CHECK_HIP( hipEventRecord( e, ctx->stream ) );
      t.open.push_back( { std::move( path ), e } );
      t.stack.push_back( name );
   } );
}

int hg_timing_stop( hg_ctx* h, const char* name )
{
   return hg_guard( h, [&]( Context* ctx ) {
      TimingTree& t = ctx->tree;
      if ( !t.enabled || ctx->capturing )
         return;
      if ( !name || t.stack.empty() || t.stack.back() != name )
         throw Error( HG_ERR_STATE, std::string( "timing: stop of a region that is not the innermost open one: " ) +
                                        ( name ? name : "(null)" ) );
      hipEvent_t e = treeEvent( ctx );
      HG_CHECK_HIP( hipEventRecord( e, ctx->stream ) );
      t.closed.push_back( { t.open.back().path, t.open.back().ev, e } );
      t.open.pop_back();
      t.stack.pop_back();
      if ( t.closed.size() >= 4096 )
         treeCollect( ctx );
   } );
}

int hg_timing_report( hg_ctx* h, char* buf, int64_t size, int64_t* needed )
{
   return hg_guard( h, [&]( Context* ctx ) {
      treeCollect( ctx );
      std::string out;
      for ( const auto& kv : ctx->tree.sums )
      {
         char line[64];
         std::snprintf( line, sizeof( line ), "\t%lld\t%.6f\n", (long long) kv.second.first, kv.second.second );
         out += kv.first + line;
      }
      if ( needed )
         *needed = int64_t( out.size() ) + 1;
      if ( buf && size > 0 )
      {
         const size_t n = std::min< size_t >( out.size(), size_t( size - 1 ) );
         std::memcpy( buf, out.data(), n );
         buf[n] = 0;
      }
   } );
}

int hg_timing_reset( hg_ctx* h )
{
   return hg_guard( h, [&]( Context* ctx ) {
      treeCollect( ctx );
      ctx->tree.sums.clear();
   } );
}

int hg_function_scale_batched( hg_ctx* h, hg_handle fn, const double* scales, int level, int flag )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& f = ctx->function( fn );
      checkLevel( f, level );
      if ( !scales )
         throw Error( HG_ERR_ARG, "null scales" );
      Lattice&      l = f.lat( ctx, level );
      const int64_t n = ( flag & kInner ? l.nc * l.sc : 0 ) + l.nb;
      if ( n == 0 )
         return;
      double* dev = ctx->partials.p + ( int64_t( kDotBlocks ) << kMaxLv );
      HG_CHECK_HIP( hipMemcpyAsync( dev, scales, sizeof( double ) * f.nvec(), hipMemcpyHostToDevice, ctx->stream ) );
      dispatchType( f.dtype, [&]( auto tag ) {
         using T = decltype( tag );
         kScaleVectors< T ><<< gridFor( n << f.lv ), kBlock, 0, ctx->stream >>>( l.nc * l.sc, l.nb, l.bndFlag.p, flag,
                                                                                 dev, f.lv, (T*) f.ptr( level ) );
      } );
      HG_CHECK_HIP( hipGetLastError() );
      // the scales live in the reduction scratch: wait before the next reduction reuses it
      HG_CHECK_HIP( hipStreamSynchronize( ctx->stream ) );
   } );
}

int hg_p1_function_create_batched( hg_ctx* h, hg_handle s, int minLevel, int maxLevel, int dtype, int nvec,
                                   hg_handle* out )
{
   return functionCreate( h, s, minLevel, maxLevel, dtype, 0, nvec, out );
}

int hg_p2_function_create_batched( hg_ctx* h, hg_handle s, int minLevel, int maxLevel, int dtype, int nvec,
                                   hg_handle* out )
{
   return functionCreate( h, s, minLevel, maxLevel, dtype, 1, nvec, out );
}

int hg_function_copy_vector( hg_ctx* h, hg_handle dst, int kdst, hg_handle src, int ksrc, int level )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& d = ctx->function( dst );
      Function& s = ctx->function( src );
      checkLevel( d, level );
      checkLevel( s, level );
      if ( d.st != s.st || d.off != s.off || d.dtype != s.dtype )
         throw Error( HG_ERR_ARG, "copy_vector needs functions of one storage, space and value type" );
      if ( kdst < 0 || kdst >= d.nvec() || ksrc < 0 || ksrc >= s.nvec() )
         throw Error( HG_ERR_ARG, "vector index out of range" );
      Lattice&      l = d.lat( ctx, level );
      const int64_t n = l.nstore;
      dispatchType( d.dtype, [&]( auto tag ) {
         using T = decltype( tag );
         kCopyVector< T ><<< gridFor( n ), kBlock, 0, ctx->stream >>>( n, (const T*) s.ptr( level ), s.lv, ksrc,
                                                                      (T*) d.ptr( level ), d.lv, kdst );
      } );
      HG_CHECK_HIP( hipGetLastError() );
   } );
}

int hg_function_nvec( hg_ctx* h, hg_handle fn, int* nvec )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !nvec )
         throw Error( HG_ERR_ARG, "null output" );
      *nvec = ctx->function( fn ).nvec();
   } );
}

int hg_function_copy_from( hg_ctx* h, hg_handle dst, hg_handle src, int level )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& d = ctx->function( dst );
      Function& s = ctx->function( src );
      checkLevel( d, level );
      checkLevel( s, level );
      if ( d.st != s.st || d.off != s.off || d.lv != s.lv )
         throw Error( HG_ERR_ARG, "copyFrom needs functions of one storage, space and vector count" );
      Lattice& l = d.lat( ctx, level );
      castCopy( ctx, s.ptr( level ), s.dtype, d.ptr( level ), d.dtype, l.nstore << d.lv );
   } );
}

int hg_function_communicate( hg_ctx* h, hg_handle fn, int level )
{
   return hg_guard( h, [&]( Context* ctx ) {
      Function& f = ctx->function( fn );
      checkLevel( f, level );
      pullHalos( ctx, *f.st, f.lat( ctx, level ), f.ptr( level ), f.dtype, f.lv );
   } );
}

} // extern "C"
