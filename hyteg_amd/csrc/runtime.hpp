// Internal runtime of the hyteg_amd library: context, device buffers,
// per-level lattice layout and halo maps, functions.  Not part of the ABI.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "lattice.hpp"
#include "topology.hpp"

namespace hg {

struct Error : std::runtime_error
{
   int code;
   Error( int c, const std::string& m )
   : std::runtime_error( m )
   , code( c )
   {}
};

#define HG_CHECK_HIP( expr )                                                                                       \
   do                                                                                                             \
   {                                                                                                              \
      hipError_t hg_e_ = ( expr );                                                                                \
      if ( hg_e_ != hipSuccess )                                                                                  \
         throw ::hg::Error( -2, std::string( #expr ) + ": " + hipGetErrorString( hg_e_ ) );                        \
   } while ( 0 )

struct Context;

// Owning device allocation, accounted in its context.
template < class T >
struct DevBuf
{
   T*       p   = nullptr;
   size_t   n   = 0;
   Context* ctx = nullptr;

   DevBuf() = default;
   DevBuf( const DevBuf& ) = delete;
   DevBuf& operator=( const DevBuf& ) = delete;
   DevBuf( DevBuf&& o ) noexcept { swap( o ); }
   DevBuf& operator=( DevBuf&& o ) noexcept
   {
      swap( o );
      return *this;
   }
   ~DevBuf() { release(); }
   void swap( DevBuf& o ) noexcept
   {
      std::swap( p, o.p );
      std::swap( n, o.n );
      std::swap( ctx, o.ctx );
   }
   void alloc( Context* c, size_t count );
   void upload( Context* c, const std::vector< T >& h );
   void release();
};

// Per-cell topology as seen from the cell: ids of its faces/edges/vertices and
// which cell-local vertex each of their (sorted) vertices is.  Level-free.
struct CellMeta
{
   int32_t face[4];
   int8_t  fperm[4][3];
   int32_t edge[6];
   int8_t  eperm[6][2];
   int32_t vert[4];
};

// One storage, one level: lattice layout and owners.  Storage of a function on
// this level is
//   [ nc macro-cell lattices of sc points each | nb owned vertex/edge/face DoFs ]
// where the owned part is in canonical order (vertices, edge interiors, face
// interiors) and the lattice points on a cell's boundary are halo copies of
// owned DoFs.
struct Lattice
{
   int            L = 0;
   int64_t        mm = 1, w = 2, sc = 0, nc = 0, nb = 0, nstore = 0;
   LevelNumbering num;

   DevBuf< uint8_t >  bndFlag;  // DoFType of each owned vertex/edge/face DoF
   DevBuf< uint32_t > cellMask; // bit set on cell-owned (interior) lattice points
   DevBuf< double >   ones;     // nstore ones (sums as dot products), built on first use
   DevBuf< double >   bndNormal; // x/|x| of each owned boundary DoF (radial normals), built on first use
   std::vector< uint8_t > hBndFlag;

   // interior rows (y,z) of a cell lattice, y,z >= 1, y+z <= mm-2, grouped by
   // (y%2, z%2); group g = (y%2) + 2*(z%2) occupies [rowGroup[g], rowGroup[g+1])
   DevBuf< int32_t > rows; // 2 ints per row
   DevBuf< int32_t > rowsNat; // the same rows in natural order (z-major, then y)
   int64_t           rowGroup[5] = { 0, 0, 0, 0, 0 };
   int64_t           nrows       = 0;

   // cell-interior points (x | y << 10 | z << 20), slice z by slice, and in a
   // slice grouped by parity class q = x%2 + 2 (y%2) + 4 (z%2), each group
   // ordered by (y, x) and padded to whole wavefronts (0xFFFFFFFF), so a
   // wavefront holds one class; classWave[8] = wavefronts per cell.  Built on
   // first use (P2 restriction).
   DevBuf< uint32_t > classPoints;
   int32_t            classWave[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };

   // cell-local offsets of the points with a coordinate divisible by 8 (the
   // points the 9^3-point boxes of the P2 tile kernels share), ordered by
   // (z, y, x).  Built on first use (FP16 reduced-form apply).
   DevBuf< int32_t > surfacePoints;

   // cell-interior points (x | y << 10 | z << 20) in chunks of 4 wavefronts: the
   // rows of a chunk share (y mod 4, z mod 4), wavefront r of a chunk holds
   // their points with x mod 4 = r (padded with 0xFFFFFFFF).  Built on first use
   // (P2 prolongation into this lattice).
   DevBuf< uint32_t > prolongChunks;

   // tile lists of the P2 tile kernels on this lattice (all tiles, interior
   // tiles, last-layer tiles), built on first use
   DevBuf< uint64_t > tileLists[3];

   // bricks of the P1 cell-interior apply on this lattice (built on first use)
   DevBuf< uint64_t > p1Bricks;

   int64_t cellPos( int c, int64_t x, int64_t y, int64_t z ) const { return c * sc + tetIndex( w, x, y, z ); }
   int64_t ownerPos( const Topology& t, int c, int64_t x, int64_t y, int64_t z ) const
   {
      const int64_t id = num.locate( t, c, x, y, z );
      return id < num.coff ? nc * sc + id : cellPos( c, x, y, z );
   }
   // halo points per cell: 4 face interiors, 6 edge interiors, 4 vertices
   int64_t haloPerCell() const { return 4 * num.fin + 6 * num.ein + 4; }
};

struct OperatorBase
{
   virtual ~OperatorBase() = default;
};

struct Storage
{
   Topology                                     topo;
   // grid-transfer data keyed by fine level (built on first use)
   std::map< int, std::unique_ptr< OperatorBase > > transfers;
   std::map< int, std::unique_ptr< Lattice > >  lattices;
   DevBuf< CellMeta >                           meta;
   // 1 / (number of cells) of each cell's local faces (0..3), edges (4..9) and
   // vertices (10..13): nc*14
   DevBuf< double >                             clsInvN;
   // cells incident to each face / edge / vertex: CSR of (cell, local index)
   DevBuf< int32_t > faceCellPtr, faceCells, edgeCellPtr, edgeCells, vertCellPtr, vertCells;
   Lattice&                                     lattice( Context* ctx, int L );
   void                                         init( Context* ctx );
};

// A function on levels [minL, maxL] of one storage.  P1 functions of level L
// live on lattice level L; P2 functions of level L on lattice level L+1 (their
// nodes are the micro-vertices and micro-edge midpoints of level L), so
// `off` = 0 for P1 and 1 for P2.
struct Function
{
   Storage*                           st    = nullptr;
   int                                dtype = 0; // 0 f64, 1 f32, 2 f16
   int                                off   = 0;
   int                                lv    = 0; // log2 of the number of interleaved vectors
   int                                minL = 0, maxL = 0;
   std::map< int, DevBuf< uint8_t > > data; // by lattice level, nstore elements
   size_t                             elemSize() const { return dtype == 0 ? 8 : ( dtype == 1 ? 4 : 2 ); }
   int                                nvec() const { return 1 << lv; }
   void*                              ptr( int L )
   {
      auto it = data.find( L + off );
      if ( it == data.end() )
         throw Error( -1, "function has no data on level " + std::to_string( L ) );
      return it->second.p;
   }
   Lattice& lat( Context* ctx, int L ) { return st->lattice( ctx, L + off ); }
};

// Nested named timing regions on a context's stream (the role of waLBerla's
// WcTimingTree in HyTeG's solvers): each region is bracketed by two HIP events;
// elapsed times are read and summed per path ("A/B/C") when a report is asked
// for, so recording a region does not wait for the device.
struct TimingTree
{
   bool                       enabled = false;
   std::vector< std::string > stack;
   struct Open
   {
      std::string path;
      hipEvent_t  ev;
   };
   struct Closed
   {
      std::string path;
      hipEvent_t  a, b;
   };
   std::vector< Open >                                   open;
   std::vector< Closed >                                 closed;
   std::vector< hipEvent_t >                             spare;
   std::map< std::string, std::pair< int64_t, double > > sums; // path -> (count, ms)
};

struct Context
{
   int                                                device = 0;
   hipStream_t                                        stream = nullptr;
   std::string                                        err;
   int64_t                                            bytes      = 0;
   int64_t                                            nextHandle = 1;
   std::map< int64_t, std::unique_ptr< Storage > >      storages;
   std::map< int64_t, hipGraphExec_t >                  graphs; // instantiated captures
   bool                                                 capturing = false; // holds the device gate
   bool                                                 useGraphs = true;
   std::map< int64_t, std::unique_ptr< Function > >     functions;
   std::map< int64_t, std::unique_ptr< OperatorBase > > operators;

   // FP64 reduction scratch
   DevBuf< double > partials;
   double*          hostResult = nullptr; // pinned

   // device-resident FP64 scalars of Krylov loops (dot results and the step
   // ratios formed from them), handed out in ranges by hg_slots_alloc
   DevBuf< double >    slots;
   std::vector< char > slotUsed;

   bool       timing = false;
   hipEvent_t ev0 = nullptr, ev1 = nullptr;
   double     lastMs = 0.0;
   TimingTree tree;

   // Freed device blocks are kept for allocations of the same size (the
   // context's work runs on its one stream, so a reused block's earlier work is
   // ordered before its new use): hipFree waits for every stream of the device,
   // which a solve that builds and drops per-request operators would otherwise
   // pay on every request.  At most kPoolBytes are kept, in blocks of at most
   // kPoolBlock bytes; the rest is freed.
   static constexpr size_t            kPoolBytes = size_t( 2 ) << 30;
   static constexpr size_t            kPoolBlock = size_t( 256 ) << 20;
   std::multimap< size_t, void* >     pool;
   size_t                             poolBytes = 0;
   // blocks released while a graph capture is open: the graph may still use
   // them, so they are only freed when the context is destroyed
   std::vector< void* >               heldByCapture;
   // device memory for a DevBuf; refused (HG_ERR_STATE) while this context
   // captures a graph: an allocation's zero fill would become part of the graph
   void*                              deviceAlloc( size_t bytes );
   void                               deviceFree( void* p, size_t bytes );
   void                               drainPool();

   Storage&  storage( int64_t h );
   Function& function( int64_t h );
   template < class Op >
   Op& op( int64_t h )
   {
      auto it = operators.find( h );
      if ( it == operators.end() )
         throw Error( -3, "unknown operator handle" );
      auto* p = dynamic_cast< Op* >( it->second.get() );
      if ( !p )
         throw Error( -3, "operator handle of wrong kind" );
      return *p;
   }
   void beginTiming();
   void endTiming();
};

// ---- device gate ------------------------------------------------------------
// A stream capture must not overlap with another thread's capture, nor with
// calls that set up or tear down device state (hipFree waits for every stream
// of the device): such overlaps invalidate the capture.  A capture holds this
// gate from its begin to its end; device allocation, deallocation, graph
// instantiation and destruction, and stream and pinned-memory set-up hold it
// for the call.  A thread that is capturing skips the gate (it must not
// allocate then anyway: the HIP call fails instead of dead-locking).
std::mutex&             deviceGate();
inline thread_local int tlsCaptures = 0;
struct DeviceGateLock
{
   std::unique_lock< std::mutex > lk;
   DeviceGateLock()
   {
      if ( tlsCaptures == 0 )
         lk = std::unique_lock< std::mutex >( deviceGate() );
   }
};

// zero `bytes` bytes at p (a device allocation) with a fill kernel on the
// context's stream
void zeroDevice( Context* c, void* p, size_t bytes );

// ---- DevBuf members (need the complete Context) ------------------------
template < class T >
inline void DevBuf< T >::alloc( Context* c, size_t count )
{
   release();
   ctx = c;
   n   = count;
   if ( count == 0 )
      return;
   p = static_cast< T* >( c->deviceAlloc( count * sizeof( T ) ) );
   zeroDevice( c, p, count * sizeof( T ) );
   c->bytes += int64_t( count * sizeof( T ) );
}
template < class T >
inline void DevBuf< T >::upload( Context* c, const std::vector< T >& h )
{
   alloc( c, h.size() );
   if ( !h.empty() )
   {
      // a copy from pageable memory and the wait for it must not run while
      // another thread captures (work lists are built lazily by serving workers)
      DeviceGateLock gate;
      HG_CHECK_HIP( hipMemcpyAsync( p, h.data(), h.size() * sizeof( T ), hipMemcpyHostToDevice, c->stream ) );
      HG_CHECK_HIP( hipStreamSynchronize( c->stream ) );
   }
}
template < class T >
inline void DevBuf< T >::release()
{
   if ( p )
   {
      if ( ctx )
      {
         ctx->deviceFree( p, n * sizeof( T ) );
         ctx->bytes -= int64_t( n * sizeof( T ) );
      }
      else
      {
         DeviceGateLock gate;
         (void) hipFree( p );
      }
   }
   p = nullptr;
   n = 0;
}

// ---- device helpers used across translation units ----------------------
constexpr int kBlock      = 256;
constexpr int kDotBlocks  = 1024;
constexpr int kMaxLv      = 6; // at most 64 interleaved vectors per function
constexpr int kSlots      = 256; // device scalar slots per context

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2): blockIdx b runs on XCD b % 8.  xcdBlock maps the hardware block id to a
// logical one so that each XCD receives a contiguous range of logical blocks
// (neighbouring rows / tiles share L2 lines); the grid is padded to a multiple
// of 8 by xcdGrid and logical ids >= n must return.
__device__ __forceinline__ int64_t xcdBlock( int64_t n )
{
   const int64_t per = ( n + 7 ) / 8;
   return int64_t( blockIdx.x % 8 ) * per + int64_t( blockIdx.x / 8 );
}
inline unsigned xcdGrid( int64_t n )
{
   const int64_t g = 8 * ( ( n + 7 ) / 8 );
   if ( g > 2147483647LL )
      throw Error( -1, "grid too large" );
   return unsigned( g );
}

inline unsigned gridFor( int64_t n, int block = kBlock )
{
   int64_t g = ( n + block - 1 ) / block;
   if ( g < 1 )
      g = 1;
   if ( g > 2147483647LL )
      throw Error( -1, "grid too large" );
   return unsigned( g );
}

// Refresh every halo copy in the cell lattices from its owner (the
// Vertex->Edge->Face->Cell communicate chain of VertexDoFFunction).
// lv: log2 of the number of vectors interleaved per lattice point (0: one vector)
void pullHalos( Context* ctx, Storage& st, Lattice& lat, void* data, int dtype, int lv = 0 );
// dst[i] = src[i] converted between value types (HG_F64 / HG_F32 / HG_F16)
void castCopy( Context* ctx, const void* src, int sdtype, void* dst, int ddtype, int64_t n );
// Sum the halo copies of every owned vertex/edge/face DoF into the owner
// (communicateAdditively<Cell, Face/Edge/Vertex>).  Owners whose type matches
// excludeFlag are left untouched; zeroOut replaces instead of adding.
void pushAdditive( Context* ctx, Storage& st, Lattice& lat, void* data, int dtype, int excludeFlag, bool zeroOut,
                   int lv = 0 );
// f = beta f on the DoFs of type flag (assign({beta}, {f}, level, flag))
void scaleFlagged( Context* ctx, Function& f, int level, int flag, double beta );
// Jacobi step of smooth_jac (P1Operator.hpp:432-446, P2ElementwiseOperator.cpp:
// 342-360) after dst = A src: dst = src + relax * invDiag * (rhs - dst) on the
// DoFs of type flag (cell-lattice positions count as Inner), each intermediate
// rounded to the storage type as the assign / multElementwise / assign
// sequence of the reference rounds it.  invDiag: FP64, one value per storage
// position of one vector (shared by the 2^lv interleaved vectors).
void jacobiUpdate( Context* ctx, Lattice& lat, int dtype, int lv, int flag, double relax, const void* rhs,
                   const void* src, const double* invDiag, void* dst );
// FP64 dot over owned DoFs of matching type
double dotDevice( Context* ctx, Lattice& lat, const void* a, const void* b, int dtype, int flag );
// per-vector dot products of two functions with 2^lv interleaved vectors
void dotDeviceBatched( Context* ctx, Lattice& lat, const void* a, const void* b, int dtype, int flag, int lv, double* out );
// storage type dispatch
template < class F >
void dispatchType( int dtype, F&& f )
{
   switch ( dtype )
   {
   case 0:
      f( double() );
      break;
   case 1:
      f( float() );
      break;
   case 2:
      f( __half() );
      break;
   default:
      throw Error( -1, "unknown storage type" );
   }
}
// arithmetic type used for a storage type: FP64 for FP64, FP32 otherwise
template < class T >
struct Acc
{
   using type = float;
};
template <>
struct Acc< double >
{
   using type = double;
};
template < class T >
__host__ __device__ __forceinline__ typename Acc< T >::type toAcc( T v )
{
   return static_cast< typename Acc< T >::type >( v );
}
template <>
__host__ __device__ __forceinline__ float toAcc< __half >( __half v )
{
   return __half2float( v );
}
template < class T >
__host__ __device__ __forceinline__ T fromAcc( typename Acc< T >::type v )
{
   return static_cast< T >( v );
}
template <>
__host__ __device__ __forceinline__ __half fromAcc< __half >( float v )
{
   return __float2half( v );
}

} // namespace hg
