// Structured tetrahedral lattice helpers, shared by host set-up code and HIP
// kernels.  A macro-cell on refinement level L carries a lattice of width
// w = 2^L + 1 points per macro-edge (Levelinfo.hpp:36-40); its points (x,y,z),
// x,y,z >= 0, x+y+z <= w-1, are stored x-fastest, then rows y, then slices z
// (indexing/MacroCellIndexing.hpp:40-53).
#pragma once

#include <cstdint>

#if defined( __HIPCC__ )
#define HG_HD __host__ __device__ __forceinline__
#else
#define HG_HD inline
#endif

namespace hg {

HG_HD int64_t tri( int64_t n )
{
   return n <= 0 ? 0 : n * ( n + 1 ) / 2;
}
HG_HD int64_t tet( int64_t n )
{
   return n <= 0 ? 0 : n * ( n + 1 ) * ( n + 2 ) / 6;
}
// index of (x,y) in a triangle of width n
HG_HD int64_t triIndex( int64_t n, int64_t x, int64_t y )
{
   return y * n - y * ( y - 1 ) / 2 + x;
}
// index of (x,y,z) in a tetrahedral lattice of width n
HG_HD int64_t tetIndex( int64_t n, int64_t x, int64_t y, int64_t z )
{
   return tet( n ) - tet( n - z ) + y * ( n - z ) - y * ( y - 1 ) / 2 + x;
}
// 32-bit variant of tetRowStart for in-lattice offsets of lattices with
// n <= 1024 (tet(n) < 2^31 and all intermediate products < 2^31)
HG_HD int32_t tetRowStart32( int32_t n, int32_t y, int32_t z )
{
   const int32_t m = n - z; // tet(n) - tet(n - z) = z (3 n^2 - 3 n z + 6 n + z^2 - 3 z + 2) / 6
   return int32_t( ( int64_t( z ) * ( 3 * n * n - 3 * n * z + 6 * n + z * z - 3 * z + 2 ) ) / 6 ) + y * m -
          y * ( y - 1 ) / 2;
}
// start of row (y, z) in a tetrahedral lattice of width n (x = 0)
HG_HD int64_t tetRowStart( int64_t n, int64_t y, int64_t z )
{
   return tet( n ) - tet( n - z ) + y * ( n - z ) - y * ( y - 1 ) / 2;
}

// Vertex offsets of the 6 micro-cell types WHITE_UP, BLUE_UP, GREEN_UP,
// WHITE_DOWN, BLUE_DOWN, GREEN_DOWN (volumedofspace/CellDoFIndexing.hpp:150-200)
// and the number of micro-cells per row of each type relative to 2^L
// (CellDoFIndexing.hpp:64-83).
struct MicroCells
{
   static constexpr int vof[6][4][3] = {
       { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } },
       { { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 }, { 1, 0, 1 } },
       { { 1, 0, 0 }, { 0, 1, 0 }, { 1, 0, 1 }, { 0, 0, 1 } },
       { { 1, 1, 0 }, { 1, 1, 1 }, { 0, 1, 1 }, { 1, 0, 1 } },
       { { 1, 0, 1 }, { 0, 1, 1 }, { 0, 0, 1 }, { 0, 1, 0 } },
       { { 0, 1, 0 }, { 1, 1, 0 }, { 1, 0, 1 }, { 0, 1, 1 } },
   };
   static constexpr int rowShift[6] = { 0, 1, 1, 2, 1, 1 };
};

// The 15 neighbour offsets of the P1 vertex stencil, centre first
// (apply_3D_macrocell_vertexdof_to_vertexdof_replace.cpp:36-50 lists the same set).
struct Stencil15
{
   static constexpr int off[15][3] = { { 0, 0, 0 },  { -1, 0, 0 }, { -1, 0, 1 }, { -1, 1, -1 }, { -1, 1, 0 },
                                       { 0, -1, 0 }, { 0, -1, 1 }, { 0, 0, -1 }, { 0, 0, 1 },   { 0, 1, -1 },
                                       { 0, 1, 0 },  { 1, -1, 0 }, { 1, -1, 1 }, { 1, 0, -1 },  { 1, 0, 0 } };
};

} // namespace hg
