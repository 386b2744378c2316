// Macro-mesh topology on the host: the primitives HyTeG's SetupPrimitiveStorage
// derives from a MeshInfo (primitivestorage/SetupPrimitiveStorage.cpp:98-380),
// in the same order (PrimitiveIDs are handed out vertices -> edges -> faces ->
// cells, each in MeshInfo map order), with the boundary flags of
// setMeshBoundaryFlagsOnBoundary(1, 0, true) mapped through create0123BC
// (boundary/BoundaryConditions.cpp:26-33).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace hg {

enum DoFType : int
{
   kInner     = 1,
   kDirichlet = 2,
   kNeumann   = 4,
   kFreeslip  = 8,
   kAll       = 15
};

// local edges / faces of a macro-cell in terms of its local vertices
// (SetupPrimitiveStorage.cpp:290-300)
constexpr int kCellEdge[6][2] = { { 0, 1 }, { 0, 2 }, { 1, 2 }, { 0, 3 }, { 1, 3 }, { 2, 3 } };
constexpr int kCellFace[4][3] = { { 0, 1, 2 }, { 0, 1, 3 }, { 0, 2, 3 }, { 1, 2, 3 } };

// Blending of the icosahedral shell (geometry/IcosahedralShellMap.hpp): per
// macro-cell data; unused (affine) cells carry kind 0.
struct ShellCellMap
{
   int    kind = 0;
   double rayVertex[3], refVertex[3], thrVertex[3], forVertex[3];
   double radRefVertex, radRayVertex, prismNormal[3];
};

struct Topology
{
   int nv = 0, ne = 0, nf = 0, nc = 0;

   std::vector< double >                 X;  // nv*3
   std::vector< std::array< int, 2 > >   ev; // sorted vertex pairs
   std::vector< std::array< int, 3 > >   fv; // sorted vertex triples
   std::vector< std::array< int, 4 > >   cv; // cell-local vertex order
   std::vector< std::array< int, 6 > >   ce;
   std::vector< std::array< int, 4 > >   cf;
   std::vector< int >                    vflag, eflag, fflag;
   std::vector< int >                    vnc, enc, fnc; // neighbour cell counts
   std::vector< ShellCellMap >           blend;         // empty: affine mesh

   static Topology fromCells( int nv, const double* coords, int nc, const int64_t* cells );
   static Topology cuboid( const double lo[3], const double hi[3], int nx, int ny, int nz );
   // MeshInfo::meshSphericalShell(ntan, nrad, rmin, rmax) (SHELLMESH_CLASSIC) and
   // IcosahedralShellMap::setMap on every cell
   static Topology shell( int ntan, int nrad, double rmin, double rmax );
   // boundary flags of a shell by surface: primitives of boundary faces nearer
   // the inner radius get innerFlag, the others outerFlag (the mesh boundary
   // flags HyTeG applications set with setMeshBoundaryFlagsInner/Outer)
   void setShellBoundaryFlags( double rmid, int innerFlag, int outerFlag );
   void            computeShellBlend();

   int findEdge( int a, int b ) const;
   int findFace( int a, int b, int c ) const;
};

// Canonical numbering of the P1 DoFs of one level:
// [ vertices | edge interiors | face interiors | cell interiors ], each block in
// primitive order and, inside a primitive, lexicographic in its own lattice.
struct LevelNumbering
{
   int     level = 0;
   int64_t mm = 1, ein = 0, fin = 0, cin = 0;
   int64_t eoff = 0, foff = 0, coff = 0, n = 0;

   LevelNumbering() = default;
   LevelNumbering( const Topology& t, int L );

   // owner-canonical id of lattice point (x,y,z) of macro-cell c
   int64_t locate( const Topology& t, int c, int64_t x, int64_t y, int64_t z ) const;
   // kind 0..3 = vertex/edge/face/cell and primitive index of a canonical id
   int     owner( int64_t id, int64_t* prim ) const;
   int     flagOf( const Topology& t, int64_t id ) const;
   int     numNeighbourCells( const Topology& t, int64_t id ) const;
};

} // namespace hg
