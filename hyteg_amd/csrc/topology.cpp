#include "topology.hpp"

#include <algorithm>
#include <map>
#include <stdexcept>

#include "lattice.hpp"

namespace hg {

Topology Topology::fromCells( int nv, const double* coords, int nc, const int64_t* cells )
{
   Topology t;
   t.nv = nv;
   t.nc = nc;
   t.X.assign( coords, coords + 3 * nv );
   t.cv.resize( nc );
   for ( int c = 0; c < nc; ++c )
      for ( int k = 0; k < 4; ++k )
      {
         const int64_t v = cells[4 * c + k];
         if ( v < 0 || v >= nv )
            throw std::invalid_argument( "cell references a vertex id out of range" );
         t.cv[c][k] = int( v );
      }
   // MeshInfo keys cells_ by their vertex vector (MeshInfo.cpp:74-84)
   std::sort( t.cv.begin(), t.cv.end() );
   if ( std::adjacent_find( t.cv.begin(), t.cv.end() ) != t.cv.end() )
      throw std::invalid_argument( "duplicate cell" );

   std::map< std::array< int, 2 >, int > edges;
   std::map< std::array< int, 3 >, int > faces;
   for ( const auto& c : t.cv )
   {
      for ( const auto& le : kCellEdge )
      {
         std::array< int, 2 > e{ c[le[0]], c[le[1]] };
         std::sort( e.begin(), e.end() );
         edges.emplace( e, 0 );
      }
      for ( const auto& lf : kCellFace )
      {
         std::array< int, 3 > f{ c[lf[0]], c[lf[1]], c[lf[2]] };
         std::sort( f.begin(), f.end() );
         faces.emplace( f, 0 );
      }
   }
   for ( auto& [k, v] : edges )
   {
      v = int( t.ev.size() );
      t.ev.push_back( k );
   }
   for ( auto& [k, v] : faces )
   {
      v = int( t.fv.size() );
      t.fv.push_back( k );
   }
   t.ne = int( t.ev.size() );
   t.nf = int( t.fv.size() );

   t.ce.resize( nc );
   t.cf.resize( nc );
   t.vnc.assign( nv, 0 );
   t.enc.assign( t.ne, 0 );
   t.fnc.assign( t.nf, 0 );
   for ( int c = 0; c < nc; ++c )
   {
      for ( int k = 0; k < 6; ++k )
      {
         t.ce[c][k] = t.findEdge( t.cv[c][kCellEdge[k][0]], t.cv[c][kCellEdge[k][1]] );
         t.enc[t.ce[c][k]]++;
      }
      for ( int k = 0; k < 4; ++k )
      {
         t.cf[c][k] = t.findFace( t.cv[c][kCellFace[k][0]], t.cv[c][kCellFace[k][1]], t.cv[c][kCellFace[k][2]] );
         t.fnc[t.cf[c][k]]++;
      }
      for ( int k = 0; k < 4; ++k )
         t.vnc[t.cv[c][k]]++;
   }
   // SetupPrimitiveStorage::onBoundary (3D): a face with one neighbour cell
   // is on the boundary, lower-dimensional primitives are when one of their
   // faces is (SetupPrimitiveStorage.cpp:1092-1160).
   t.vflag.assign( nv, kInner );
   t.eflag.assign( t.ne, kInner );
   t.fflag.assign( t.nf, kInner );
   for ( int f = 0; f < t.nf; ++f )
   {
      if ( t.fnc[f] != 1 )
         continue;
      t.fflag[f]   = kDirichlet;
      const auto v = t.fv[f];
      for ( int j = 0; j < 3; ++j )
         t.vflag[v[j]] = kDirichlet;
      t.eflag[t.findEdge( v[0], v[1] )] = kDirichlet;
      t.eflag[t.findEdge( v[0], v[2] )] = kDirichlet;
      t.eflag[t.findEdge( v[1], v[2] )] = kDirichlet;
   }
   return t;
}

Topology Topology::cuboid( const double lo[3], const double hi[3], int nx, int ny, int nz )
{
   if ( nx < 1 || ny < 1 || nz < 1 || !( lo[0] < hi[0] && lo[1] < hi[1] && lo[2] < hi[2] ) )
      throw std::invalid_argument( "invalid cuboid" );
   // vertex ids z-y-x ordered and 6 tets per hexahedron with the node table of
   // MeshGenCuboid.cpp:86-101
   const int           nv = ( nx + 1 ) * ( ny + 1 ) * ( nz + 1 );
   std::vector< double > X( 3 * size_t( nv ) );
   const double        h[3] = { ( hi[0] - lo[0] ) / nx, ( hi[1] - lo[1] ) / ny, ( hi[2] - lo[2] ) / nz };
   int                 id   = 0;
   for ( int iz = 0; iz <= nz; ++iz )
      for ( int iy = 0; iy <= ny; ++iy )
         for ( int ix = 0; ix <= nx; ++ix, ++id )
         {
            X[3 * id + 0] = lo[0] + ix * h[0];
            X[3 * id + 1] = lo[1] + iy * h[1];
            X[3 * id + 2] = lo[2] + iz * h[2];
         }
   constexpr int hexTet[6][4]  = { { 4, 5, 7, 3 }, { 0, 3, 1, 4 }, { 4, 1, 5, 3 }, { 5, 6, 7, 2 }, { 1, 3, 2, 5 }, { 7, 2, 3, 5 } };
   constexpr int hexNode[8][3] = { { 0, 0, 0 }, { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 },
                                   { 0, 0, 1 }, { 1, 0, 1 }, { 1, 1, 1 }, { 0, 1, 1 } };
   std::vector< int64_t > cells;
   for ( int iz = 0; iz < nz; ++iz )
      for ( int iy = 0; iy < ny; ++iy )
         for ( int ix = 0; ix < nx; ++ix )
            for ( const auto& tt : hexTet )
               for ( int k : tt )
               {
                  const int* o = hexNode[k];
                  cells.push_back( int64_t( iz + o[2] ) * ( nx + 1 ) * ( ny + 1 ) + int64_t( iy + o[1] ) * ( nx + 1 ) +
                                   ( ix + o[0] ) );
               }
   return fromCells( nv, X.data(), int( cells.size() / 4 ), cells.data() );
}

int Topology::findEdge( int a, int b ) const
{
   std::array< int, 2 > k{ std::min( a, b ), std::max( a, b ) };
   auto                 it = std::lower_bound( ev.begin(), ev.end(), k );
   if ( it == ev.end() || *it != k )
      throw std::logic_error( "edge not found" );
   return int( it - ev.begin() );
}

int Topology::findFace( int a, int b, int c ) const
{
   std::array< int, 3 > k{ a, b, c };
   std::sort( k.begin(), k.end() );
   auto it = std::lower_bound( fv.begin(), fv.end(), k );
   if ( it == fv.end() || *it != k )
      throw std::logic_error( "face not found" );
   return int( it - fv.begin() );
}

LevelNumbering::LevelNumbering( const Topology& t, int L )
: level( L )
{
   mm   = int64_t( 1 ) << L;
   ein  = mm - 1;
   fin  = tri( mm - 2 );
   cin  = tet( mm - 3 );
   eoff = t.nv;
   foff = eoff + int64_t( t.ne ) * ein;
   coff = foff + int64_t( t.nf ) * fin;
   n    = coff + int64_t( t.nc ) * cin;
}

int64_t LevelNumbering::locate( const Topology& t, int c, int64_t x, int64_t y, int64_t z ) const
{
   // barycentric lattice weights on the cell's local vertices: the owner is
   // the primitive spanned by the vertices with non-zero weight, the local
   // coordinates are the weights of that primitive's 2nd/3rd vertex.
   const int64_t w[4] = { mm - x - y - z, x, y, z };
   int           S[4], ns = 0;
   for ( int k = 0; k < 4; ++k )
      if ( w[k] != 0 )
         S[ns++] = k;
   const auto& cv       = t.cv[c];
   auto        weightOf = [&]( int g ) {
      for ( int k = 0; k < 4; ++k )
         if ( cv[k] == g )
            return w[k];
      return int64_t( -1 );
   };
   if ( ns == 1 )
      return cv[S[0]];
   if ( ns == 2 )
   {
      int e = -1;
      for ( int k = 0; k < 6; ++k )
         if ( kCellEdge[k][0] == S[0] && kCellEdge[k][1] == S[1] )
            e = t.ce[c][k];
      return eoff + e * ein + weightOf( t.ev[e][1] ) - 1;
   }
   if ( ns == 3 )
   {
      int f = -1;
      for ( int k = 0; k < 4; ++k )
         if ( kCellFace[k][0] == S[0] && kCellFace[k][1] == S[1] && kCellFace[k][2] == S[2] )
            f = t.cf[c][k];
      const int64_t fx = weightOf( t.fv[f][1] ), fy = weightOf( t.fv[f][2] );
      return foff + f * fin + triIndex( mm - 2, fx - 1, fy - 1 );
   }
   return coff + c * cin + tetIndex( mm - 3, x - 1, y - 1, z - 1 );
}

int LevelNumbering::owner( int64_t id, int64_t* prim ) const
{
   if ( id < eoff )
   {
      *prim = id;
      return 0;
   }
   if ( id < foff )
   {
      *prim = ( id - eoff ) / ein;
      return 1;
   }
   if ( id < coff )
   {
      *prim = ( id - foff ) / fin;
      return 2;
   }
   *prim = ( id - coff ) / cin;
   return 3;
}

int LevelNumbering::flagOf( const Topology& t, int64_t id ) const
{
   int64_t p;
   switch ( owner( id, &p ) )
   {
   case 0:
      return t.vflag[p];
   case 1:
      return t.eflag[p];
   case 2:
      return t.fflag[p];
   default:
      return kInner;
   }
}

int LevelNumbering::numNeighbourCells( const Topology& t, int64_t id ) const
{
   int64_t p;
   switch ( owner( id, &p ) )
   {
   case 0:
      return t.vnc[p];
   case 1:
      return t.enc[p];
   case 2:
      return t.fnc[p];
   default:
      return 1;
   }
}

} // namespace hg
