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

// ---------------------------------------------------------------------------
// icosahedral spherical shell (mesh/MeshGenSphericalShell.cpp:38-700)
// ---------------------------------------------------------------------------
#include <cmath>

namespace hg {
namespace {

using Vec3 = std::array< double, 3 >;

struct UnitSphereDiamonds
{
   int                  ntan;
   Vec3                 ico[12];
   int                  diamond[10][4];
   std::vector< Vec3 >  node; // [is1][is2][id]
   Vec3&                at( int i1, int i2, int id ) { return node[( size_t( i1 ) * ntan + i2 ) * 10 + id]; }

   explicit UnitSphereDiamonds( int n )
   : ntan( n )
   , node( size_t( n ) * n * 10 )
   {
      const double fifthpi = 0.4 * std::asin( 1.0 );
      const double w       = 2.0 * std::acos( 1.0 / ( 2.0 * std::sin( fifthpi ) ) );
      ico[0]               = { 0.0, 0.0, 1.0 };
      ico[11]              = { 0.0, 0.0, -1.0 };
      for ( int k = 1; k <= 5; ++k )
      {
         const double a = 2.0 * ( k - 0.5 ) * fifthpi, b = 2.0 * ( k - 1 ) * fifthpi;
         ico[k]         = { std::sin( w ) * std::cos( a ), std::sin( w ) * std::sin( a ), std::cos( w ) };
         ico[k + 5]     = { std::sin( w ) * std::cos( b ), std::sin( w ) * std::sin( b ), -std::cos( w ) };
      }
      const int d[10][4] = { { 0, 5, 6, 1 },  { 0, 1, 7, 2 },  { 0, 2, 8, 3 },  { 0, 3, 9, 4 },  { 0, 4, 10, 5 },
                             { 11, 7, 1, 6 }, { 11, 8, 2, 7 }, { 11, 9, 3, 8 }, { 11, 10, 4, 9 }, { 11, 6, 5, 10 } };
      std::copy( &d[0][0], &d[0][0] + 40, &diamond[0][0] );
      for ( int id = 0; id < 10; ++id )
      {
         const int left = id < 5 ? 1 : 3, right = id < 5 ? 3 : 1;
         at( 0, 0, id )               = ico[diamond[id][0]];
         at( n - 1, 0, id )           = ico[diamond[id][left]];
         at( n - 1, n - 1, id )       = ico[diamond[id][2]];
         at( 0, n - 1, id )           = ico[diamond[id][right]];
         // recursive midpoint refinement: rows, columns, diagonals per step
         for ( int m = 1; m < n - 1; m *= 2 )
         {
            const int l = ( n - 1 ) / m, h = l / 2;
            auto      mid = [&]( int i1, int i2, int a1, int a2, int b1, int b2 ) {
               const Vec3& A = at( a1, a2, id );
               const Vec3& B = at( b1, b2, id );
               Vec3        s{ A[0] + B[0], A[1] + B[1], A[2] + B[2] };
               const double r = std::sqrt( s[0] * s[0] + s[1] * s[1] + s[2] * s[2] );
               at( i1, i2, id ) = { s[0] / r, s[1] / r, s[2] / r };
            };
            for ( int j1 = 0; j1 <= m; ++j1 )
               for ( int j2 = 0; j2 < m; ++j2 )
                  mid( j1 * l, j2 * l + h, j1 * l, j2 * l, j1 * l, j2 * l + l );
            for ( int j1 = 0; j1 <= m; ++j1 )
               for ( int j2 = 0; j2 < m; ++j2 )
                  mid( j2 * l + h, j1 * l, j2 * l, j1 * l, j2 * l + l, j1 * l );
            for ( int j1 = 0; j1 < m; ++j1 )
               for ( int j2 = 0; j2 < m; ++j2 )
                  mid( j1 * l + h, j2 * l + h, j1 * l, j2 * l + l, j1 * l + l, j2 * l );
         }
      }
   }
};

// global vertex index of the tuple (is1, is2, diamond, layer), folding the
// diamond's left edges onto their owning neighbours
int shellVertexIndex( int ntan, int nrad, int is1, int is2, int id, int ir )
{
   if ( is1 == 0 && is2 == 0 )
      return id < 5 ? ir : nrad + ir;
   if ( is2 == 0 )
   {
      is2 = is1;
      is1 = 0;
      id  = id < 5 ? ( id + 4 ) % 5 : ( id + 4 ) % 5 + 5;
   }
   else if ( is1 == ntan - 1 )
   {
      is1 = ( ntan - 1 ) - is2;
      is2 = ntan - 1;
      id  = id < 5 ? ( id + 4 ) % 5 + 5 : id % 5;
   }
   const int q = ( ntan - 1 ) * ( ntan - 1 );
   return 2 * nrad + ir * q * 10 + id * q + ( is2 - 1 ) * ( ntan - 1 ) + is1;
}

} // namespace

void Topology::setShellBoundaryFlags( double rmid, int innerFlag, int outerFlag )
{
   for ( int f = 0; f < nf; ++f )
   {
      if ( fnc[f] != 1 )
         continue;
      const auto v = fv[f];
      double     r = 0.0;
      for ( int j = 0; j < 3; ++j )
         r += std::sqrt( X[3 * v[j]] * X[3 * v[j]] + X[3 * v[j] + 1] * X[3 * v[j] + 1] + X[3 * v[j] + 2] * X[3 * v[j] + 2] ) /
              3.0;
      const int flag = r < rmid ? innerFlag : outerFlag;
      fflag[f]       = flag;
      for ( int j = 0; j < 3; ++j )
         vflag[v[j]] = flag;
      eflag[findEdge( v[0], v[1] )] = flag;
      eflag[findEdge( v[0], v[2] )] = flag;
      eflag[findEdge( v[1], v[2] )] = flag;
   }
}

Topology Topology::shell( int ntan, int nrad, double rmin, double rmax )
{
   if ( ntan < 2 || nrad < 2 || !( rmin > 0.0 && rmin < rmax ) )
      throw std::invalid_argument( "invalid shell parameters" );
   if ( ( ( ntan - 1 ) & ( ntan - 2 ) ) != 0 )
      throw std::invalid_argument( "ntan - 1 must be a power of two" );
   UnitSphereDiamonds  sph( ntan );
   const int           q  = ( ntan - 1 ) * ( ntan - 1 );
   const int           nv = ( 10 * q + 2 ) * nrad;
   std::vector< double > X( 3 * size_t( nv ) );
   for ( int j = 0; j < nv; ++j )
   {
      int is1 = 0, is2 = 0, id, ir;
      if ( j < nrad )
         id = 0, ir = j;
      else if ( j < 2 * nrad )
         id = 5, ir = j - nrad;
      else
      {
         int r = j - 2 * nrad;
         ir    = r / ( 10 * q );
         r -= ir * 10 * q;
         id = r / q;
         r -= id * q;
         is2 = r / ( ntan - 1 ) + 1;
         is1 = r % ( ntan - 1 );
      }
      const double radius = rmin + ( ( rmax - rmin ) / ( nrad - 1 ) ) * ir;
      const Vec3&  p      = sph.at( is1, is2, id );
      for ( int d = 0; d < 3; ++d )
         X[3 * j + d] = p[d] * radius;
   }
   // six tetrahedra per radial prism cell, node tables for both hemispheres
   constexpr int corner[8][3] = { { 0, 0, 1 }, { 1, 0, 1 }, { 1, 1, 1 }, { 0, 1, 1 },
                                  { 0, 0, 0 }, { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 } };
   constexpr int north[6][4]  = { { 0, 1, 3, 7 }, { 4, 7, 5, 0 }, { 0, 5, 1, 7 }, { 1, 2, 3, 6 }, { 5, 7, 6, 1 }, { 3, 6, 7, 1 } };
   constexpr int south[6][4]  = { { 4, 7, 5, 1 }, { 0, 1, 3, 4 }, { 4, 3, 7, 1 }, { 7, 6, 5, 2 }, { 3, 1, 2, 7 }, { 5, 2, 1, 7 } };
   std::vector< int64_t > cells;
   const int              ncell = q * ( nrad - 1 ) * 60;
   cells.reserve( 4 * size_t( ncell ) );
   for ( int e = 0; e < ncell; ++e )
   {
      int r        = e;
      const int ir = r / ( 60 * q );
      r -= ir * 60 * q;
      const int id = r / ( 6 * q );
      r -= id * 6 * q;
      const int is2 = r / ( 6 * ( ntan - 1 ) );
      r -= is2 * 6 * ( ntan - 1 );
      const int is1 = r / 6, it = r % 6;
      int       v8[8];
      for ( int k = 0; k < 8; ++k )
      {
         const int a = id <= 4 ? corner[k][0] : corner[k][1];
         const int b = id <= 4 ? corner[k][1] : corner[k][0];
         v8[k]       = shellVertexIndex( ntan, nrad, is1 + a, is2 + b, id, ir + corner[k][2] );
      }
      for ( int k = 0; k < 4; ++k )
         cells.push_back( v8[id <= 4 ? north[it][k] : south[it][k]] );
   }
   Topology t = fromCells( nv, X.data(), ncell, cells.data() );
   t.computeShellBlend();
   return t;
}

// IcosahedralShellMap::classifyVertices (geometry/IcosahedralShellMap.hpp:529-780):
// the two vertices on one radial ray, ordered by tet type; a "skew" tet (two
// vertices on each sphere) takes the map of the tet across its first face that
// does not contain both radial vertices.
void Topology::computeShellBlend()
{
   blend.assign( nc, ShellCellMap{} );
   auto classify = [&]( int c, ShellCellMap& out, auto& self, int depth ) -> void {
      const auto& cvv = cv[c];
      double      P[4][3], r[4], rin = 1e300, rout = 0.0;
      for ( int k = 0; k < 4; ++k )
      {
         for ( int d = 0; d < 3; ++d )
            P[k][d] = X[3 * cvv[k] + d];
         r[k] = std::sqrt( P[k][0] * P[k][0] + P[k][1] * P[k][1] + P[k][2] * P[k][2] );
         rin  = std::min( rin, r[k] );
         rout = std::max( rout, r[k] );
      }
      int outer = 0;
      for ( int k = 0; k < 4; ++k )
         outer += ( rout - r[k] ) < ( rout - rin ) * 0.001 ? 1 : 0;
      int ref = -1, ray = -1;
      for ( int k = 0; k < 4 && ref < 0; ++k )
         for ( int j = k + 1; j < 4 && ref < 0; ++j )
         {
            const double s = r[k] * r[j];
            const double cr[3] = { P[k][1] * P[j][2] - P[k][2] * P[j][1], P[k][2] * P[j][0] - P[k][0] * P[j][2],
                                   P[k][0] * P[j][1] - P[k][1] * P[j][0] };
            if ( std::fabs( cr[0] ) / s > 1e-14 || std::fabs( cr[1] ) / s > 1e-14 || std::fabs( cr[2] ) / s > 1e-14 )
               continue;
            ref = k;
            ray = j;
         }
      if ( ref < 0 )
         throw std::runtime_error( "shell map: no radial vertex pair in cell" );
      if ( outer == 2 )
      {
         if ( depth > 0 )
            throw std::runtime_error( "shell map: skew neighbour of a skew cell" );
         for ( int k = 0; k < 4; ++k )
         {
            const auto& f = fv[cf[c][k]];
            const bool  hasRef = std::count( f.begin(), f.end(), cvv[ref] ) > 0;
            const bool  hasRay = std::count( f.begin(), f.end(), cvv[ray] ) > 0;
            if ( hasRef && hasRay )
               continue;
            for ( int o = 0; o < nc; ++o )
               if ( o != c && std::count( cf[o].begin(), cf[o].end(), cf[c][k] ) > 0 )
               {
                  self( o, out, self, depth + 1 );
                  return;
               }
         }
         throw std::runtime_error( "shell map: no non-skew neighbour" );
      }
      int thr = -1, fo = -1;
      for ( int k = 0; k < 4; ++k )
         if ( k != ref && k != ray )
            thr = k;
      for ( int k = 0; k < 4; ++k )
         if ( k != ref && k != ray && k != thr )
            fo = k;
      if ( ( outer == 1 && r[ref] < r[ray] ) || ( outer == 3 && r[ref] > r[ray] ) )
         std::swap( ref, ray );
      out.kind = 1;
      for ( int d = 0; d < 3; ++d )
      {
         out.rayVertex[d] = P[ray][d];
         out.refVertex[d] = P[ref][d];
         out.thrVertex[d] = P[thr][d];
         out.forVertex[d] = P[fo][d];
      }
      out.radRayVertex = r[ray];
      out.radRefVertex = r[ref];
   };
   for ( int c = 0; c < nc; ++c )
      classify( c, blend[c], classify, 0 );
}

} // namespace hg
