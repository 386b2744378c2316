// VTK output of device functions (VTKOutput with VTKP1Writer / VTKP2Writer,
// dataexport/VTKOutput/VTKOutput.cpp, VTKP1Writer.cpp:34-120,
// VTKP2Writer.cpp:34-160, VTKMeshWriter.cpp:41-1070): one XML unstructured-grid
// file (.vtu) per call.  As in the reference, every macro-cell writes its own
// copy of its lattice points (points on shared primitives appear once per
// cell), the points are mapped by the storage's blending (the icosahedral shell
// map), P1 functions are written on the level-L micro-tetrahedra (VTK_TETRA)
// and P2 functions on the level-L micro-tetrahedra as quadratic tetrahedra
// (VTK_QUADRATIC_TETRA, vertex and edge DoFs).  Host code: the values are
// downloaded once per function in the canonical order and placed per cell.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "abi_util.hpp"
#include "hyteg_amd.h"
#include "lattice.hpp"
#include "runtime.hpp"

namespace hg {
namespace {

// point of the level lattice (x,y,z) of macro-cell c, mapped by the blending
void latticePoint( const Topology& t, int c, int64_t mm, int64_t x, int64_t y, int64_t z, double out[3] )
{
   const double* v[4];
   for ( int k = 0; k < 4; ++k )
      v[k] = &t.X[3 * t.cv[c][k]];
   const double h = 1.0 / double( mm );
   for ( int d = 0; d < 3; ++d )
      out[d] = v[0][d] + h * ( double( x ) * ( v[1][d] - v[0][d] ) + double( y ) * ( v[2][d] - v[0][d] ) +
                               double( z ) * ( v[3][d] - v[0][d] ) );
   if ( t.blend.empty() )
      return;
   // IcosahedralShellMap::evalF (IcosahedralShellMap.hpp:109-146): the point moves
   // along its ray to the radius interpolated between the ray and reference vertex
   const ShellCellMap& b = t.blend[c];
   double              a[3], tt[3], f[3], xr[3];
   for ( int d = 0; d < 3; ++d )
   {
      a[d]  = b.refVertex[d] - b.rayVertex[d];
      tt[d] = b.thrVertex[d] - b.rayVertex[d];
      f[d]  = b.forVertex[d] - b.rayVertex[d];
      xr[d] = out[d] - b.rayVertex[d];
   }
   const double n[3]   = { tt[1] * f[2] - tt[2] * f[1], tt[2] * f[0] - tt[0] * f[2], tt[0] * f[1] - tt[1] * f[0] };
   const double volT   = a[0] * n[0] + a[1] * n[1] + a[2] * n[2];
   const double volX   = xr[0] * n[0] + xr[1] * n[1] + xr[2] * n[2];
   const double bary   = std::fabs( volX / volT );
   const double oldRad = std::sqrt( out[0] * out[0] + out[1] * out[1] + out[2] * out[2] );
   const double newRad = b.radRayVertex + bary * ( b.radRefVertex - b.radRayVertex );
   for ( int d = 0; d < 3; ++d )
      out[d] *= newRad / oldRad;
}

const char kB64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

std::string base64( const unsigned char* p, size_t n )
{
   std::string out;
   out.reserve( 4 * ( n + 2 ) / 3 );
   for ( size_t i = 0; i < n; i += 3 )
   {
      const uint32_t b0 = p[i], b1 = i + 1 < n ? p[i + 1] : 0, b2 = i + 2 < n ? p[i + 2] : 0;
      const uint32_t w  = ( b0 << 16 ) | ( b1 << 8 ) | b2;
      out += kB64[( w >> 18 ) & 63];
      out += kB64[( w >> 12 ) & 63];
      out += i + 1 < n ? kB64[( w >> 6 ) & 63] : '=';
      out += i + 2 < n ? kB64[w & 63] : '=';
   }
   return out;
}

// one <DataArray>: ASCII, or VTK's inline binary (base64 of a UInt32 byte count
// followed by the raw little-endian values)
template < class T >
void dataArray( std::ostream& o, const char* type, const std::string& name, int ncomp, const std::vector< T >& v,
                bool binary )
{
   o << "<DataArray type=\"" << type << "\"";
   if ( !name.empty() )
      o << " Name=\"" << name << "\"";
   if ( ncomp > 1 )
      o << " NumberOfComponents=\"" << ncomp << "\"";
   o << " format=\"" << ( binary ? "binary" : "ascii" ) << "\">\n";
   if ( binary )
   {
      const uint32_t             bytes = uint32_t( v.size() * sizeof( T ) );
      std::vector< unsigned char > buf( sizeof( uint32_t ) + bytes );
      std::memcpy( buf.data(), &bytes, sizeof( uint32_t ) );
      if ( bytes )
         std::memcpy( buf.data() + sizeof( uint32_t ), v.data(), bytes );
      o << base64( buf.data(), buf.size() ) << "\n";
   }
   else
   {
      o.precision( 17 );
      for ( size_t i = 0; i < v.size(); ++i )
         o << +v[i] << ( ( i + 1 ) % 12 == 0 ? "\n" : " " );
      o << "\n";
   }
   o << "</DataArray>\n";
}

} // namespace
} // namespace hg

using namespace hg;

extern "C" {

int hg_vtk_write( hg_ctx* h, const char* path, hg_handle storage, int level, int nfun, const hg_handle* functions,
                  const char* const* names, int binary )
{
   return hg_guard( h, [&]( Context* ctx ) {
      if ( !path || level < 0 || nfun < 0 || ( nfun > 0 && ( !functions || !names ) ) )
         throw Error( HG_ERR_ARG, "vtk: invalid arguments" );
      Storage&        st  = ctx->storage( storage );
      const Topology& t   = st.topo;
      int             off = -1;
      std::vector< Function* > fs;
      for ( int i = 0; i < nfun; ++i )
      {
         Function& f = ctx->function( functions[i] );
         if ( f.st != &st || f.lv != 0 || level < f.minL || level > f.maxL || ( off >= 0 && f.off != off ) || !names[i] )
            throw Error( HG_ERR_ARG, "vtk: the functions must be single-vector functions of one space on the storage, "
                                     "with the level in their range and a name each" );
         off = f.off;
         fs.push_back( &f );
      }
      if ( off < 0 )
         off = 0;
      const int     Lw  = level + off; // the lattice the values live on
      const int64_t mw  = int64_t( 1 ) << Lw, w = mw + 1, ppc = tet( w );
      const int64_t mm  = int64_t( 1 ) << level;
      LevelNumbering num( t, Lw );
      // values in the canonical order of the lattice
      std::vector< std::vector< double > > vals( fs.size(), std::vector< double >( size_t( num.n ) ) );
      for ( size_t i = 0; i < fs.size(); ++i )
      {
         const int st2 = hg_function_download( h, functions[i], level, vals[i].data() );
         if ( st2 != HG_OK )
            throw Error( st2, "vtk: download failed" );
      }
      const int64_t        np = int64_t( t.nc ) * ppc;
      std::vector< double > pts( size_t( 3 * np ) );
      std::vector< std::vector< double > > pdata( fs.size(), std::vector< double >( size_t( np ) ) );
      for ( int c = 0; c < t.nc; ++c )
         for ( int64_t z = 0; z < w; ++z )
            for ( int64_t y = 0; y + z < w; ++y )
               for ( int64_t x = 0; x + y + z < w; ++x )
               {
                  const int64_t p = c * ppc + tetIndex( w, x, y, z );
                  latticePoint( t, c, mw, x, y, z, &pts[size_t( 3 * p )] );
                  const int64_t id = num.locate( t, c, x, y, z );
                  for ( size_t i = 0; i < fs.size(); ++i )
                     pdata[i][size_t( p )] = vals[i][size_t( id )];
               }
      // cells: the level-L micro-tetrahedra of every macro-cell
      const int             npc = off ? 10 : 4;
      std::vector< int64_t > conn, offs;
      std::vector< uint8_t > types;
      // VTK_QUADRATIC_TETRA node order after the corners: edges 01, 12, 02, 03, 13, 23
      const int qe[6][2] = { { 0, 1 }, { 1, 2 }, { 0, 2 }, { 0, 3 }, { 1, 3 }, { 2, 3 } };
      for ( int c = 0; c < t.nc; ++c )
         for ( int ty = 0; ty < 6; ++ty )
         {
            const int64_t wt = mm - MicroCells::rowShift[ty];
            for ( int64_t z = 0; z < wt; ++z )
               for ( int64_t y = 0; y + z < wt; ++y )
                  for ( int64_t x = 0; x + y + z < wt; ++x )
                  {
                     int64_t v[4][3];
                     for ( int k = 0; k < 4; ++k )
                     {
                        v[k][0] = ( x + MicroCells::vof[ty][k][0] ) << off;
                        v[k][1] = ( y + MicroCells::vof[ty][k][1] ) << off;
                        v[k][2] = ( z + MicroCells::vof[ty][k][2] ) << off;
                        conn.push_back( c * ppc + tetIndex( w, v[k][0], v[k][1], v[k][2] ) );
                     }
                     if ( off )
                        for ( const auto& e : qe )
                           conn.push_back( c * ppc + tetIndex( w, ( v[e[0]][0] + v[e[1]][0] ) / 2,
                                                               ( v[e[0]][1] + v[e[1]][1] ) / 2,
                                                               ( v[e[0]][2] + v[e[1]][2] ) / 2 ) );
                     offs.push_back( int64_t( conn.size() ) );
                     types.push_back( off ? 24 : 10 );
                  }
         }
      std::ofstream o( path );
      if ( !o )
         throw Error( HG_ERR_ARG, std::string( "vtk: cannot open " ) + path );
      o << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
           "header_type=\"UInt32\">\n<UnstructuredGrid>\n<Piece NumberOfPoints=\""
        << np << "\" NumberOfCells=\"" << types.size() << "\">\n<Points>\n";
      dataArray( o, "Float64", "", 3, pts, binary != 0 );
      o << "</Points>\n<Cells>\n";
      dataArray( o, "Int64", "connectivity", 1, conn, binary != 0 );
      dataArray( o, "Int64", "offsets", 1, offs, binary != 0 );
      dataArray( o, "UInt8", "types", 1, types, binary != 0 );
      o << "</Cells>\n<PointData>\n";
      for ( size_t i = 0; i < fs.size(); ++i )
         dataArray( o, "Float64", names[i], 1, pdata[i], binary != 0 );
      o << "</PointData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
      if ( !o )
         throw Error( HG_ERR_STATE, "vtk: write failed" );
   } );
}

} // extern "C"
