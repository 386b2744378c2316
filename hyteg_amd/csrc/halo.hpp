// Device-side geometry of owned macro-vertex/edge/face DoFs and their halo
// copies in the incident macro-cell lattices.  Replaces HyTeG's pack infos
// (VertexDoFPackInfo.hpp:192-700) by arithmetic on per-cell orientation data:
// an owned DoF is identified by its primitive and its local lattice
// coordinates; its position in an incident cell follows from which cell-local
// vertex each of the primitive's (sorted) vertices is.
#pragma once

#include "runtime.hpp"

namespace hg {

struct BndGeom
{
   const CellMeta* meta;
   const int32_t * fptr, *fcells, *eptr, *ecells, *vptr, *vcells;
   int64_t         mm, w, sc, fin, ein, eoff, foff;

   static BndGeom of( const Storage& st, const Lattice& l )
   {
      return BndGeom{ st.meta.p,
                      st.faceCellPtr.p,
                      st.faceCells.p,
                      st.edgeCellPtr.p,
                      st.edgeCells.p,
                      st.vertCellPtr.p,
                      st.vertCells.p,
                      l.mm,
                      l.w,
                      l.sc,
                      l.num.fin,
                      l.num.ein,
                      l.num.eoff,
                      l.num.foff };
   }
};

// (x, y) of index i in a triangle of width n (rows of n, n-1, ... points)
__device__ __forceinline__ void decodeTri( int64_t n, int64_t i, int64_t& x, int64_t& y )
{
   const double b  = 2.0 * double( n ) + 1.0;
   int64_t      yy = int64_t( floor( 0.5 * ( b - sqrt( b * b - 8.0 * double( i ) ) ) ) );
   if ( yy < 0 )
      yy = 0;
   while ( yy > 0 && triIndex( n, 0, yy ) > i )
      --yy;
   while ( triIndex( n, 0, yy + 1 ) <= i )
      ++yy;
   y = yy;
   x = i - triIndex( n, 0, yy );
}

// barycentric lattice weights t[0..3] in cell-local vertex order; (x,y,z) = t[1..3]
__device__ __forceinline__ void faceToCell( const CellMeta& m, int k, int64_t mm, int64_t fx, int64_t fy, int64_t t[4] )
{
   t[0] = t[1] = t[2] = t[3] = 0;
   t[m.fperm[k][0]] = mm - fx - fy;
   t[m.fperm[k][1]] = fx;
   t[m.fperm[k][2]] = fy;
}
__device__ __forceinline__ void edgeToCell( const CellMeta& m, int k, int64_t mm, int64_t s, int64_t t[4] )
{
   t[0] = t[1] = t[2] = t[3] = 0;
   t[m.eperm[k][0]] = mm - s;
   t[m.eperm[k][1]] = s;
}

// Calls f(c, x, y, z, cls) for every macro-cell incident to owned boundary DoF
// b (0 <= b < nb), in a fixed order; cls = local face k (0..3), 4 + local
// edge k, or 10 + local vertex k.
template < class F >
__device__ __forceinline__ void forEachIncident( const BndGeom& g, int64_t b, F&& f )
{
   int64_t t[4];
   if ( b < g.eoff )
   {
      for ( int32_t e = g.vptr[b]; e < g.vptr[b + 1]; ++e )
      {
         const int64_t c = g.vcells[2 * e];
         const int     k = g.vcells[2 * e + 1];
         t[0] = t[1] = t[2] = t[3] = 0;
         t[k]                      = g.mm;
         f( c, t[1], t[2], t[3], 10 + k );
      }
   }
   else if ( b < g.foff )
   {
      const int64_t eid = ( b - g.eoff ) / g.ein;
      const int64_t s   = ( b - g.eoff ) - eid * g.ein + 1;
      for ( int32_t e = g.eptr[eid]; e < g.eptr[eid + 1]; ++e )
      {
         const int64_t c = g.ecells[2 * e];
         const int     k = g.ecells[2 * e + 1];
         edgeToCell( g.meta[c], k, g.mm, s, t );
         f( c, t[1], t[2], t[3], 4 + k );
      }
   }
   else
   {
      const int64_t fid = ( b - g.foff ) / g.fin;
      const int64_t i   = ( b - g.foff ) - fid * g.fin;
      int64_t       a, bb;
      decodeTri( g.mm - 2, i, a, bb );
      for ( int32_t e = g.fptr[fid]; e < g.fptr[fid + 1]; ++e )
      {
         const int64_t c = g.fcells[2 * e];
         const int     k = g.fcells[2 * e + 1];
         faceToCell( g.meta[c], k, g.mm, a + 1, bb + 1, t );
         f( c, t[1], t[2], t[3], k );
      }
   }
}

// the same incidences by index: entries [e0, e1) of owned boundary DoF b, and
// the cell point (c; x, y, z) of entry e (the order of forEachIncident)
__device__ __forceinline__ void incidentRange( const BndGeom& g, int64_t b, int32_t& e0, int32_t& e1 )
{
   if ( b < g.eoff )
   {
      e0 = g.vptr[b];
      e1 = g.vptr[b + 1];
   }
   else if ( b < g.foff )
   {
      const int64_t eid = ( b - g.eoff ) / g.ein;
      e0                = g.eptr[eid];
      e1                = g.eptr[eid + 1];
   }
   else
   {
      const int64_t fid = ( b - g.foff ) / g.fin;
      e0                = g.fptr[fid];
      e1                = g.fptr[fid + 1];
   }
}
__device__ __forceinline__ int64_t incidentAt( const BndGeom& g, int64_t b, int32_t e, int64_t& x, int64_t& y,
                                               int64_t& z )
{
   int64_t t[4] = { 0, 0, 0, 0 };
   int64_t c;
   if ( b < g.eoff )
   {
      c    = g.vcells[2 * e];
      t[g.vcells[2 * e + 1]] = g.mm;
   }
   else if ( b < g.foff )
   {
      const int64_t eid = ( b - g.eoff ) / g.ein;
      const int64_t s   = ( b - g.eoff ) - eid * g.ein + 1;
      c                 = g.ecells[2 * e];
      edgeToCell( g.meta[c], g.ecells[2 * e + 1], g.mm, s, t );
   }
   else
   {
      const int64_t fid = ( b - g.foff ) / g.fin;
      const int64_t i   = ( b - g.foff ) - fid * g.fin;
      int64_t       a, bb;
      decodeTri( g.mm - 2, i, a, bb );
      c = g.fcells[2 * e];
      faceToCell( g.meta[c], g.fcells[2 * e + 1], g.mm, a + 1, bb + 1, t );
   }
   x = t[1];
   y = t[2];
   z = t[3];
   return c;
}

// class of a cell lattice point: -1 interior, 0..3 face, 4..9 edge, 10..13 vertex
__device__ __forceinline__ int pointClass( int64_t mm, int64_t x, int64_t y, int64_t z )
{
   const int64_t t[4] = { mm - x - y - z, x, y, z };
   int           nz = 0, zmask = 0;
   for ( int k = 0; k < 4; ++k )
      if ( t[k] == 0 )
      {
         ++nz;
         zmask |= 1 << k;
      }
   if ( nz == 0 )
      return -1;
   if ( nz == 1 )
   {
      // face k excludes vertex 3-k: faces (0,1,2), (0,1,3), (0,2,3), (1,2,3)
      for ( int k = 0; k < 4; ++k )
         if ( zmask == ( 1 << ( 3 - k ) ) )
            return k;
   }
   if ( nz == 2 )
   {
      // edges (0,1), (0,2), (1,2), (0,3), (1,3), (2,3): zero weights are the other two
      const int keep = ( ~zmask ) & 15;
      const int pairs[6] = { 3, 5, 6, 9, 10, 12 };
      for ( int k = 0; k < 6; ++k )
         if ( keep == pairs[k] )
            return 4 + k;
   }
   const int keep = ( ~zmask ) & 15;
   for ( int k = 0; k < 4; ++k )
      if ( keep == ( 1 << k ) )
         return 10 + k;
   return -1;
}

} // namespace hg
