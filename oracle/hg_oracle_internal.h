/*
 * hg_oracle_internal.h -- shared internals of the CPU oracle (test
 * infrastructure only; see hg_oracle.h).
 */
#ifndef HG_ORACLE_INTERNAL_H
#define HG_ORACLE_INTERNAL_H

#include "hg_oracle.h"

/* ------------------------------------------------------------------------- */
/* lattice helpers                                                            */
/* ------------------------------------------------------------------------- */
static inline int64_t tri( int64_t n ) { return n <= 0 ? 0 : n * ( n + 1 ) / 2; }
static inline int64_t tet( int64_t n ) { return n <= 0 ? 0 : n * ( n + 1 ) * ( n + 2 ) / 6; }
/* MacroCellIndexing.hpp:40-53, width n (points per edge) */
static inline int64_t tri_index( int64_t n, int64_t x, int64_t y ) { return y * n - y * ( y - 1 ) / 2 + x; }
static inline int64_t tet_index( int64_t n, int64_t x, int64_t y, int64_t z )
{
   return tet( n ) - tet( n - z ) + y * ( n - z ) - y * ( y - 1 ) / 2 + x;
}

/* 6 micro-cell types (CellDoFIndexing.hpp:38-60, 150-200) */
static const int VOF[6][4][3] = {
    { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } }, /* WHITE_UP   */
    { { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 }, { 1, 0, 1 } }, /* BLUE_UP    */
    { { 1, 0, 0 }, { 0, 1, 0 }, { 1, 0, 1 }, { 0, 0, 1 } }, /* GREEN_UP   */
    { { 1, 1, 0 }, { 1, 1, 1 }, { 0, 1, 1 }, { 1, 0, 1 } }, /* WHITE_DOWN */
    { { 1, 0, 1 }, { 0, 1, 1 }, { 0, 0, 1 }, { 0, 1, 0 } }, /* BLUE_DOWN  */
    { { 0, 1, 0 }, { 1, 1, 0 }, { 1, 0, 1 }, { 0, 1, 1 } }, /* GREEN_DOWN */
};
/* micro-cells per row by type (CellDoFIndexing.hpp:64-83): m - WSHIFT */
static const int WSHIFT[6] = { 0, 1, 1, 2, 1, 1 };

/* 15-point stencil offsets (apply_3D_macrocell_vertexdof_to_vertexdof_replace.cpp:36-50) */
static const int OFF15[15][3] = { { 0, 0, 0 },  { -1, 0, 0 }, { -1, 0, 1 }, { -1, 1, -1 }, { -1, 1, 0 },
                                  { 0, -1, 0 }, { 0, -1, 1 }, { 0, 0, -1 }, { 0, 0, 1 },   { 0, 1, -1 },
                                  { 0, 1, 0 },  { 1, -1, 0 }, { 1, -1, 1 }, { 1, 0, -1 },  { 1, 0, 0 } };
/* local edges / faces of a cell (SetupPrimitiveStorage.cpp:290-300) */
static const int LEDGE[6][2] = { { 0, 1 }, { 0, 2 }, { 1, 2 }, { 0, 3 }, { 1, 3 }, { 2, 3 } };
static const int LFACE[4][3] = { { 0, 1, 2 }, { 0, 1, 3 }, { 0, 2, 3 }, { 1, 2, 3 } };

/* ------------------------------------------------------------------------- */
/* mesh                                                                       */
/* ------------------------------------------------------------------------- */
struct or_mesh
{
   int     nv, ne, nf, nc;
   double* X;
   int*    ev; /* ne*2, sorted */
   int*    fv; /* nf*3, sorted */
   int*    cv; /* nc*4, local order */
   int*    ce; /* nc*6 */
   int*    cf; /* nc*4 */
   int *   vflag, *eflag, *fflag;
   int *   vnc, *enc, *fnc;
   /* icosahedral-shell blending per cell, NULL for affine meshes:
    * ray[3], ref[3], thr[3], for[3], radRay, radRef (IcosahedralShellMap.hpp:608-731) */
   double* blend;
};

typedef struct
{
   int64_t mm;                         /* 2^L micro-edges per macro-edge */
   int64_t eoff, foff, coff, n;        /* block offsets and total */
   int64_t ein, fin, cin;              /* interior points per edge/face/cell */
} lvl_t;


/* k(x) = exp( sum_m a_m sin(omega_m . x + phi_m) ); nm == 0 means k == 1 */
typedef struct
{
   int     nm;
   double* a;     /* nm */
   double* omega; /* nm*3 */
   double* phi;   /* nm */
} kfun_t;

struct or_p2op
{
   const or_mesh* m;
   int            L; /* P2 level; storage lattice is level L+1 */
   lvl_t          l1;
   int64_t        sc1, w1;
   int64_t*       cmap; /* nc * sc1 */
   kfun_t         K;
   int32_t*       flags;
   int            reduced; /* element matrices of or_p2_reduced_element_matrix */
};

lvl_t    or_lvl_info( const or_mesh* m, int L );
int64_t  or_locate_l( const or_mesh* m, const lvl_t* l, int c, int64_t x, int64_t y, int64_t z );
int      or_owner_flag( const or_mesh* m, const lvl_t* l, int64_t id );
double   or_inv_ncells( const or_mesh* m, const lvl_t* l, int64_t id );
int64_t* or_build_cmap( const or_mesh* m, int L, int64_t* sc_out );
void     or_axpby_f( int64_t n, double alpha, const double* x, double beta, double* y, const int32_t* flags, int flag );

#endif
