/*
 * xsmd_pair_geometry.h -- the pair-geometry contract shared by the GPU kernels
 * and the CPU oracle.
 *
 * exaStamp handles periodicity with ghost cells: a periodic image of atom j is
 * a ghost particle whose position is r_j + k*L (computed once per ghost update),
 * and the pair displacement is ghost - r_i (include/exaStamp/fields.h:32-37 for
 * the fields; ghost updates in data/config/config_move_particles.msp:84-95,134-135).
 * This library stores the image k in the neighbour entry instead of a ghost
 * copy.  Because k*L is exact for k in {-1,0,1}, fma(k, L, x_j) rounds exactly
 * once, like the ghost position x_j + k*L, so the FP64 displacement is the same
 * number either way.
 *
 * Every function here is the single definition of a quantity whose bits decide
 * list membership (cell index, squared distance).  Both the HIP kernels and the
 * oracle include this file, so a neighbour list built on the GPU is bit-exact
 * against the oracle's.
 *
 * Includers may define XSMD_HD (e.g. "__host__ __device__ inline") before the
 * include; it defaults to "static inline" for plain C/C++.
 */
#ifndef XSMD_PAIR_GEOMETRY_H
#define XSMD_PAIR_GEOMETRY_H

#include <stdint.h>
#include <math.h>

#ifndef XSMD_HD
#define XSMD_HD static inline
#endif

/* Neighbour entry: bits 0..25 = neighbour atom index (cell-sorted order),
 * bits 26..27 / 28..29 / 30..31 = periodic image along x / y / z as a 2-bit
 * two's-complement integer in {-1,0,1}.  An entry with no image shift is just
 * the index, so "entry >> 26 == 0" tests for the common interior case. */
#define XSMD_NBR_INDEX_BITS 26
#define XSMD_NBR_INDEX_MASK ((uint32_t)((1u << XSMD_NBR_INDEX_BITS) - 1u))
#define XSMD_NBR_MAX_ATOMS ((int64_t)1 << XSMD_NBR_INDEX_BITS)

XSMD_HD uint32_t xsmd_nbr_encode(uint32_t j, int ix, int iy, int iz)
{
  return j | ((uint32_t)(ix & 3) << 26) | ((uint32_t)(iy & 3) << 28) | ((uint32_t)(iz & 3) << 30);
}

XSMD_HD uint32_t xsmd_nbr_index(uint32_t e) { return e & XSMD_NBR_INDEX_MASK; }
XSMD_HD int xsmd_nbr_img_x(uint32_t e) { return ((int32_t)(e << 4)) >> 30; }
XSMD_HD int xsmd_nbr_img_y(uint32_t e) { return ((int32_t)(e << 2)) >> 30; }
XSMD_HD int xsmd_nbr_img_z(uint32_t e) { return ((int32_t)e) >> 30; }

/* Coordinate of the periodic image k of x (k in {-1,0,1}): x + k*L, one rounding. */
XSMD_HD double xsmd_image(double x, int k, double L) { return fma((double)k, L, x); }

/* Squared distance of a displacement; fixed FMA order. */
XSMD_HD double xsmd_d2(double dx, double dy, double dz)
{
  return fma(dz, dz, fma(dy, dy, dx * dx));
}

/* Cell coordinate along one axis: floor((x - lo) * (n / L)), clamped to [0, n-1].
 * inv_cs = n / L is computed once on the host in double. */
XSMD_HD int xsmd_cell_coord(double x, double lo, double inv_cs, int n)
{
  double t = (x - lo) * inv_cs;
  int c = (t < 0.0) ? -1 : (int)t;
  return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

/* Gap along one axis between coordinate x and the slab of cell c (n cells over
 * [lo, lo + L)), the slab widened by 1e-6 A to cover the rounding of cell
 * assignment; on a non-periodic axis the edge cells also hold the atoms clamped
 * into them, so their slabs are unbounded on the outer side.  0 inside. */
XSMD_HD double xsmd_cell_gap(double x, double lo, double L, int n, int c, int periodic)
{
  const double cs = L / n;
  const double margin = 1e-6;
  double a = lo + c * cs - margin;
  double b = lo + (c + 1) * cs + margin;
  double g = 0.0;
  if (!(c == 0 && !periodic) && x < a) g = a - x;
  if (!(c == n - 1 && !periodic) && x > b) g = x - b;
  return g;
}

/* Deformed domain.  Positions live in grid space; the domain's xform m (row-major
 * 3x3, exaNBody's Domain::xform) maps a grid-space displacement d to the physical one,
 * m*d, as LinearXForm does for every pair (pair_potential_impl.hxx:388-401 selects the
 * LinearXForm traversal when the xform is not the identity).  Fixed operation order,
 * so the GPU list build and the oracle agree bit for bit. */
XSMD_HD void xsmd_xform_apply(const double* m, double* x, double* y, double* z)
{
  const double a = *x, b = *y, c = *z;
  *x = fma(m[2], c, fma(m[1], b, m[0] * a));
  *y = fma(m[5], c, fma(m[4], b, m[3] * a));
  *z = fma(m[8], c, fma(m[7], b, m[6] * a));
}

/* Grid-space reach along axis a of a physical distance of 1 under xform m, given
 * mi = m^-1: a physical displacement d has grid component (mi d)_a, at most
 * |row a of mi| * |d| in magnitude.  Cells of grid width >= rl * reach therefore
 * hold every physical neighbour within rl in the 27-cell stencil. */
XSMD_HD double xsmd_xform_reach(const double* mi, int a)
{
  const double* r = mi + 3 * a;
  return sqrt(fma(r[2], r[2], fma(r[1], r[1], r[0] * r[0])));
}

/* Number of cells along an axis of length L for a minimum cell size cs. */
XSMD_HD int xsmd_cells_along(double L, double cs)
{
  int n = (int)(L / cs);
  return n < 1 ? 1 : n;
}

/* Wrap one coordinate back into [lo, lo+L) after at most one box crossing,
 * as particle migration does at a neighbour-list rebuild. */
XSMD_HD double xsmd_wrap(double x, double lo, double L)
{
  if (x < lo) return x + L;
  if (x >= lo + L) return x - L;
  return x;
}

#endif /* XSMD_PAIR_GEOMETRY_H */
