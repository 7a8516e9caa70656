// pair_math.hpp -- device arithmetic shared by the global-list kernels (k_pair.hip)
// and the cell-block kernels (k_block.hip): spline evaluation (eam_alloy.h:198-313),
// the EAM pair scalars (eam_alloy_mm_force), the per-atom output epilogue.
//
// Real is the type of the pair and spline arithmetic: double for the plain FP64
// form, float for the mixed form (FP64 displacement, d^2, cutoff test and all
// accumulations in either case).
#pragma once

#include "xsmd_internal.hpp"

namespace xsmd {
namespace pm {

template <class Real>
struct Q4 {
  Real a, b, c, d;
};

__device__ __forceinline__ Q4<float> load4(const float* __restrict__ p)
{
  const float4 v = *reinterpret_cast<const float4*>(p);
  return {v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ Q4<double> load4(const double* __restrict__ p)
{
  const double2 u = reinterpret_cast<const double2*>(p)[0];
  const double2 w = reinterpret_cast<const double2*>(p)[1];
  return {u.x, u.y, w.x, w.y};
}

// sqrt(d2) and 1/sqrt(d2) correctly rounded in FP64 (what plain FP64 code compiles to)
__device__ __forceinline__ void dist(double d2, double& r, double& rinv)
{
  r = sqrt(d2);
  rinv = 1.0 / r;
}
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }

// XCD-aware block order: hardware hands workgroup b to XCD b % 8, so map it to
// logical block (b % 8) * (G/8) + b / 8 and each XCD's L2 sees one contiguous
// (hence spatially compact) range of atoms.  G is padded to a multiple of 8.
__device__ __forceinline__ int64_t logical_block()
{
  const unsigned b = blockIdx.x, g = gridDim.x;
  return (int64_t)(b & 7u) * (g >> 3) + (b >> 3);
}

struct Outputs {
  double* ax;
  double* ay;
  double* az;
  double* ep;
  double* vir;
  const double* mass;  // fused epilogue
  const uint8_t* type;
  double* vx;
  double* vy;
  double* vz;
  double kick_dt;
  int64_t npad;
};

// Per-atom epilogue: add the force (operator form) or write a = f/m and kick v
// (fused MD form); energy and virial when requested.
template <bool ENERGY, bool VIRIAL, bool FUSED>
__device__ __forceinline__ void write_outputs(const Outputs& o, int64_t i, double fx, double fy, double fz,
                                              double e, const double* v)
{
  if constexpr (FUSED) {
    const double m = o.mass[o.type[i]];
    const double a0 = fx / m, a1 = fy / m, a2 = fz / m;
    o.ax[i] = a0;
    o.ay[i] = a1;
    o.az[i] = a2;
    o.vx[i] = fma(a0, o.kick_dt, o.vx[i]);
    o.vy[i] = fma(a1, o.kick_dt, o.vy[i]);
    o.vz[i] = fma(a2, o.kick_dt, o.vz[i]);
  } else {
    o.ax[i] += fx;
    o.ay[i] += fy;
    o.az[i] += fz;
  }
  if constexpr (ENERGY) o.ep[i] += e;
  if constexpr (VIRIAL) {
#pragma unroll
    for (int q = 0; q < 9; q++) {
      if constexpr (FUSED) o.vir[q * o.npad + i] = v[q];
      else o.vir[q * o.npad + i] += v[q];
    }
  }
}

// Next-step drift fused into the MD-loop epilogue: the arithmetic of k_drift_kick
// applied to the velocity this step's kick produced.  The loop only needs to know
// whether some atom moved more than the rebuild threshold, so an atom past it
// stores 1 to *over (every writer stores the same value).
struct Drift {
  Pos4* pos_out;
  const double* rrx;
  const double* rry;
  const double* rrz;
  double dt, dt2, hdt;
  double thr2;
  unsigned long long* over;
};

// Fused epilogue with drift: a = f/m, v += a*kick_dt (this step), then
// push_f_v_r(dt) and push_f_v(dt/2) of the next step from the record p of atom i,
// and the rebuild test on the displacement since the last list build.
__device__ __forceinline__ void write_outputs_drift(const Outputs& o, const Drift& d, int64_t i, const Pos4& p,
                                                    double fx, double fy, double fz)
{
  const double m = o.mass[o.type[i]];
  const double a0 = fx / m, a1 = fy / m, a2 = fz / m;
  o.ax[i] = a0;
  o.ay[i] = a1;
  o.az[i] = a2;
  const double v0 = fma(a0, o.kick_dt, o.vx[i]);
  const double v1 = fma(a1, o.kick_dt, o.vy[i]);
  const double v2 = fma(a2, o.kick_dt, o.vz[i]);
  Pos4 q = p;
  q.x += fma(v0, d.dt, a0 * d.dt2);
  q.y += fma(v1, d.dt, a1 * d.dt2);
  q.z += fma(v2, d.dt, a2 * d.dt2);
  q.w = 0.0;
  store_pos(d.pos_out, i, q);
  o.vx[i] = fma(a0, d.hdt, v0);
  o.vy[i] = fma(a1, d.hdt, v1);
  o.vz[i] = fma(a2, d.hdt, v2);
  if (xsmd_d2(q.x - d.rrx[i], q.y - d.rry[i], q.z - d.rrz[i]) > d.thr2) *d.over = 1ull;
}

// onika tensor(fe, dr) * -0.5 added to v (row-major m11 m12 ... m33)
__device__ __forceinline__ void add_virial(double* v, double fe0, double fe1, double fe2, double dx, double dy,
                                           double dz)
{
  const double fe[3] = {fe0, fe1, fe2};
  const double dr[3] = {dx, dy, dz};
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) v[3 * a + b] += (fe[a] * dr[b]) * -0.5;
}

template <class Real>
__device__ __forceinline__ void knot_r(const EamTabDev<Real>& t, Real r, int& m, Real& p)
{
  p = r * t.rdr + Real(1);
  m = (int)p;
  m = max(0, min(m, t.nr - 1));
  p -= (Real)m;
  p = fmin(p, Real(1));
}

// Distance, 1/r and spline knot of a pair from its FP64 d^2 (the first steps of
// eam_alloy_rho_noderiv / eam_alloy_mm_force, eam_alloy.h:198-265: p = r rdr + 1,
// m = min(int p, nr - 1), p -= m, p = min(p, 1)).
// FP64 form: correctly rounded sqrt and division, as plain FP64 code.
// Mixed form: the knot fraction of r ~ 5 A at 1000 knots per A needs more than FP32
// (its ulp at p ~ 5000 is 5e-4 of a knot, which puts ~1e-5 of relative error on the
// forces of atoms whose pair forces nearly cancel), so 1/sqrt(d2) is the FP32
// hardware rsq refined by one FP64 Newton step (relative error ~1e-14), r and p are
// formed in FP64, and only the fraction (in [0, 1]) and 1/r are rounded to FP32 for
// the spline arithmetic.
__device__ __forceinline__ void knot_dist(const EamTabDev<double>& t, double d2, int& m, double& p, double& rinv)
{
  double r;
  dist(d2, r, rinv);
  knot_r(t, r, m, p);
}
__device__ __forceinline__ void knot_dist(const EamTabDev<float>& t, double d2, int& m, float& p, float& rinv)
{
  const double y0 = (double)__builtin_amdgcn_rsqf((float)d2);
  const double y = y0 * fma(-0.5 * d2 * y0, y0, 1.5);
  double q = fma(d2 * y, t.rdr_d, 1.0);
  m = (int)q;
  m = max(0, min(m, t.nr - 1));
  q -= (double)m;
  p = (float)fmin(q, 1.0);
  rinv = (float)y;
}

// Element `i` of a table whose base is the same for every lane: the byte offset is
// formed in 32 bits (every table here is far below 4 GiB and its indices are
// non-negative), so the loads take the scalar-base form with one 32-bit offset per
// lane instead of 64-bit address arithmetic.
template <class T>
__device__ __forceinline__ const T* at32(const T* base, uint32_t i)
{
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (i * (uint32_t)sizeof(T)));
}

template <class Real>
__device__ __forceinline__ const Real* row_r(const Real* tab, const EamTabDev<Real>& t, int which, int m)
{
  return at32(tab, ((uint32_t)which * (uint32_t)(t.nr + 1) + (uint32_t)m) * 8u);
}

template <class Real>
__device__ __forceinline__ Real cubic(const Q4<Real>& c, Real p)
{
  return ((c.a * p + c.b) * p + c.c) * p + c.d;
}
template <class Real>
__device__ __forceinline__ Real quad(const Q4<Real>& c, Real p)
{
  return (c.a * p + c.b) * p + c.c;
}
// derivative of the value cubic with respect to p
template <class Real>
__device__ __forceinline__ Real dcubic(const Q4<Real>& c, Real p)
{
  return (Real(3) * c.a * p + Real(2) * c.b) * p + c.c;
}
// The FP32 forms spell out their fused multiply-adds, so every kernel that shares
// them (block and global lists, full and half lists) rounds them alike whatever the
// compiler's contraction choices around them.
__device__ __forceinline__ float cubic(const Q4<float>& c, float p) { return fmaf(fmaf(fmaf(c.a, p, c.b), p, c.c), p, c.d); }
__device__ __forceinline__ float quad(const Q4<float>& c, float p) { return fmaf(fmaf(c.a, p, c.b), p, c.c); }
__device__ __forceinline__ float dcubic(const Q4<float>& c, float p)
{
  return fmaf(fmaf(3.0f * c.a, p, 2.0f * c.b), p, c.c);
}

// types slot of eam_alloy_force (m_pair_enabled, eam_potential_multimat.cu:154-166):
// species `type` takes part when its bit is set; a pair counts when both species do,
// and an atom of a disabled species keeps rho = 0, F' = 0 and gets no force.
template <class Real>
__device__ __forceinline__ bool eam_on(const EamTabDev<Real>& t, int type)
{
  return (t.types >> type) & 1u;
}

// F(rho), F'(rho) in eV-converted internal units, with the linear extrapolation
// of F above rhomax (eam_alloy.h:289-313).
template <class Real>
__device__ __forceinline__ void embed(const EamTabDev<Real>& t, int ti, double rho, Real& F, Real& Fp)
{
  const Real rr = (Real)rho;
  Real p;
  if constexpr (sizeof(Real) == 4) p = fmaf(rr, t.rdrho, 1.0f);
  else p = rr * t.rdrho + Real(1);
  int m = (int)p;
  m = max(1, min(m, t.nrho - 1));
  p -= (Real)m;
  p = fmin(p, Real(1));
  const Real* row = t.frho + ((int64_t)t.t2frho[ti + 1] * (t.nrho + 1) + m) * 8;
  const Q4<Real> v = load4(row);
  const Q4<Real> dv = load4(row + 4);
  Fp = quad(dv, p);
  F = cubic(v, p);
  if (rr > t.rhomax) {
    if constexpr (sizeof(Real) == 4) F = fmaf(Fp, rr - t.rhomax, F);
    else F += Fp * (rr - t.rhomax);
  }
  F *= t.conv;
  Fp *= t.conv;
}

// Single-element pair table: two planes of (nr + 1) rows of 4 values, the rho value
// cubic [c3 c4 c5 c6] of knot m at row m of plane 0 and the z2r value cubic at row m
// of plane 1.  Rows of 4 values keep 8 (FP32) or 4 (FP64) knots per 128-byte line,
// so the clustered knots that the lanes of a wave look up touch few lines.
template <class Real>
__device__ __forceinline__ const Real* pair_rho_row(const Real* __restrict__ pair, int m)
{
  return at32(pair, (uint32_t)m * 4u);
}
template <class Real>
__device__ __forceinline__ const Real* pair_z2_row(const Real* __restrict__ pair, const EamTabDev<Real>& t, int m)
{
  return at32(pair, (uint32_t)(t.nr + 1 + m) * 4u);
}

// Single-element psi'(r)/r and phi(r) from the two rows of knot m:
// rc = rho value cubic, zc = z2r value cubic; the derivative quadratics follow
// from the value cubics (times 1/dr).
template <class Real>
__device__ __forceinline__ void eam_pair_rows(const EamTabDev<Real>& t, const Q4<Real>& rc, const Q4<Real>& zc, Real p,
                                              Real rinv, Real fpi, Real fpj, Real& fpair, Real& phi)
{
  const Real rhop = dcubic(rc, p) * t.rdr;
  const Real z2p = dcubic(zc, p) * t.rdr;
  const Real z2 = cubic(zc, p);
  phi = z2 * rinv;
  Real phip, psip;
  if constexpr (sizeof(Real) == 4) {
    phip = fmaf(z2p, rinv, -(phi * rinv)) * t.conv;
    psip = fmaf(fpi + fpj, rhop, phip);
  } else {
    phip = (z2p * rinv - phi * rinv) * t.conv;
    psip = (fpi + fpj) * rhop + phip;
  }
  phi *= t.conv;
  fpair = psip * rinv;
}

// ---------------------------------------------------------------------------
// FP32 tables in s = r^2 (single element, mixed form, force pass)
// ---------------------------------------------------------------------------
// The mixed force pass of one element reads, instead of the knot rows, cubic pieces
// in s = r^2 over [s_lo, s_hi] = [(0.3 rmax)^2, rmax^2] (rmax = (nr - 1) dr), uniform
// in s: piece k covers [s_lo + k ds, s_lo + (k+1) ds] in the local coordinate
// p = (s - s_lo) / ds - k.  Each piece is the Hermite cubic through the reference
// spline's FP64 values and s-slopes at its two ends (value and slope continuous from
// piece to piece), of
//   force plane: rho'(r)/r | phi'(r) conv / r   [A B C 0 | A B C 0]: the s-derivative
//                quadratics times 2/ds, so (A p + B) p + C is the function itself
//   phi plane:   phi(r) conv = z2r(r) conv / r  [a b c d]  (energy)
// so a pair needs no square root and no division: its index comes from the FP64
// d^2 (p formed in FP64, rounded to FP32 in [0, 1]), and
// psi'(r)/r = (F'_i + F'_j) rho'(r)/r + phi'(r)/r is one FMA of two quadratics.
// The host takes the fewest pieces (256 .. 8192, doubling) whose FP32 evaluation stays
// within 8 FP32 ulps of the largest |rho'/r|, |phi'/r| and 4 of the largest |phi| of
// the FP64 spline over [s_lo, s_hi] (the FP32 knot rows are at about 1-4 of each), or
// no s-table (then every pair reads the knot rows).  A pair with d^2 outside
// [s_lo, s_hi] (closer than 0.3 rmax, or past the table's end) takes the knot rows:
// the reference evaluates r beyond rmax at the table's end, which the s-table cannot
// express.  An FP64 cubic in s could not reproduce the spline to FP64 precision, so
// the FP64 form keeps the knot rows; the density pass, a single cubic per pair, keeps
// them too (its FP32 s-table form measured no faster: the pass is bound by its
// gathers, which the s-table does not reduce).

// row and local coordinate of d^2 in the s-table; `out` = d^2 outside [s_lo, s_hi]
// (with no s-table the host sets a range no d^2 falls in, so every pair is `out`)
__device__ __forceinline__ void s_knot(const EamTabDev<float>& t, double d2, int& m, float& p, bool& out)
{
  const double x0 = fma(d2, t.s_rds, t.s_off);
  const double x = fmin(fmax(x0, 0.0), t.s_xmax);
  out = x != x0;
  m = min((int)x, t.sprows - 1);
  p = (float)(x - (double)m);
}

// d^2 outside the s-table's range [s_lo, s_hi]
__device__ __forceinline__ bool s_out(const EamTabDev<float>& t, double d2)
{
  const double x = fma(d2, t.s_rds, t.s_off);
  return !(x >= 0.0 && x <= t.s_xmax);
}

// the slope quadratic [A B C] of an s-table force row (its fourth value is padding)
struct Q3 {
  float a, b, c;
};
__device__ __forceinline__ Q3 load3(const float* __restrict__ p)
{
  const float3 v = *reinterpret_cast<const float3*>(p);
  return {v.x, v.y, v.z};
}
__device__ __forceinline__ float quad3(const Q3& c, float p) { return fmaf(fmaf(c.a, p, c.b), p, c.c); }

// The FP32 single-element form reads the s-table; every other form the knot rows.
template <class Real, bool MULTI>
constexpr bool kSTable = !MULTI && sizeof(Real) == 4;

// rho_{type_j}(r) contribution to atom i (SymRhoOp, eam_force_op_multimat.h:142-171)
// from the knot rows (single element: plane 0 of the pair table).  `in` false
// evaluates at d2 = 1 instead (a guard for list padding whose d2 may be 0); the value
// of a pair that does not count is arbitrary and callers mask it.  Callers whose d2 is
// always finite and positive pass true.
template <class Real, bool MULTI>
__device__ __forceinline__ Real eam_rho_knot(const EamTabDev<Real>& t, const Real* __restrict__ pair, int ti, int tj,
                                             double d2, bool in)
{
  Real p, rinv;
  int m;
  knot_dist(t, in ? d2 : 1.0, m, p, rinv);
  const Real* row;
  if constexpr (MULTI) row = row_r(t.rhor, t, t.t2rhor[tj + 1][ti + 1], m);
  else row = pair_rho_row(pair, m);
  return cubic(load4(row), p);
}

// psi'(r)/r and phi(r) of eam_alloy_mm_force (eam_alloy.h:222-265) from the knot
// rows: psi' = F'_i rho'_j + F'_j rho'_i + phi', phi = z2r / r.
template <class Real, bool MULTI>
__device__ __forceinline__ void eam_pair_knot(const EamTabDev<Real>& t, const Real* __restrict__ pair, int ti, int tj,
                                              Real fpi, Real fpj, double d2, bool in, Real& fpair, Real& phi)
{
  Real p, rinv;
  int m;
  knot_dist(t, in ? d2 : 1.0, m, p, rinv);
  if constexpr (MULTI) {
    const Real rhoip = quad(load4(row_r(t.rhor, t, t.t2rhor[ti + 1][tj + 1], m) + 4), p);
    const Real rhojp = quad(load4(row_r(t.rhor, t, t.t2rhor[tj + 1][ti + 1], m) + 4), p);
    const Real* zrow = row_r(t.z2r, t, t.t2z2r[ti + 1][tj + 1], m);
    const Real z2p = quad(load4(zrow + 4), p);
    const Real z2 = cubic(load4(zrow), p);
    phi = z2 * rinv;
    Real phip, psip;
    if constexpr (sizeof(Real) == 4) {
      phip = fmaf(z2p, rinv, -(phi * rinv)) * t.conv;
      psip = fmaf(fpi, rhojp, fmaf(fpj, rhoip, phip));
    } else {
      phip = (z2p * rinv - phi * rinv) * t.conv;
      psip = fpi * rhojp + fpj * rhoip + phip;
    }
    phi *= t.conv;
    fpair = psip * rinv;
  } else {
    eam_pair_rows(t, load4(pair_rho_row(pair, m)), load4(pair_z2_row(pair, t, m)), p, rinv, fpi, fpj, fpair, phi);
  }
}

// Pair scalars psi'(r)/r (and, with PHI, phi(r)) of U pairs (nb[u].d2, nb[u].in,
// nb[u].w = F'_j; neighbour species tj[u]).  The s-table form evaluates all U from
// the s-table (their row loads in flight together: psi'/r = (F'_i + F'_j) rho'(r)/r +
// phi'(r)/r from the two slope quadratics of one 32-byte row) and flags slow[u] = a
// counted pair outside the s-table's range, which the caller evaluates with
// eam_pair_knot, only in the rare wave where some slow[u] is set.  Every other form
// reads the knot rows and leaves slow[] false.
template <class Real, bool MULTI, bool PHI, int U, class Nb>
__device__ __forceinline__ void eam_pair_batch(const EamTabDev<Real>& t, const Real* __restrict__ pair, int ti,
                                               const int (&tj)[U], Real fpi, const Nb (&nb)[U], Real (&fpair)[U],
                                               Real (&phi)[U], bool (&slow)[U])
{
  if constexpr (kSTable<Real, MULTI>) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      int m;
      float p;
      bool out;
      s_knot(t, nb[u].d2, m, p, out);
      const Real* row = at32(pair + t.s_force, (uint32_t)m * 8u);
      const Q3 gr = load3(row), gz = load3(row + 4);
      fpair[u] = fmaf(fpi + (Real)nb[u].w, quad3(gr, p), quad3(gz, p));
      if constexpr (PHI) phi[u] = cubic(load4(at32(pair + t.s_phi, (uint32_t)m * 4u)), p);
      slow[u] = nb[u].in && out;
    }
  } else {
#pragma unroll
    for (int u = 0; u < U; u++) {
      eam_pair_knot<Real, MULTI>(t, pair, ti, tj[u], fpi, (Real)nb[u].w, nb[u].d2, nb[u].in, fpair[u], phi[u]);
      slow[u] = false;
    }
  }
}

// Lennard-Jones pair (lennard_jones.h:40-50, force_op_impl2.hxx:33-78) written in
// r^2 only: (sigma/r)^2 = sigma^2 / d2, de/r = -24 eps (2 s^12 - s^6) / d2, e shifted
// by E(rcut) (pair_potential_impl.hxx:487-498).  Pair math in Real, FP64 out.
template <class Real>
struct LJ {
  Real sigma2, eps4, eps24, ecut;
};

// lj_multi_force: parameters per unordered pair of species (packed triangle,
// entry b*(b+1)/2 + a for a <= b) and each pair's cutoff radius; a pair counts
// when r <= its rcut (pair_potential_force_op_multiparam.h:82-84).
struct LJMultiDev {
  const void* rows;          // LJ<Real> per type pair
  const double* __restrict__ rcut;
};
__device__ __forceinline__ int type_pair(int a, int b)
{
  const int lo = min(a, b), hi = max(a, b);
  return hi * (hi + 1) / 2 + lo;
}
template <class Real>
__device__ __forceinline__ LJ<Real> lj_multi_row(const LJMultiDev& m, int pid)
{
  return reinterpret_cast<const LJ<Real>*>(m.rows)[pid];
}
__device__ __forceinline__ bool lj_multi_in(const LJMultiDev& m, int pid, double d2)
{
  return sqrt(d2) <= m.rcut[pid];
}

template <class Real>
__device__ __forceinline__ void lj_pair(const LJ<Real>& lp, double d2, bool in, double& de_r, double& e)
{
  const Real ir2 = recip((Real)(in ? d2 : 1.0));
  const Real s2 = lp.sigma2 * ir2;
  const Real s6 = s2 * s2 * s2;
  const Real s12 = s6 * s6;
  de_r = in ? (double)(-lp.eps24 * (Real(2) * s12 - s6) * ir2) : 0.0;
  e = in ? 0.5 * (double)(lp.eps4 * (s12 - s6) - lp.ecut) : 0.0;
}

}  // namespace pm

pm::Outputs make_outputs(Ctx& c, const ForceEpilogue& ep);

// device view of the lj_multi_force tables in the precision Real
template <class Real>
inline pm::LJMultiDev ljm_dev(const Ctx& c)
{
  pm::LJMultiDev m;
  if constexpr (sizeof(Real) == 4) m.rows = c.ljm_f.p;
  else m.rows = c.ljm_d.p;
  m.rcut = c.ljm_rcut.p;
  return m;
}

}  // namespace xsmd
