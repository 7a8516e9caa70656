// setfl.cpp -- setfl (.eam.alloy) reader, spline coefficient build and upload of
// the EAM tables to the device in FP64 and FP32.
#include "xsmd_internal.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <limits>

// Coefficients must be bit-identical to the reference's, so no FMA contraction.
#pragma clang fp contract(off)

namespace xsmd {

// Cubic spline through knots f[0..n-1] (knot m+1 of the 1-based scheme), with the
// slopes from the fourth-order central difference in the interior and lower order
// at the ends: the construction of EamAlloyTools::interpolate
// (eam_alloy.cpp:29-58), evaluated in the same order so every coefficient has the
// same bits.  coeffs: (n+1) rows of 8, row 0 unused; row m = [c0..c6, 0] with
// value(p) = ((c3 p + c4) p + c5) p + c6 and derivative = (c0 p + c1) p + c2.
void spline_interpolate(int n, double delta, const double* f, double* coeffs)
{
  if (n < 5) throw Error(XS_E_ARG, "spline needs at least 5 knots");
  std::vector<double> v(n + 2, 0.0), s(n + 2, 0.0), a(n + 2, 0.0), b(n + 2, 0.0);
  for (int m = 1; m <= n; m++) v[m] = f[m - 1];
  s[1] = v[2] - v[1];
  s[2] = 0.5 * (v[3] - v[1]);
  s[n - 1] = 0.5 * (v[n] - v[n - 2]);
  s[n] = v[n] - v[n - 1];
  for (int m = 3; m <= n - 2; m++) s[m] = ((v[m - 2] - v[m + 2]) + 8.0 * (v[m + 1] - v[m - 1])) / 12.0;
  for (int m = 1; m <= n - 1; m++) {
    const double dv = v[m + 1] - v[m];
    b[m] = 3.0 * dv - 2.0 * s[m] - s[m + 1];   // p^2 coefficient
    a[m] = s[m] + s[m + 1] - 2.0 * dv;          // p^3 coefficient
  }
  for (int k = 0; k < 8; k++) coeffs[k] = 0.0;
  for (int m = 1; m <= n; m++) {
    double* row = coeffs + 8 * (size_t)m;
    row[6] = v[m];
    row[5] = s[m];
    row[4] = b[m];
    row[3] = a[m];
    row[2] = s[m] / delta;
    row[1] = 2.0 * b[m] / delta;
    row[0] = 3.0 * a[m] / delta;
    row[7] = 0.0;
  }
}

// setfl reader with the reference's token order (eam_alloy.cpp:86-201): three
// comment lines; "nel name..."; "nrho drho nr dr rcut"; per element
// "Z mass a0 lattice" then F(rho)[nrho] and rho(r)[nr]; then r*phi(r)[nr] for the
// packed lower triangle 11, 21, 22, 31, ...
void read_setfl(const char* path, int& nel, int& nrho, double& drho, int& nr, double& dr,
                double& rcut, std::vector<double>& frho, std::vector<double>& rhor,
                std::vector<double>& z2r, std::vector<double>& masses)
{
  std::ifstream in(path);
  if (!in) throw Error(XS_E_IO, std::string("cannot open setfl file ") + path);
  for (int i = 0; i < 3; ++i) in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
  long long ne = 0;
  in >> ne;
  if (!in || ne < 1 || ne > 7) throw Error(XS_E_IO, "setfl: bad element count");
  nel = (int)ne;
  for (int i = 0; i < nel; i++) {
    std::string name;
    in >> name;
  }
  long long a = 0, b = 0;
  in >> a >> drho >> b >> dr >> rcut;
  if (!in || a < 5 || b < 5) throw Error(XS_E_IO, "setfl: bad grid line");
  nrho = (int)a;
  nr = (int)b;
  frho.assign((size_t)nel * nrho, 0.0);
  rhor.assign((size_t)nel * nr, 0.0);
  masses.assign(nel, 0.0);
  for (int e = 0; e < nel; e++) {
    int z;
    double a0;
    std::string lat;
    in >> std::ws >> z >> masses[e] >> a0 >> lat;
    for (int k = 0; k < nrho; k++) in >> frho[(size_t)e * nrho + k];
    for (int k = 0; k < nr; k++) in >> rhor[(size_t)e * nr + k];
  }
  const int nz = nel * (nel + 1) / 2;
  z2r.assign((size_t)nz * nr, 0.0);
  for (int t = 0; t < nz; t++)
    for (int k = 0; k < nr; k++) in >> z2r[(size_t)t * nr + k];
  if (!in) throw Error(XS_E_IO, "setfl: file truncated");
}

namespace {

// [c0..c6,0] rows -> device order [c3 c4 c5 c6 c0 c1 c2 0]
template <class Real>
void reorder_rows(const std::vector<double>& src, std::vector<Real>& dst)
{
  const size_t rows = src.size() / 8;
  dst.resize(src.size());
  for (size_t r = 0; r < rows; r++) {
    const double* s = &src[8 * r];
    Real* d = &dst[8 * r];
    d[0] = (Real)s[3];
    d[1] = (Real)s[4];
    d[2] = (Real)s[5];
    d[3] = (Real)s[6];
    d[4] = (Real)s[0];
    d[5] = (Real)s[1];
    d[6] = (Real)s[2];
    d[7] = (Real)0;
  }
}

template <class Real>
void upload(Ctx& c, DBuf<Real>& buf, const std::vector<double>& coeffs)
{
  std::vector<Real> tmp;
  reorder_rows(coeffs, tmp);
  buf.reserve(tmp.size());
  XS_HIP(hipMemcpyAsync(buf.p, tmp.data(), tmp.size() * sizeof(Real), hipMemcpyHostToDevice, c.stream));
  XS_HIP(hipStreamSynchronize(c.stream));
}

// The reference spline of one table (FP64 rows [c0..c6, 0], row m = knot m) at r:
// value and slope, with eam_alloy.h's index and clamp steps.
void spline_at(const double* rows, int nr, double rdr, double r, double& v, double& d)
{
  double p = r * rdr + 1.0;
  const int m = std::max(0, std::min((int)p, nr - 1));
  p -= (double)m;
  p = std::min(p, 1.0);
  const double* c = rows + 8 * (size_t)m;
  v = ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
  d = (c[0] * p + c[1]) * p + c[2];
}

// rho(r), rho'(r)/r, phi(r) conv and phi'(r) conv / r of the single-element spline
// (phi = z2r / r) and the s-slopes (d/ds = d/dr / (2 r)) of rho and phi conv.
struct SFuncs {
  double rho, grho, phi, gphi, drho_ds, dphi_ds;
};
SFuncs s_funcs(const double* cr, const double* cz, int nr, double rdr, double r)
{
  double rv, rd, zv, zd;
  spline_at(cr, nr, rdr, r, rv, rd);
  spline_at(cz, nr, rdr, r, zv, zd);
  const double phi = zv / r * kEV;
  const double phip = (zd / r - zv / (r * r)) * kEV;
  return {rv, rd / r, phi, phip / r, rd / (2.0 * r), phip / (2.0 * r)};
}

// The FP32 s = r^2 tables (layout and the accuracy rule in pair_math.hpp) appended to
// out: the force plane, the slope quadratics of rho and phi conv times 2/ds
// [A B C 0 | A B C 0], then the phi plane, phi conv [a b c d].
void s_tables(Ctx& c, const double* cr, const double* cz, std::vector<float>& out)
{
  EamHost& e = c.eam;
  e.s_pieces = 0;
  for (int k = 0; k < 3; k++) e.s_err[k] = 0.0;
  if (c.s_table_disabled) return;
  const int nr = e.nr;
  const double rdr = 1.0 / e.dr, rmax = (nr - 1) * e.dr;
  const double s_lo = (0.3 * rmax) * (0.3 * rmax), s_hi = rmax * rmax;
  for (int npc = 256; npc <= 8192; npc *= 2) {
    const double ds = (s_hi - s_lo) / npc, rds = npc / (s_hi - s_lo), off = -s_lo * rds, g2 = 2.0 / ds;
    std::vector<SFuncs> node(npc + 1);
    for (int k = 0; k <= npc; k++) node[k] = s_funcs(cr, cz, nr, rdr, std::sqrt(s_lo + k * ds));
    std::vector<float> frc(8 * (size_t)npc), phi(4 * (size_t)npc);
    // Hermite cubic through (v0, d0 ds) and (v1, d1 ds) on p in [0, 1]: its value
    // coefficients and, times 2/ds, its p-slope quadratic
    auto hermite = [&](double v0, double v1, double d0, double d1, float* val, float* slope) {
      const double D0 = d0 * ds, D1 = d1 * ds;
      const double A = 2.0 * (v0 - v1) + D0 + D1, B = 3.0 * (v1 - v0) - 2.0 * D0 - D1;
      if (val) {
        val[0] = (float)A;
        val[1] = (float)B;
        val[2] = (float)D0;
        val[3] = (float)v0;
      }
      slope[0] = (float)(3.0 * A * g2);
      slope[1] = (float)(2.0 * B * g2);
      slope[2] = (float)(D0 * g2);
      slope[3] = 0.0f;
    };
    for (int k = 0; k < npc; k++) {
      const SFuncs &a = node[k], &b = node[k + 1];
      hermite(a.rho, b.rho, a.drho_ds, b.drho_ds, nullptr, &frc[8 * (size_t)k]);
      hermite(a.phi, b.phi, a.dphi_ds, b.dphi_ds, &phi[4 * (size_t)k], &frc[8 * (size_t)k + 4]);
    }
    // FP32 evaluation (the device's steps) against the FP64 spline, 8 points per piece
    double err[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (int k = 0; k < npc; k++)
      for (int q = 0; q < 8; q++) {
        const double sq = s_lo + (k + (q + 0.5) / 8.0) * ds;
        const SFuncs x = s_funcs(cr, cz, nr, rdr, std::sqrt(sq));
        const double xx = std::fma(sq, rds, off);
        const int m = std::min((int)xx, npc - 1);
        const float p = (float)(xx - m);
        const float* fv = &frc[8 * (size_t)m];
        const float* pv = &phi[4 * (size_t)m];
        const float got[3] = {std::fma(std::fma(fv[0], p, fv[1]), p, fv[2]),
                              std::fma(std::fma(std::fma(pv[0], p, pv[1]), p, pv[2]), p, pv[3]),
                              std::fma(std::fma(fv[4], p, fv[5]), p, fv[6])};
        const double want[3] = {x.grho, x.phi, x.gphi};
        for (int f = 0; f < 3; f++) {
          err[f] = std::max(err[f], std::abs((double)got[f] - want[f]));
          mx[f] = std::max(mx[f], std::abs(want[f]));
        }
      }
    const double ulp = std::ldexp(1.0, -24);
    double u[3];
    for (int f = 0; f < 3; f++) u[f] = mx[f] > 0.0 ? err[f] / (ulp * mx[f]) : 0.0;
    if (u[0] <= 8.0 && u[1] <= 4.0 && u[2] <= 8.0) {
      while (out.size() % 8) out.push_back(0.0f);  // 32-byte aligned planes
      e.s_pieces = npc;
      e.s_lo = s_lo;
      e.s_hi = s_hi;
      for (int f = 0; f < 3; f++) e.s_err[f] = u[f];
      out.insert(out.end(), frc.begin(), frc.end());
      out.insert(out.end(), phi.begin(), phi.end());
      return;
    }
  }
}

std::vector<double> build_all(int ntab, int n, double delta, const std::vector<double>& knots)
{
  std::vector<double> out((size_t)ntab * (n + 1) * 8, 0.0);
  for (int t = 0; t < ntab; t++)
    spline_interpolate(n, delta, knots.data() + (size_t)t * n, out.data() + (size_t)t * (n + 1) * 8);
  return out;
}

}  // namespace

// Type maps of initialize_types_table (eam_alloy.h:70-133) for nel types mapped
// to elements in order, then both precisions of every table to the device.
void eam_upload(Ctx& c)
{
  EamHost& e = c.eam;
  const int nel = e.nel;
  std::memset(e.t2rhor, 0, sizeof(e.t2rhor));
  std::memset(e.t2z2r, 0, sizeof(e.t2z2r));
  std::memset(e.t2frho, 0, sizeof(e.t2frho));
  for (int i = 1; i <= nel; i++) {
    e.t2frho[i] = (int8_t)(i - 1);
    for (int j = 1; j <= nel; j++) {
      e.t2rhor[i][j] = (int8_t)(i - 1);
      int hiv = i - 1, lov = j - 1;
      if (hiv < lov) std::swap(hiv, lov);
      e.t2z2r[i][j] = (int8_t)(hiv * (hiv + 1) / 2 + lov);
    }
  }
  const int nz = nel * (nel + 1) / 2;
  const std::vector<double> cf = build_all(nel, e.nrho, e.drho, e.frho_knots);
  const std::vector<double> cr = build_all(nel, e.nr, e.dr, e.rhor_knots);
  const std::vector<double> cz = build_all(nz, e.nr, e.dr, e.z2r_knots);
  upload(c, e.frho_d, cf);
  upload(c, e.rhor_d, cr);
  upload(c, e.z2r_d, cz);
  upload(c, e.frho_f, cf);
  upload(c, e.rhor_f, cr);
  upload(c, e.z2r_f, cz);

  // pair table of element 0 / pair 00 for single-element kernels: plane 0 holds the
  // rho value cubics [c3 c4 c5 c6] of rows 0..nr, plane 1 the z2r value cubics
  const size_t rows = (size_t)e.nr + 1;
  std::vector<double> pair(8 * rows);
  for (size_t m = 0; m < rows; m++)
    for (int k = 0; k < 4; k++) {
      pair[4 * m + k] = cr[8 * m + 3 + k];
      pair[4 * (rows + m) + k] = cz[8 * m + 3 + k];
    }
  auto up = [&](auto& buf, const auto& tmp) {
    buf.reserve(tmp.size());
    XS_HIP(hipMemcpyAsync(buf.p, tmp.data(), tmp.size() * sizeof(tmp[0]), hipMemcpyHostToDevice, c.stream));
    XS_HIP(hipStreamSynchronize(c.stream));
  };
  up(e.pair_d, pair);
  // FP32: the same knot planes, then the s = r^2 tables of the mixed passes
  std::vector<float> pf(pair.size());
  for (size_t q = 0; q < pair.size(); q++) pf[q] = (float)pair[q];
  s_tables(c, cr.data(), cz.data(), pf);
  up(e.pair_f, pf);
}

namespace {

// Slope s[m] (1-based knot m) of spline_interpolate from the knots f[0..n-1].
__device__ __forceinline__ double spline_slope_dev(const double* __restrict__ f, int n, int m)
{
  if (m == 1) return f[1] - f[0];
  if (m == 2) return 0.5 * (f[2] - f[0]);
  if (m == n - 1) return 0.5 * (f[n - 1] - f[n - 3]);
  if (m == n) return f[n - 1] - f[n - 2];
  return ((f[m - 3] - f[m + 1]) + 8.0 * (f[m] - f[m - 2])) / 12.0;
}

// One coefficient row per thread (row m of table t, rows 0..n): spline_interpolate's
// arithmetic in its order, with contraction off in this file, so every coefficient
// has the host build's bits; written in device order [c3 c4 c5 c6 c0 c1 c2 0] in
// FP64 and FP32.
__global__ void k_spline_rows(int ntab, int n, double delta, const double* __restrict__ knots,
                              double* __restrict__ rows_d, float* __restrict__ rows_f)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)n + 1;
  if (q >= (int64_t)ntab * per) return;
  const int t = (int)(q / per), m = (int)(q - t * per);
  const double* f = knots + (int64_t)t * n;
  double c[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (m >= 1) {
    const double sm = spline_slope_dev(f, n, m);
    double a = 0.0, b = 0.0;
    if (m <= n - 1) {
      const double sm1 = spline_slope_dev(f, n, m + 1);
      const double dv = f[m] - f[m - 1];
      b = 3.0 * dv - 2.0 * sm - sm1;
      a = sm + sm1 - 2.0 * dv;
    }
    c[0] = a;
    c[1] = b;
    c[2] = sm;
    c[3] = f[m - 1];
    c[4] = 3.0 * a / delta;
    c[5] = 2.0 * b / delta;
    c[6] = sm / delta;
  }
  double* d = rows_d + 8 * q;
  float* g = rows_f + 8 * q;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    d[k] = c[k];
    g[k] = (float)c[k];
  }
}

// The knot planes of the single-element pair table (rho and z2r value cubics of
// table 0, eam_upload's layout) from the device rows, both precisions.
__global__ void k_pair_planes(int64_t rows, const double* __restrict__ rhor, const double* __restrict__ z2r,
                              double* __restrict__ pair_d, float* __restrict__ pair_f)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 8 * rows) return;
  const int64_t plane = q / (4 * rows), m = (q >> 2) - plane * rows, k = q & 3;
  const double v = (plane == 0 ? rhor : z2r)[8 * m + k];
  pair_d[q] = v;
  pair_f[q] = (float)v;
}

void spline_rows_dev(Ctx& c, int ntab, int n, double delta, const double* knots, DBuf<double>& rows_d,
                     DBuf<float>& rows_f)
{
  const int64_t total = (int64_t)ntab * (n + 1);
  k_spline_rows<<<grid_for(total), kBlock, 0, c.stream>>>(ntab, n, delta, knots, rows_d.p, rows_f.p);
  XS_HIP(hipGetLastError());
}

}  // namespace

void eam_theta_dev(Ctx& c, const double* theta)
{
  EamHost& e = c.eam;
  const int nel = e.nel, nz = nel * (nel + 1) / 2;
  const double* f = theta;
  const double* r = f + (size_t)nel * e.nrho;
  const double* z = r + (size_t)nel * e.nr;
  spline_rows_dev(c, nel, e.nrho, e.drho, f, e.frho_d, e.frho_f);
  spline_rows_dev(c, nel, e.nr, e.dr, r, e.rhor_d, e.rhor_f);
  spline_rows_dev(c, nz, e.nr, e.dr, z, e.z2r_d, e.z2r_f);
  const int64_t rows = (int64_t)e.nr + 1;
  k_pair_planes<<<grid_for(8 * rows), kBlock, 0, c.stream>>>(rows, e.rhor_d.p, e.z2r_d.p, e.pair_d.p, e.pair_f.p);
  XS_HIP(hipGetLastError());
  // the s = r^2 tables were built (and checked) from the previous knots on the host
  e.s_pieces = 0;
  e.s_lo = e.s_hi = 0.0;
  e.s_err[0] = e.s_err[1] = e.s_err[2] = 0.0;
  c.rho_state = RhoState::None;
  c.rho_ckpt = false;
}

}  // namespace xsmd
