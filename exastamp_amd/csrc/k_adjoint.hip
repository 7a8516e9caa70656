// k_adjoint.hip -- reverse mode of the EAM force/energy with respect to the
// tabulated knots (no counterpart in exaStamp; the fitting capability built on
// its eam_alloy tables).
//
// For a loss L(F, E) with g_i = dL/dF_i and per-replica gE = dL/dE:
//   dL/dtheta = sum_i gE_i dE/dtheta - dD/dtheta,   D = sum_i g_i . dE/dx_i,
// because F_i = -dE/dx_i.  With u_ij = (x_i - x_j)/r and s_ij = u_ij . (g_i - g_j),
//   D = sum_i F'_i drho_i + 1/2 sum_i sum_j phi'_ij s_ij,   drho_i = sum_j rho'_tj(r_ij) s_ij.
// Differentiating through both EAM passes (rho_i inside F'_i is the mixed second
// derivative) gives, per atom and per pair of the full list,
//   F table at rho_i : value weight  gE_i conv,
//                      slope weight  conv (gE_i (rho_i - rhomax)^+ - drho_i)
//   rho table at r_ij: value weight  gE_i F'_i - lambda_i,  slope weight  -F'_i s_ij,
//                      lambda_i = F''(rho_i) drho_i
//   z2r table at r_ij: value weight  conv (gE_i / (2r) + s_ij / (2r^2)),
//                      slope weight  -conv s_ij / (2r)
// where "slope" multiplies the derivative of the spline with respect to r (or rho).
// rho_i and F'_i come from the forward pass (checkpointed, not recomputed).
//
// Each weight is spread on the four value coefficients c3..c6 of the spline row
// it falls in (the slope coefficients c0..c2 are 3c3/D, 2c4/D, c5/D), accumulated
// in FP64; a last kernel applies the transpose of the knot -> coefficient map.
//
// Accumulation.  All pairs of a shell fall on a few hundred rows, and all atoms on
// a few F rows, so atomics on the gradient rows in HBM serialise.  Instead a
// launch covers a set of row windows (slots) that fit in LDS: persistent
// workgroups walk all atoms (or pairs), add the contributions that fall in a window
// into their LDS copy (LDS FP64 atomics), and write it out as a partial; a second
// kernel sums the partials over the workgroups in a fixed order.  Windows cover
// every row of every table, one launch per group of windows.  The first launch walks
// the list for B1 (lambda_i stays in registers) and then B2 per atom; when every row
// of every table fits in it (small tables) that is the whole sweep, otherwise it
// stores lambda_i and the F-table weights for the F-row launches and the later B2
// launches.
#include "xsmd_internal.hpp"

#include <algorithm>

namespace xsmd {
namespace {

__device__ __forceinline__ int64_t lblock()
{
  const unsigned b = blockIdx.x, g = gridDim.x;
  return (int64_t)(b & 7u) * (g >> 3) + (b >> 3);
}

// Only the table storage follows Real: distances, knot fractions, weights and
// accumulations are FP64 (the sweeps are bound by gathers and atomics, and the
// knot fraction of an FP32 r would carry the rounding of r times the knot count).
template <class Real>
struct AdjTab {
  const Real* frho;  // reference layout rows [c3 c4 c5 c6 | c0 c1 c2 0]
  const Real* rhor;
  const Real* z2r;
  int nr, nrho;
  double rdr, rdrho, rhomax, conv;
  int8_t t2rhor[8][8];
  int8_t t2z2r[8][8];
  int8_t t2frho[8];
  uint32_t types;  // types slot of eam_alloy_force: bit t set = species t takes part
};

template <class Real>
__device__ __forceinline__ bool adj_on(const AdjTab<Real>& t, int type)
{
  return (t.types >> type) & 1u;
}

struct AdjArgs {
  int64_t n, npad;
  const Pos4* pos;        // x y z F'_i
  const Pos4* gvec;       // g_i (dL/dF_i) in cell order, w unused
  const double* rho;      // checkpointed rho_i
  const uint8_t* type;
  const uint16_t* rep;
  const BoxDev* box;
  const uint32_t* nbr;
  const int* nnb;
  const double* gE;       // per replica
  double rc2;
  double* lam;            // lambda_i (B1 out, B2 in)
  double* fc;             // F-table contribution of atom i (4 coefficients; B1 out)
  int* fm;                // its F-table row
  int nel;                // elements: table ids F 0..nel-1, rho nel..2nel-1, z2r 2nel..
  const XformDev* xf;     // physical displacements m * (r_j(image) - r_i) when not NULL
};

// LDS row windows of one launch: rows [lo, lo + rows) of table `table` at LDS
// rows off .. off + rows - 1
constexpr int kAdjSlots = 8;
constexpr int kAdjRowsMax = 5120;  // LDS rows per launch: 32 bytes each, 160 KB at most
constexpr int kAdjThreads = 1024;
struct AdjSlot {
  int table, lo, rows, off;
  int64_t g0;  // row of the gradient array that holds row lo of the table
};
struct AdjSlots {
  int n, total;
  int stride;  // rows of one workgroup's partial
  AdjSlot s[kAdjSlots];
};

// the four coefficient weights of a value weight wv and a slope weight ws at
// fraction p: v = c3 p^3 + c4 p^2 + c5 p + c6, dv/dx = (3 c3 p^2 + 2 c4 p + c5) rscale
__device__ __forceinline__ void coeff4(double wv, double ws, double p, double rscale, double (&o)[4])
{
  const double s = ws * rscale;
  o[0] = wv * p * p * p + s * 3.0 * p * p;
  o[1] = wv * p * p + s * 2.0 * p;
  o[2] = wv * p + s;
  o[3] = wv;
}

// add o to row m of table T if a window of this launch holds it
__device__ __forceinline__ void lds_add(double* acc, const AdjSlots& S, int T, int m, const double (&o)[4])
{
  for (int k = 0; k < S.n; k++) {
    const AdjSlot& w = S.s[k];
    if (w.table == T && m >= w.lo && m < w.lo + w.rows) {
      double* row = acc + 4 * (w.off + m - w.lo);
      atomicAdd(row + 0, o[0]);
      atomicAdd(row + 1, o[1]);
      atomicAdd(row + 2, o[2]);
      atomicAdd(row + 3, o[3]);
    }
  }
}

__device__ __forceinline__ void lds_zero(double* acc, int total)
{
  for (int q = threadIdx.x; q < 4 * total; q += blockDim.x) acc[q] = 0.0;
  __syncthreads();
}

__device__ __forceinline__ void lds_out(const double* acc, const AdjSlots& S, double* __restrict__ part)
{
  __syncthreads();
  double* dst = part + (int64_t)blockIdx.x * S.stride * 4;
  for (int q = threadIdx.x; q < 4 * S.total; q += blockDim.x) dst[q] = acc[q];
}

__device__ __forceinline__ void knot(double x, double rscale, int nmax1, int mmin, int& m, double& p,
                                     bool& clamped)
{
  p = x * rscale + 1.0;
  m = (int)p;
  m = max(mmin, min(m, nmax1));
  p -= (double)m;
  clamped = p > 1.0;
  p = fmin(p, 1.0);
}

template <class Real>
__device__ __forceinline__ double quad4(const Real* row, double p)
{
  return ((double)row[4] * p + (double)row[5]) * p + (double)row[6];
}

// B1 for atom i: drho_i and lambda_i, and the atom's F-table contribution o (four
// coefficient weights) at row m of F table T.  L lanes share the atom (consecutive
// lanes, sub = 0 .. L - 1, entries sub, sub + L, ...); every lane gets the results.
template <class Real, bool MULTI, int L>
__device__ __forceinline__ void adj_b1_atom(const AdjArgs& a, const AdjTab<Real>& t, int64_t i, int sub,
                                            double& lam, double (&o)[4], int& m, int& T)
{
  const int ti = MULTI ? a.type[i] : 0;
  const int rep = a.rep ? a.rep[i] : 0;
  const double* Lb = a.box[rep].L;
  const Pos4 pi = load_pos(a.pos, i);
  const Pos4 gi = load_pos(a.gvec, i);
  double drho = 0.0;
  const bool oni = !MULTI || adj_on(t, ti);
  const int cnt = oni ? a.nnb[i] : 0;
  for (int k = sub; k < cnt; k += L) {
    const uint32_t e = a.nbr[(int64_t)k * a.npad + i];
    const uint32_t j = xsmd_nbr_index(e);
    const Pos4 pj = load_pos(a.pos, j);
    double dx = xsmd_image(pj.x, xsmd_nbr_img_x(e), Lb[0]) - pi.x;
    double dy = xsmd_image(pj.y, xsmd_nbr_img_y(e), Lb[1]) - pi.y;
    double dz = xsmd_image(pj.z, xsmd_nbr_img_z(e), Lb[2]) - pi.z;
    if (a.xf) xsmd_xform_apply(a.xf->m, &dx, &dy, &dz);
    const double d2 = xsmd_d2(dx, dy, dz);
    if (d2 > a.rc2) continue;
    const Pos4 gj = load_pos(a.gvec, j);
    const double r = sqrt(d2);
    // s_ij = u_ij . (g_i - g_j), u_ij = -(dx,dy,dz)/r
    const double s = (dx * (gj.x - gi.x) + dy * (gj.y - gi.y) + dz * (gj.z - gi.z)) / r;
    const int tj = MULTI ? a.type[j] : 0;
    if (MULTI && !adj_on(t, tj)) continue;
    int mr;
    double p;
    bool cl;
    knot(r, t.rdr, t.nr - 1, 0, mr, p, cl);
    const Real* row = t.rhor + ((int64_t)t.t2rhor[tj + 1][ti + 1] * (t.nr + 1) + mr) * 8;
    drho += quad4(row, p) * s;
  }
#pragma unroll
  for (int off = 1; off < L; off <<= 1) drho += __shfl_xor(drho, off);
  const double rho = a.rho[i];
  double p;
  bool cl;
  knot(rho, t.rdrho, t.nrho - 1, 1, m, p, cl);
  T = t.t2frho[ti + 1];
  const Real* row = t.frho + ((int64_t)T * (t.nrho + 1) + m) * 8;
  // F'' = d/drho of (c0 p + c1) p + c2, zero where p was clamped
  const double fpp = cl ? 0.0 : (2.0 * (double)row[4] * p + (double)row[5]) * t.rdrho * t.conv;
  lam = fpp * drho;
  const double ge = a.gE[rep];
  const double conv = t.conv;
  const double over = rho > t.rhomax ? rho - t.rhomax : 0.0;
  coeff4(ge * conv, conv * (ge * over - drho), p, t.rdrho, o);
  if (!oni) {  // a disabled species has no embedding term (lambda and drho are 0 already)
    o[0] = o[1] = o[2] = o[3] = 0.0;
    lam = 0.0;
  }
}

// F-table rows: the per-atom contributions of B1 into this launch's windows
template <class Real, bool MULTI>
__global__ void __launch_bounds__(kAdjThreads) k_adj_fsum(AdjArgs a, AdjTab<Real> t, AdjSlots S,
                                                          double* __restrict__ part)
{
  extern __shared__ double acc[];
  lds_zero(acc, S.total);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * blockDim.x) {
    const int T = t.t2frho[(MULTI ? a.type[i] : 0) + 1];
    const Pos4 c = load_pos(reinterpret_cast<const Pos4*>(a.fc), i);
    const double o[4] = {c.x, c.y, c.z, c.w};
    lds_add(acc, S, T, a.fm[i], o);
  }
  lds_out(acc, S, part);
}

// B2: per pair, the rho- and z2r-table gradients.
template <class Real, bool MULTI, int L>
__device__ __forceinline__ void adj_b2_atom(const AdjArgs& a, const AdjTab<Real>& t, const AdjSlots& S,
                                            double* acc, int64_t i, int sub, double lam)
{
  const int ti = MULTI ? a.type[i] : 0;
  const int rep = a.rep ? a.rep[i] : 0;
  const double* Lb = a.box[rep].L;
  const Pos4 pi = load_pos(a.pos, i);
  const Pos4 gi = load_pos(a.gvec, i);
  const double fpi = pi.w;
  const double ge = a.gE[rep];
  const double conv = t.conv;
  const double rdr = t.rdr;
  const int cnt = (!MULTI || adj_on(t, ti)) ? a.nnb[i] : 0;
  for (int k = sub; k < cnt; k += L) {
    const uint32_t e = a.nbr[(int64_t)k * a.npad + i];
    const uint32_t j = xsmd_nbr_index(e);
    const Pos4 pj = load_pos(a.pos, j);
    double dx = xsmd_image(pj.x, xsmd_nbr_img_x(e), Lb[0]) - pi.x;
    double dy = xsmd_image(pj.y, xsmd_nbr_img_y(e), Lb[1]) - pi.y;
    double dz = xsmd_image(pj.z, xsmd_nbr_img_z(e), Lb[2]) - pi.z;
    if (a.xf) xsmd_xform_apply(a.xf->m, &dx, &dy, &dz);
    const double d2 = xsmd_d2(dx, dy, dz);
    if (d2 > a.rc2) continue;
    const Pos4 gj = load_pos(a.gvec, j);
    const double r = sqrt(d2);
    const double s = (dx * (gj.x - gi.x) + dy * (gj.y - gi.y) + dz * (gj.z - gi.z)) / r;
    const int tj = MULTI ? a.type[j] : 0;
    if (MULTI && !adj_on(t, tj)) continue;
    int m;
    double p;
    bool cl;
    knot(r, t.rdr, t.nr - 1, 0, m, p, cl);
    double o[4];
    coeff4(ge * fpi - lam, -fpi * s, p, rdr, o);
    lds_add(acc, S, a.nel + t.t2rhor[tj + 1][ti + 1], m, o);
    const double ir = 1.0 / r;
    coeff4(conv * (0.5 * ge * ir + 0.5 * s * ir * ir), -conv * 0.5 * s * ir, p, rdr, o);
    lds_add(acc, S, 2 * a.nel + t.t2z2r[ti + 1][tj + 1], m, o);
  }
}

template <class Real, bool MULTI>
__global__ void __launch_bounds__(kAdjThreads) k_adj_b2(AdjArgs a, AdjTab<Real> t, AdjSlots S,
                                                        double* __restrict__ part)
{
  extern __shared__ double acc[];
  lds_zero(acc, S.total);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * blockDim.x)
    adj_b2_atom<Real, MULTI, 1>(a, t, S, acc, i, 0, a.lam[i]);
  lds_out(acc, S, part);
}

// B1 and B2 in one launch, kFusedLanes lanes per atom walking its list twice: B1
// (lambda_i kept in registers), then B2 into this launch's windows.  When the rows of
// every table fit the launch (small tables, small systems) the F-table contribution
// goes into LDS too and this is the whole sweep; with STORE (the first of several
// launches) lambda_i and the F-table contribution are stored for k_adj_fsum and the
// later k_adj_b2 launches instead, so the list is walked once less.
constexpr int kFusedLanes = 4;
template <class Real, bool MULTI, bool STORE>
__global__ void __launch_bounds__(kAdjThreads) k_adj_fused(AdjArgs a, AdjTab<Real> t, AdjSlots S,
                                                           double* __restrict__ part)
{
  extern __shared__ double acc[];
  lds_zero(acc, S.total);
  const int sub = threadIdx.x % kFusedLanes;
  const int64_t step = (int64_t)gridDim.x * (blockDim.x / kFusedLanes);
  // the lanes of an atom share i, so they leave the loop together (their shuffles)
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kFusedLanes; i < a.n; i += step) {
    double lam, o[4];
    int m, T;
    adj_b1_atom<Real, MULTI, kFusedLanes>(a, t, i, sub, lam, o, m, T);
    if constexpr (STORE) {
      if (sub == 0) {
        a.lam[i] = lam;
        store_pos(reinterpret_cast<Pos4*>(a.fc), i, Pos4{o[0], o[1], o[2], o[3]});
        a.fm[i] = m;
      }
    } else {
      if (sub == 0) lds_add(acc, S, T, m, o);
    }
    adj_b2_atom<Real, MULTI, kFusedLanes>(a, t, S, acc, i, sub, lam);
  }
  lds_out(acc, S, part);
}

// Sum of the partials of `groups` workgroups into the gradient rows of this launch's
// windows (each row belongs to exactly one window of one launch).  A workgroup takes
// 64 consecutive values; its 16 waves sum the partials w = wave, wave + 16, ...
// (coalesced), and the 16 sums are added in wave order, so the result does not
// depend on scheduling.
constexpr int kRedThreads = 1024;
__global__ void __launch_bounds__(kRedThreads) k_adj_reduce(AdjSlots S, int groups, const double* __restrict__ part,
                                                            double* __restrict__ g)
{
  __shared__ double red[kRedThreads];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = blockIdx.x * 64 + lane;
  const bool in = q < 4 * S.total;
  double v = 0.0;
  if (in)
    for (int w = wave; w < groups; w += kRedThreads / 64) v += part[(int64_t)w * S.stride * 4 + q];
  red[threadIdx.x] = v;
  __syncthreads();
  if (wave != 0 || !in) return;
  v = red[lane];
  for (int k = 1; k < kRedThreads / 64; k++) v += red[k * 64 + lane];
  const int row = q >> 2;
  int k = 0;
  while (k + 1 < S.n && row >= S.s[k + 1].off) k++;
  g[(S.s[k].g0 + row - S.s[k].off) * 4 + (q & 3)] = v;
}
// g (dL/dF) from id order into cell order as position-like records
// (and times `scale`); the first nrep threads scale the per-replica dL/dE copy in place
__global__ void k_gather_g(int64_t n, const int64_t* __restrict__ id, const double* __restrict__ g_id,
                           Pos4* __restrict__ gvec, double scale, double* __restrict__ gE, int nrep)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nrep) gE[i] *= scale;
  if (i >= n) return;
  const int64_t q = id[i];
  store_pos(gvec, i, Pos4{g_id[3 * q] * scale, g_id[3 * q + 1] * scale, g_id[3 * q + 2] * scale, 0.0});
}

// ---------------------------------------------------------------------------
// Lennard-Jones adjoint with respect to (epsilon, sigma).
// With w(r) = e'(r)/r (the force factor, F_i = sum_j w(r_ij) (x_j - x_i)) and the
// pair energy 1/2 (e(r) - e(rc)) per list entry (pair_potential_impl.hxx:487-498,
// lennard_jones.h:40-50), for theta in {epsilon, sigma}:
//   dL/dtheta = sum over list entries (i, j), r_ij <= rc, of
//               1/2 gE_rep(i) (de/dtheta(r) - de/dtheta(rc)) + dw/dtheta(r) g_i.(x_j - x_i)
// with s6 = (sigma/r)^6, s12 = s6^2:
//   de/deps = 4 (s12 - s6),             de/dsigma = 4 eps (12 s12 - 6 s6) / sigma,
//   dw/deps = -24 (2 s12 - s6) / r^2,   dw/dsigma = -24 eps (24 s12 - 6 s6) / (sigma r^2).
// One thread per atom walks the global list in FP64; each workgroup writes its two
// sums, and one workgroup adds the partials in a fixed order (no atomics, so the
// result does not depend on scheduling).
// ---------------------------------------------------------------------------
struct LjAdjArgs {
  int64_t n, npad;
  const Pos4* pos;
  const Pos4* gvec;       // dL/dF_i in cell order
  const uint16_t* rep;
  const BoxDev* box;
  const uint32_t* nbr;
  const int* nnb;
  const double* gE;       // per replica
  double rc2, eps, sigma, sigma2;
  double decut_eps, decut_sigma;  // de/deps and de/dsigma at r = rc
  const XformDev* xf;
};

constexpr int kLjAdjThreads = 256;

// a and b summed over the workgroup in a fixed order (shuffle tree within each wave,
// then the waves in index order through red); every thread calls it and gets both sums
__device__ __forceinline__ void block_sum2(double a, double b, double* red, double& ra, double& rb)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[2 * w] = a;
    red[2 * w + 1] = b;
  }
  __syncthreads();
  ra = rb = 0.0;
  for (int k = 0; k < nw; k++) {
    ra += red[2 * k];
    rb += red[2 * k + 1];
  }
}

__global__ void __launch_bounds__(kLjAdjThreads) k_lj_adjoint(LjAdjArgs a, double* __restrict__ part)
{
  __shared__ double red[2 * (kLjAdjThreads / 64)];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double ge = 0.0, gs = 0.0;
  if (i < a.n) {
    const int rep = a.rep ? a.rep[i] : 0;
    const double* L = a.box[rep].L;
    const Pos4 pi = load_pos(a.pos, i);
    const Pos4 gi = load_pos(a.gvec, i);
    const double hE = 0.5 * a.gE[rep];
    const int cnt = a.nnb[i];
    for (int k = 0; k < cnt; k++) {
      const uint32_t e = a.nbr[(int64_t)k * a.npad + i];
      const Pos4 pj = load_pos(a.pos, xsmd_nbr_index(e));
      double dx = xsmd_image(pj.x, xsmd_nbr_img_x(e), L[0]) - pi.x;
      double dy = xsmd_image(pj.y, xsmd_nbr_img_y(e), L[1]) - pi.y;
      double dz = xsmd_image(pj.z, xsmd_nbr_img_z(e), L[2]) - pi.z;
      if (a.xf) xsmd_xform_apply(a.xf->m, &dx, &dy, &dz);
      const double d2 = xsmd_d2(dx, dy, dz);
      if (d2 > a.rc2) continue;
      const double ir2 = 1.0 / d2;
      const double s2 = a.sigma2 * ir2;
      const double s6 = s2 * s2 * s2;
      const double s12 = s6 * s6;
      const double gd = gi.x * dx + gi.y * dy + gi.z * dz;
      ge += hE * (4.0 * (s12 - s6) - a.decut_eps) - 24.0 * (2.0 * s12 - s6) * ir2 * gd;
      gs += hE * (4.0 * a.eps * (12.0 * s12 - 6.0 * s6) / a.sigma - a.decut_sigma) -
            24.0 * a.eps * (24.0 * s12 - 6.0 * s6) / a.sigma * ir2 * gd;
    }
  }
  double ra, rb;
  block_sum2(ge, gs, red, ra, rb);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ra;
    part[2 * blockIdx.x + 1] = rb;
  }
}

// out[0..1] = sum of the nparts partials, each thread a fixed stride of them
__global__ void __launch_bounds__(kLjAdjThreads) k_lj_adjoint_reduce(int64_t nparts, const double* __restrict__ part,
                                                                     double* __restrict__ out)
{
  __shared__ double red[2 * (kLjAdjThreads / 64)];
  double a = 0.0, b = 0.0;
  for (int64_t k = threadIdx.x; k < nparts; k += blockDim.x) {
    a += part[2 * k];
    b += part[2 * k + 1];
  }
  double ra, rb;
  block_sum2(a, b, red, ra, rb);
  if (threadIdx.x == 0) {
    out[0] = ra;
    out[1] = rb;
  }
}

template <class Real>
AdjTab<Real> adj_tab(Ctx& c);

template <>
AdjTab<double> adj_tab<double>(Ctx& c)
{
  AdjTab<double> t;
  t.frho = c.eam.frho_d.p;
  t.rhor = c.eam.rhor_d.p;
  t.z2r = c.eam.z2r_d.p;
  t.rdr = c.eam.rdr;
  t.rdrho = c.eam.rdrho;
  t.rhomax = c.eam.rhomax;
  t.conv = kEV;
  t.nr = c.eam.nr;
  t.nrho = c.eam.nrho;
  std::memcpy(t.t2rhor, c.eam.t2rhor, sizeof(t.t2rhor));
  std::memcpy(t.t2z2r, c.eam.t2z2r, sizeof(t.t2z2r));
  std::memcpy(t.t2frho, c.eam.t2frho, sizeof(t.t2frho));
  t.types = c.eam_types;
  return t;
}

template <>
AdjTab<float> adj_tab<float>(Ctx& c)
{
  AdjTab<float> t;
  t.frho = c.eam.frho_f.p;
  t.rhor = c.eam.rhor_f.p;
  t.z2r = c.eam.z2r_f.p;
  t.rdr = c.eam.rdr;
  t.rdrho = c.eam.rdrho;
  t.rhomax = c.eam.rhomax;
  t.conv = kEV;
  t.nr = c.eam.nr;
  t.nrho = c.eam.nrho;
  std::memcpy(t.t2rhor, c.eam.t2rhor, sizeof(t.t2rhor));
  std::memcpy(t.t2z2r, c.eam.t2z2r, sizeof(t.t2z2r));
  std::memcpy(t.t2frho, c.eam.t2frho, sizeof(t.t2frho));
  t.types = c.eam_types;
  return t;
}

// Row windows of every table (F tables use their rows_f rows, rho and z2r tables
// rows_r), packed into launches of at most `limit` LDS rows and kAdjSlots windows;
// `what` selects the F tables (0), the rho and z2r tables (1) or all of them (2).
std::vector<AdjSlots> adj_launches(int nel, int nz, int64_t rows_f, int64_t rows_r, int what, int limit)
{
  std::vector<AdjSlots> out;
  AdjSlots cur{};
  cur.stride = limit;
  auto add = [&](int T, int64_t rows, int64_t g0) {
    for (int64_t lo = 0; lo < rows; lo += limit) {
      const int w = (int)std::min<int64_t>(limit, rows - lo);
      if (cur.n == kAdjSlots || cur.total + w > limit) {
        out.push_back(cur);
        cur = AdjSlots{};
        cur.stride = limit;
      }
      cur.s[cur.n++] = AdjSlot{T, (int)lo, w, cur.total, g0 + lo};
      cur.total += w;
    }
  };
  const int64_t f_end = nel * rows_f;
  if (what != 1)
    for (int t = 0; t < nel; t++) add(t, rows_f, t * rows_f);
  if (what != 0) {
    for (int t = 0; t < nel; t++) add(nel + t, rows_r, f_end + t * rows_r);
    for (int t = 0; t < nz; t++) add(2 * nel + t, rows_r, f_end + (nel + t) * rows_r);
  }
  if (cur.n) out.push_back(cur);
  return out;
}

template <class Real>
void adj_run(Ctx& c, const AdjArgs& a)
{
  const AdjTab<Real> t = adj_tab<Real>(c);
  EamHost& e = c.eam;
  const int nel = e.nel, nz = nel * (nel + 1) / 2;
  const int64_t rows_f = e.nrho + 1, rows_r = e.nr + 1;
  const bool multi = eam_multi(c);
  // LDS windows as large as one workgroup may hold; persistent workgroups, as many as
  // fit on the CUs at the launch's LDS size (1024 threads each, 256 for systems too
  // small to give every CU 1024 atoms), but no more than the atoms fill
  const int limit = (int)std::min<size_t>(kAdjRowsMax, c.lds_max / (4 * sizeof(double)));
  const int cus = std::max(c.num_cu, 1);
  auto run = [&](auto k, AdjSlots S, int lanes) {
    const int64_t items = c.n * lanes;
    const int threads = items >= (int64_t)cus * kAdjThreads ? kAdjThreads : 256;
    const size_t lds = (size_t)S.total * 4 * sizeof(double);
    const int64_t per_cu = std::max<int64_t>(1, (int64_t)(c.lds_max / lds));
    const int groups = (int)std::max<int64_t>(1, std::min<int64_t>(cus * per_cu, (items + threads - 1) / threads));
    S.stride = S.total;
    c.adj_part.reserve((size_t)groups * S.stride * 4);
    XS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    k<<<groups, threads, lds, c.stream>>>(a, t, S, c.adj_part.p);
    XS_HIP(hipGetLastError());
    k_adj_reduce<<<(unsigned)((4 * S.total + 63) / 64), kRedThreads, 0, c.stream>>>(S, groups, c.adj_part.p,
                                                                                      c.adj_g.p);
    XS_HIP(hipGetLastError());
  };
  const std::vector<AdjSlots> all = adj_launches(nel, nz, rows_f, rows_r, 2, limit);
  if (all.size() == 1) {  // every row fits: one sweep of the list
    run(multi ? k_adj_fused<Real, true, false> : k_adj_fused<Real, false, false>, all[0], kFusedLanes);
    return;
  }
  // B1 with the first window set of the pair tables, then the F tables, then the rest
  const std::vector<AdjSlots> pairs = adj_launches(nel, nz, rows_f, rows_r, 1, limit);
  run(multi ? k_adj_fused<Real, true, true> : k_adj_fused<Real, false, true>, pairs[0], kFusedLanes);
  for (const AdjSlots& S : adj_launches(nel, nz, rows_f, rows_r, 0, limit))
    run(multi ? k_adj_fsum<Real, true> : k_adj_fsum<Real, false>, S, 1);
  for (size_t k = 1; k < pairs.size(); k++) run(multi ? k_adj_b2<Real, true> : k_adj_b2<Real, false>, pairs[k], 1);
}

}  // namespace

// Transpose of the knot -> coefficient map of spline_interpolate for one table:
// G holds dL/d(c3,c4,c5,c6) per row (rows 0..n), out receives dL/df for the n knots.
void spline_interpolate_adjoint(int n, double delta, const double* G, double* out)
{
  std::vector<double> gf(n + 2, 0.0), gs(n + 2, 0.0);
  (void)delta;  // slope coefficients were folded into c3..c5 on the device
  for (int m = 1; m <= n; m++) {
    const double* g = G + 4 * (size_t)m;
    gf[m] += g[3];  // c6 = f
    gs[m] += g[2];  // c5 = s
    if (m <= n - 1) {
      // c4 = 3 (f[m+1] - f[m]) - 2 s[m] - s[m+1]
      gf[m + 1] += 3.0 * g[1];
      gf[m] -= 3.0 * g[1];
      gs[m] -= 2.0 * g[1];
      gs[m + 1] -= g[1];
      // c3 = s[m] + s[m+1] - 2 (f[m+1] - f[m])
      gs[m] += g[0];
      gs[m + 1] += g[0];
      gf[m + 1] -= 2.0 * g[0];
      gf[m] += 2.0 * g[0];
    }
  }
  // slopes
  gf[2] += gs[1];
  gf[1] -= gs[1];
  gf[3] += 0.5 * gs[2];
  gf[1] -= 0.5 * gs[2];
  gf[n] += 0.5 * gs[n - 1];
  gf[n - 2] -= 0.5 * gs[n - 1];
  gf[n] += gs[n];
  gf[n - 1] -= gs[n];
  for (int m = 3; m <= n - 2; m++) {
    const double w = gs[m] / 12.0;
    gf[m - 2] += w;
    gf[m + 2] -= w;
    gf[m + 1] += 8.0 * w;
    gf[m - 1] -= 8.0 * w;
  }
  for (int m = 1; m <= n; m++) out[m - 1] = gf[m];
}

namespace {

// Gather form of spline_interpolate_adjoint for every table at once: one thread per
// knot k (1-based) of table t sums the rows' value-coefficient gradients G (4 per
// row, dL/d(c3, c4, c5, c6)) that depend on knot k, directly (c6 = f, and f in
// c3, c4 of rows k and k - 1) and through the slopes s[m] of rows m = k - 2 .. k + 2.
__device__ __forceinline__ double adj_gs(const double* __restrict__ G, int n, int m)
{
  // dL/ds[m]: c5 of row m, c4 and c3 of row m (s[m]) and of row m - 1 (s[m + 1] there)
  double v = G[4 * m + 2];
  if (m <= n - 1) v += -2.0 * G[4 * m + 1] + G[4 * m + 0];
  if (m >= 2) v += -G[4 * (m - 1) + 1] + G[4 * (m - 1) + 0];
  return v;
}

__global__ void k_spline_adjoint(int nel, int nz, int nrho, int nr, const double* __restrict__ G,
                                 double* __restrict__ grad)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t kf = (int64_t)nel * nrho, kr = (int64_t)nel * nr, kz = (int64_t)nz * nr;
  if (q >= kf + kr + kz) return;
  int n;
  int64_t t, k0;
  const double* g;
  if (q < kf) {
    n = nrho;
    t = q / nrho;
    k0 = q - t * nrho;
    g = G + t * (nrho + 1) * 4;
  } else {
    const int64_t u = q - kf;
    n = nr;
    t = u / nr;
    k0 = u - t * nr;
    g = G + (kf + (int64_t)nel) * 4 + t * (nr + 1) * 4;  // the F tables' rows, then nel + nz tables of nr + 1
  }
  const int k = (int)k0 + 1;
  // direct: c6 = f[k]; c4, c3 of row k (-3, +2) and of row k - 1 (+3, -2)
  double v = g[4 * k + 3];
  if (k <= n - 1) v += -3.0 * g[4 * k + 1] + 2.0 * g[4 * k + 0];
  if (k >= 2) v += 3.0 * g[4 * (k - 1) + 1] - 2.0 * g[4 * (k - 1) + 0];
  // through the slopes
  if (k == 2) v += adj_gs(g, n, 1);
  if (k == 1) v -= adj_gs(g, n, 1) + 0.5 * adj_gs(g, n, 2);
  if (k == 3) v += 0.5 * adj_gs(g, n, 2);
  if (k == n) v += 0.5 * adj_gs(g, n, n - 1) + adj_gs(g, n, n);
  if (k == n - 2) v -= 0.5 * adj_gs(g, n, n - 1);
  if (k == n - 1) v -= adj_gs(g, n, n);
  // interior slopes m in [3, n - 2]: f[m - 2] +w, f[m + 2] -w, f[m + 1] +8w, f[m - 1] -8w
  const auto interior = [&](int m) { return m >= 3 && m <= n - 2; };
  if (interior(k + 2)) v += adj_gs(g, n, k + 2) / 12.0;
  if (interior(k - 2)) v -= adj_gs(g, n, k - 2) / 12.0;
  if (interior(k - 1)) v += 8.0 * (adj_gs(g, n, k - 1) / 12.0);
  if (interior(k + 1)) v -= 8.0 * (adj_gs(g, n, k + 1) / 12.0);
  grad[q] = v;
}

}  // namespace

void eam_adjoint(Ctx& c, double rcut, const double* dLdF_id, const double* dLdE, double scale, double* grad,
                 bool mixed, bool device)
{
  ensure_global_list(c);
  EamHost& e = c.eam;
  const int nel = e.nel, nz = nel * (nel + 1) / 2;
  const size_t rows_r = (size_t)e.nr + 1, rows_f = (size_t)e.nrho + 1;
  const size_t nF = (size_t)nel * rows_f * 4, nR = (size_t)nel * rows_r * 4, nZ = (size_t)nz * rows_r * 4;
  const size_t nknots = (size_t)nel * e.nrho + (size_t)(nel + nz) * e.nr;
  c.adj_g.reserve(nF + nR + nZ);
  c.adj_lam.reserve(c.npad);
  c.adj_gvec.reserve(c.npad);
  c.adj_in.reserve(3 * (size_t)c.n + c.nrep);
  if (!device) c.adj_grad.reserve(nknots);
  const hipMemcpyKind in_kind = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  XS_HIP(hipMemsetAsync(c.adj_g.p, 0, sizeof(double) * (nF + nR + nZ), c.stream));
  XS_HIP(hipMemcpyAsync(c.adj_in.p, dLdF_id, sizeof(double) * 3 * c.n, in_kind, c.stream));
  XS_HIP(hipMemcpyAsync(c.adj_in.p + 3 * c.n, dLdE, sizeof(double) * c.nrep, in_kind, c.stream));
  if (c.n > 0) {
    k_gather_g<<<grid_for(std::max<int64_t>(c.n, c.nrep)), kBlock, 0, c.stream>>>(
        c.n, c.id.p, c.adj_in.p, c.adj_gvec.p, scale, c.adj_in.p + 3 * c.n, c.nrep);
    XS_HIP(hipGetLastError());
    AdjArgs a;
    a.n = c.n;
    a.npad = c.npad;
    a.pos = c.pos.p;
    a.gvec = c.adj_gvec.p;
    a.rho = c.rho.p;
    a.type = c.type.p;
    a.rep = c.has_rep ? c.rep.p : nullptr;
    a.box = c.box_d.p;
    a.nbr = c.nbr.p;
    a.nnb = c.nnb.p;
    a.gE = c.adj_in.p + 3 * c.n;
    a.rc2 = rcut * rcut;
    a.lam = c.adj_lam.p;
    c.adj_fc.reserve(4 * (size_t)c.npad);
    c.adj_fm.reserve(c.npad);
    a.fc = c.adj_fc.p;
    a.fm = c.adj_fm.p;
    a.nel = nel;
    a.xf = c.xf_dev();
    ProfScope prof(c.prof, c.stream, XS_PROF_ADJOINT);
    if (mixed) adj_run<float>(c, a);
    else adj_run<double>(c, a);
  }
  // dL/dtheta: the transpose of the knot -> coefficient map, on the device
  double* out = device ? grad : c.adj_grad.p;
  k_spline_adjoint<<<grid_for((int64_t)nknots), kBlock, 0, c.stream>>>(nel, nz, e.nrho, e.nr, c.adj_g.p, out);
  XS_HIP(hipGetLastError());
  if (!device) {
    XS_HIP(hipMemcpyAsync(grad, out, sizeof(double) * nknots, hipMemcpyDeviceToHost, c.stream));
    XS_HIP(hipStreamSynchronize(c.stream));
  }
}

void lj_adjoint(Ctx& c, double eps, double sigma, double rcut, const double* dLdF_id, const double* dLdE,
                double grad[2])
{
  ensure_global_list(c);
  const unsigned g = grid_for(c.n, kLjAdjThreads);
  c.adj_gvec.reserve(c.npad);
  c.adj_in.reserve(3 * (size_t)c.n + c.nrep);
  c.adj_part.reserve(2 * (size_t)std::max(g, 1u) + 2);
  XS_HIP(hipMemcpyAsync(c.adj_in.p, dLdF_id, sizeof(double) * 3 * c.n, hipMemcpyHostToDevice, c.stream));
  XS_HIP(hipMemcpyAsync(c.adj_in.p + 3 * c.n, dLdE, sizeof(double) * c.nrep, hipMemcpyHostToDevice, c.stream));
  double* out = c.adj_part.p + 2 * (size_t)std::max(g, 1u);
  if (c.n > 0) {
    k_gather_g<<<grid_for(c.n), kBlock, 0, c.stream>>>(c.n, c.id.p, c.adj_in.p, c.adj_gvec.p, 1.0, nullptr, 0);
    XS_HIP(hipGetLastError());
    LjAdjArgs a;
    a.n = c.n;
    a.npad = c.npad;
    a.pos = c.pos.p;
    a.gvec = c.adj_gvec.p;
    a.rep = c.has_rep ? c.rep.p : nullptr;
    a.box = c.box_d.p;
    a.nbr = c.nbr.p;
    a.nnb = c.nnb.p;
    a.gE = c.adj_in.p + 3 * c.n;
    a.rc2 = rcut * rcut;
    a.eps = eps;
    a.sigma = sigma;
    a.sigma2 = sigma * sigma;
    a.xf = c.xf_dev();
    // the derivatives of energy_cutoff = e(rc) (pair_potential_impl.hxx:487-498)
    const double sc2 = a.sigma2 / a.rc2, sc6 = sc2 * sc2 * sc2, sc12 = sc6 * sc6;
    a.decut_eps = 4.0 * (sc12 - sc6);
    a.decut_sigma = 4.0 * eps * (12.0 * sc12 - 6.0 * sc6) / sigma;
    ProfScope prof(c.prof, c.stream, XS_PROF_ADJOINT);
    k_lj_adjoint<<<g, kLjAdjThreads, 0, c.stream>>>(a, c.adj_part.p);
    XS_HIP(hipGetLastError());
    k_lj_adjoint_reduce<<<1, kLjAdjThreads, 0, c.stream>>>(g, c.adj_part.p, out);
    XS_HIP(hipGetLastError());
    XS_HIP(hipMemcpyAsync(grad, out, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  } else {
    grad[0] = grad[1] = 0.0;
  }
  XS_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace xsmd
