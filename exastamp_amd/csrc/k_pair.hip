// k_pair.hip -- Lennard-Jones and EAM pair kernels over the global neighbour list.
//
// One thread per central atom walks its ELL list (coalesced entry loads), gathers
// neighbour position records, and accumulates its own force only: no atomics, so
// results do not depend on scheduling.  The exception is the Newton-symmetric
// operator form (XS_SYMMETRIC) at the end of the file, which walks half of the list
// and adds each partner's share with FP64 atomics.
//
// The neighbour loop is branch-free and batched: U entries are loaded, then U
// position records (one 32-byte sector each), then U pairs are evaluated with
// the cutoff test applied as a select.  Lanes of a wavefront diverge on the
// cutoff anyway, so computing the out-of-range entries costs nothing extra, and
// the batch keeps several gathers in flight per wavefront.
//
// With a cell-block list (the default) the LJ and EAM operators run k_block.hip
// instead; these kernels serve the global list (XSMD_LIST=global, or no brick that
// fits in LDS) and the Newton-symmetric operators.
#include "pair_math.hpp"

namespace xsmd {
namespace {

using namespace pm;

struct PairCommon {
  int64_t n, npad;
  const Pos4* __restrict__ pos;
  const uint8_t* __restrict__ type;
  const uint16_t* __restrict__ rep;
  const BoxDev* __restrict__ box;
  const uint32_t* __restrict__ nbr;
  const int* __restrict__ nnb;
  double rc2;
  const XformDev* xf;  // displacements m * (r_j(image) - r_i) when not NULL
};

constexpr int kBatch = 4;

// One neighbour: FP64 displacement r_j(image) - r_i, squared norm, the neighbour's
// record scalar w, and whether it is a real entry inside the cutoff.
struct Nb {
  double x, y, z, d2, w;
  uint32_t j;
  bool in;
};

// Load entries k0..k0+U-1 (missing ones become the atom itself, masked out) and
// their position records, then form the displacements.
template <int U>
__device__ __forceinline__ void fetch(const PairCommon& pc, int64_t i, int k0, int cnt, const Pos4& pi,
                                      const double* L, Nb (&nb)[U])
{
  uint32_t e[U];
#pragma unroll
  for (int u = 0; u < U; u++) e[u] = (k0 + u < cnt) ? pc.nbr[(int64_t)(k0 + u) * pc.npad + i] : (uint32_t)i;
  Pos4 pj[U];
#pragma unroll
  for (int u = 0; u < U; u++) pj[u] = load_pos(pc.pos, xsmd_nbr_index(e[u]));
#pragma unroll
  for (int u = 0; u < U; u++) {
    double xj = pj[u].x, yj = pj[u].y, zj = pj[u].z;
    if (e[u] >> XSMD_NBR_INDEX_BITS) {
      xj = xsmd_image(xj, xsmd_nbr_img_x(e[u]), L[0]);
      yj = xsmd_image(yj, xsmd_nbr_img_y(e[u]), L[1]);
      zj = xsmd_image(zj, xsmd_nbr_img_z(e[u]), L[2]);
    }
    nb[u].x = xj - pi.x;
    nb[u].y = yj - pi.y;
    nb[u].z = zj - pi.z;
    if (pc.xf) xsmd_xform_apply(pc.xf->m, &nb[u].x, &nb[u].y, &nb[u].z);
    nb[u].d2 = xsmd_d2(nb[u].x, nb[u].y, nb[u].z);
    nb[u].w = pj[u].w;
    nb[u].j = xsmd_nbr_index(e[u]);
    nb[u].in = (k0 + u < cnt) && nb[u].d2 <= pc.rc2;
  }
}

// ------------------------------------------------------------------------------
// Lennard-Jones: lj_compute_energy (lennard_jones.h:40-50) with the energy shift
// and per-atom accumulation of pair_potential_force_op_singlemat.h:101-123.
// Written in r^2 only: (sigma/r)^2 = sigma^2 / d2 and de/r = -24 eps (2 s^12 - s^6) / d2.
// ------------------------------------------------------------------------------
template <class Real>
using LJParams = pm::LJ<Real>;

// MULTI (lj_multi_force): each pair takes the parameters and cutoff of its two species.
template <class Real, bool ENERGY, bool VIRIAL, bool FUSED, bool MULTI = false>
__global__ void __launch_bounds__(256) k_lj(PairCommon pc, LJParams<Real> lp, Outputs o, pm::LJMultiDev lm = {})
{
  const int64_t i = logical_block() * blockDim.x + threadIdx.x;
  if (i >= pc.n) return;
  const double* L = pc.box[pc.rep ? pc.rep[i] : 0].L;
  const Pos4 pi = load_pos(pc.pos, i);
  const int ti = MULTI ? pc.type[i] : 0;
  double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int cnt = pc.nnb[i];
  for (int k0 = 0; k0 < cnt; k0 += kBatch) {
    Nb nb[kBatch];
    fetch<kBatch>(pc, i, k0, cnt, pi, L, nb);
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      double de_r, eh;
      if constexpr (MULTI) {
        const int pid = pm::type_pair(ti, pc.type[nb[u].j]);
        pm::lj_pair(pm::lj_multi_row<Real>(lm, pid), nb[u].d2, nb[u].in && pm::lj_multi_in(lm, pid, nb[u].d2),
                    de_r, eh);
      } else {
        pm::lj_pair(lp, nb[u].d2, nb[u].in, de_r, eh);
      }
      const double fe0 = de_r * nb[u].x, fe1 = de_r * nb[u].y, fe2 = de_r * nb[u].z;
      fx += fe0;
      fy += fe1;
      fz += fe2;
      if constexpr (ENERGY) e += eh;
      if constexpr (VIRIAL) add_virial(v, fe0, fe1, fe2, nb[u].x, nb[u].y, nb[u].z);
    }
  }
  write_outputs<ENERGY, VIRIAL, FUSED>(o, i, fx, fy, fz, e, v);
}

// ------------------------------------------------------------------------------
// EAM (eam_alloy): pass semantics of SymRhoOp / Rho2EmbOp / SymForceOp with
// NewtonSym = false (eam_force_op_multimat.h:114-343).
// ------------------------------------------------------------------------------

// Pass 1 (+ fused pass 1b): rho_i = sum_j rho_{type_j}(r_ij) in FP64 (kept as the
// adjoint's checkpoint), then F'(rho_i) into the position record's w (and F(rho_i)
// into ep when ENERGY).
template <class Real, bool MULTI, bool RHO, bool EMB, bool ENERGY, bool SET_EP>
__global__ void __launch_bounds__(256) k_eam_rho(PairCommon pc, EamTabDev<Real> t, const Real* __restrict__ pair,
                                                 double* __restrict__ rho_io, double* __restrict__ pos_w,
                                                 double* __restrict__ ep)
{
  const int64_t i = logical_block() * blockDim.x + threadIdx.x;
  if (i >= pc.n) return;
  const int ti = MULTI ? pc.type[i] : 0;
  const bool oni = !MULTI || eam_on(t, ti);
  double rho = 0.0;
  if constexpr (RHO) {
    const double* L = pc.box[pc.rep ? pc.rep[i] : 0].L;
    const Pos4 pi = load_pos(pc.pos, i);
    const int cnt = oni ? pc.nnb[i] : 0;
    for (int k0 = 0; k0 < cnt; k0 += kBatch) {
      Nb nb[kBatch];
      fetch<kBatch>(pc, i, k0, cnt, pi, L, nb);
#pragma unroll
      for (int u = 0; u < kBatch; u++) {
        const int tj = MULTI ? pc.type[nb[u].j] : 0;
        if constexpr (MULTI) nb[u].in = nb[u].in && eam_on(t, tj);
        const Real val = eam_rho_knot<Real, MULTI>(t, pair, ti, tj, nb[u].d2, nb[u].in);
        rho += nb[u].in ? (double)val : 0.0;
      }
    }
    rho_io[i] = rho;
  } else {
    rho = rho_io[i];
  }
  if constexpr (EMB) {
    Real F, Fp;
    embed(t, ti, rho, F, Fp);
    if (!oni) F = Fp = Real(0);  // a disabled species keeps rho_dEmb = 0 and adds no F
    pos_w[4 * i + 3] = (double)Fp;
    if constexpr (ENERGY) {
      if constexpr (SET_EP) ep[i] = (double)F;
      else ep[i] += (double)F;
    }
  }
}

// Pass 2: f_i += psi'/r (r_j - r_i), with the per-atom epilogue.
template <class Real, bool MULTI, bool ENERGY, bool VIRIAL, bool FUSED>
__global__ void __launch_bounds__(256) k_eam_force(PairCommon pc, EamTabDev<Real> t, const Real* __restrict__ pair,
                                                   Outputs o)
{
  const int64_t i = logical_block() * blockDim.x + threadIdx.x;
  if (i >= pc.n) return;
  const int ti = MULTI ? pc.type[i] : 0;
  const double* L = pc.box[pc.rep ? pc.rep[i] : 0].L;
  const Pos4 pi = load_pos(pc.pos, i);
  const Real fpi = (Real)pi.w;
  double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int cnt = (!MULTI || eam_on(t, ti)) ? pc.nnb[i] : 0;
  for (int k0 = 0; k0 < cnt; k0 += kBatch) {
    Nb nb[kBatch];
    fetch<kBatch>(pc, i, k0, cnt, pi, L, nb);
    int tj[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      tj[u] = MULTI ? pc.type[nb[u].j] : 0;
      if constexpr (MULTI) nb[u].in = nb[u].in && eam_on(t, tj[u]);
    }
    Real fp_r[kBatch], phi[kBatch];
    bool slow[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) phi[u] = Real(0);
    eam_pair_batch<Real, MULTI, ENERGY, kBatch>(t, pair, ti, tj, fpi, nb, fp_r, phi, slow);
    if constexpr (kSTable<Real, MULTI>) {
#pragma unroll
      for (int u = 0; u < kBatch; u++)
        if (slow[u]) eam_pair_knot<Real, false>(t, pair, 0, 0, fpi, (Real)nb[u].w, nb[u].d2, true, fp_r[u], phi[u]);
    }
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      const double fpair = nb[u].in ? (double)fp_r[u] : 0.0;
      const double fe0 = nb[u].x * fpair, fe1 = nb[u].y * fpair, fe2 = nb[u].z * fpair;
      fx += fe0;
      fy += fe1;
      fz += fe2;
      if constexpr (ENERGY) e += nb[u].in ? 0.5 * (double)phi[u] : 0.0;
      if constexpr (VIRIAL) add_virial(v, fe0, fe1, fe2, nb[u].x, nb[u].y, nb[u].z);
    }
  }
  write_outputs<ENERGY, VIRIAL, FUSED>(o, i, fx, fy, fz, e, v);
}

// ------------------------------------------------------------------------------
// Analytic Johnson EAM (johnson_force): the single-material operator template
// (eam_potential.cu:62-193; EmbOp and ForceOp, eam_force_op_singlemat.h:39-171) with
// the functions of johnson.h:59-167.  FP64 throughout; the energy terms are always
// accumulated, as EmbOp and ForceOp do.
// ------------------------------------------------------------------------------

// One screened exponential c exp(-a (x - 1)) / (1 + (x - k)^20) of x = r / re and its
// r-derivative by the quotient rule: phi is the sum of two such terms (johnson.h:59-99),
// rho(r) is one (johnson.h:105-124).
__device__ __forceinline__ void johnson_term(double c, double a, double k, double x, double ire, double& v,
                                             double& dv)
{
  const double num = c * exp(-a * (x - 1.0));
  const double s = x - k;
  const double s20 = pow(s, 20.0);
  const double den = 1.0 + s20;
  const double dnum = -a * num;
  const double dden = 20.0 * s20 / s;
  v = num / den;
  dv = ire * (dnum * den - num * dden) / (den * den);
}

__device__ __forceinline__ void johnson_rho(const JohnsonParams& p, double r, double& rho, double& drho)
{
  johnson_term(p.fe, p.beta, p.lambda, r / p.re, 1.0 / p.re, rho, drho);
}

__device__ __forceinline__ void johnson_phi(const JohnsonParams& p, double r, double& phi, double& dphi)
{
  const double x = r / p.re, ire = 1.0 / p.re;
  double v2, d2;
  johnson_term(p.A, p.alpha, p.kappa, x, ire, phi, dphi);
  johnson_term(-p.B, p.beta, p.lambda, x, ire, v2, d2);
  phi += v2;
  dphi += d2;
}

// F(rho) and F'(rho) (johnson.h:131-167): cubic in rho / (0.85 rhoe) - 1 below
// 0.85 rhoe, cubic in rho / rhoe - 1 below 1.15 rhoe, Fo (1 - ln q^eta) q^eta above.
__device__ __forceinline__ void johnson_embed(const JohnsonParams& p, double rho, double& f, double& df)
{
  const double rhon = 0.85 * p.rhoe, rho0 = 1.15 * p.rhoe;
  const double irhoe = 1.0 / p.rhoe;
  if (rho < rhon) {
    const double t = rho / rhon - 1.0, t2 = t * t, t3 = t * t2;
    f = p.Fn0 + p.Fn1 * t + p.Fn2 * t2 + p.Fn3 * t3;
    df = (p.Fn1 + 2.0 * p.Fn2 * t + 3.0 * p.Fn3 * t2) * (1.0 / rhon);
  } else if (rho < rho0) {
    const double t = rho * irhoe - 1.0, t2 = t * t, t3 = t * t2;
    f = p.F0 + p.F1 * t + p.F2 * t2 + p.F3 * t3;
    df = (p.F1 + 2.0 * p.F2 * t + 3.0 * p.F3 * t2) * irhoe;
  } else {
    const double q = rho * irhoe;
    const double qe = pow(q, p.eta);
    const double lq = log(qe);
    f = p.Fo * (1.0 - lq) * qe;
    df = (-p.eta * qe + (1.0 - lq) * p.eta * qe) * (p.Fo * irhoe / q);
  }
}

// Pass 1 (EmbOp): rho_i over the pairs inside rc, then F'(rho_i) into the position
// record's w and F(rho_i) added to ep.
__global__ void __launch_bounds__(256) k_johnson_rho(PairCommon pc, JohnsonParams jp, double* __restrict__ rho_out,
                                                     double* __restrict__ pos_w, double* __restrict__ ep)
{
  const int64_t i = logical_block() * blockDim.x + threadIdx.x;
  if (i >= pc.n) return;
  const double* L = pc.box[pc.rep ? pc.rep[i] : 0].L;
  const Pos4 pi = load_pos(pc.pos, i);
  const int cnt = pc.nnb[i];
  double rho = 0.0;
  for (int k0 = 0; k0 < cnt; k0 += kBatch) {
    Nb nb[kBatch];
    fetch<kBatch>(pc, i, k0, cnt, pi, L, nb);
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      double v, dv;
      johnson_rho(jp, nb[u].in ? sqrt(nb[u].d2) : jp.re, v, dv);
      rho += nb[u].in ? v : 0.0;
    }
  }
  rho_out[i] = rho;
  double F, Fp;
  johnson_embed(jp, rho, F, Fp);
  pos_w[4 * i + 3] = Fp;
  ep[i] += F;
}

// Pass 2 (ForceOp): f_i += (rho'(r) (F'_i + F'_j) + phi'(r)) / r (r_j - r_i), ep += phi / 2,
// virial -1/2 f (x) dr.
template <bool VIRIAL>
__global__ void __launch_bounds__(256) k_johnson_force(PairCommon pc, JohnsonParams jp, Outputs o)
{
  const int64_t i = logical_block() * blockDim.x + threadIdx.x;
  if (i >= pc.n) return;
  const double* L = pc.box[pc.rep ? pc.rep[i] : 0].L;
  const Pos4 pi = load_pos(pc.pos, i);
  double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int cnt = pc.nnb[i];
  for (int k0 = 0; k0 < cnt; k0 += kBatch) {
    Nb nb[kBatch];
    fetch<kBatch>(pc, i, k0, cnt, pi, L, nb);
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      const double r = nb[u].in ? sqrt(nb[u].d2) : jp.re;
      double rho, drho, phi, dphi;
      johnson_rho(jp, r, rho, drho);
      johnson_phi(jp, r, phi, dphi);
      const double de = nb[u].in ? (drho * (pi.w + nb[u].w) + dphi) / r : 0.0;
      const double fe0 = de * nb[u].x, fe1 = de * nb[u].y, fe2 = de * nb[u].z;
      fx += fe0;
      fy += fe1;
      fz += fe2;
      e += nb[u].in ? 0.5 * phi : 0.0;
      if constexpr (VIRIAL) add_virial(v, fe0, fe1, fe2, nb[u].x, nb[u].y, nb[u].z);
    }
  }
  write_outputs<true, VIRIAL, false>(o, i, fx, fy, fz, e, v);
}

PairCommon make_common(Ctx& c, double rcut)
{
  PairCommon pc;
  pc.n = c.n;
  pc.npad = c.npad;
  pc.pos = c.pos.p;
  pc.type = c.type.p;
  pc.rep = c.has_rep ? c.rep.p : nullptr;
  pc.box = c.box_d.p;
  pc.nbr = c.nbr.p;
  pc.nnb = c.nnb.p;
  pc.rc2 = rcut * rcut;
  pc.xf = c.xf_dev();
  return pc;
}

unsigned pair_grid(int64_t n)
{
  const unsigned g = grid_for(n);
  return (g + 7u) / 8u * 8u;
}

template <class Real, bool MULTI>
void eam_dispatch(Ctx& c, double rcut, bool do_rho, bool do_emb, bool do_force, unsigned flags,
                  const ForceEpilogue& epi)
{
  const EamTabDev<Real> t = make_eam_tab<Real>(c);
  const Real* pair = eam_pair_table<Real>(c);
  const PairCommon pc = make_common(c, rcut);
  const unsigned g = pair_grid(c.n);
  const bool energy = flags & XS_ENERGY;
  double* pos_w = reinterpret_cast<double*>(c.pos.p);

  if (do_rho || do_emb) {
    ProfScope prof(c.prof, c.stream, XS_PROF_EAM_RHO);
#define XS_RHO(R, E, EN, SET) \
  k_eam_rho<Real, MULTI, R, E, EN, SET><<<g, kBlock, 0, c.stream>>>(pc, t, pair, c.rho.p, pos_w, c.ep.p)
    if (do_rho && do_emb) {
      if (energy && epi.fused) XS_RHO(true, true, true, true);
      else if (energy) XS_RHO(true, true, true, false);
      else XS_RHO(true, true, false, false);
    } else if (do_rho) {
      XS_RHO(true, false, false, false);
    } else {
      if (energy) XS_RHO(false, true, true, false);
      else XS_RHO(false, true, false, false);
    }
#undef XS_RHO
    XS_HIP(hipGetLastError());
  }
  if (do_force) {
    const Outputs o = make_outputs(c, epi);
    const bool vir = flags & XS_VIRIAL;
    ProfScope prof(c.prof, c.stream, XS_PROF_EAM_FORCE);
#define XS_FORCE(EN, VI, FU) k_eam_force<Real, MULTI, EN, VI, FU><<<g, kBlock, 0, c.stream>>>(pc, t, pair, o)
    if (epi.fused) {
      if (energy && vir) XS_FORCE(true, true, true);
      else if (energy) XS_FORCE(true, false, true);
      else if (vir) XS_FORCE(false, true, true);
      else XS_FORCE(false, false, true);
    } else {
      if (energy && vir) XS_FORCE(true, true, false);
      else if (energy) XS_FORCE(true, false, false);
      else if (vir) XS_FORCE(false, true, false);
      else XS_FORCE(false, false, false);
    }
#undef XS_FORCE
    XS_HIP(hipGetLastError());
  }
}

template <class Real>
void lj_dispatch(Ctx& c, double eps, double sigma, double rcut, unsigned flags, const ForceEpilogue& epi)
{
  const double ecut = lj_energy_cutoff(eps, sigma, rcut);
  LJParams<Real> lp{(Real)(sigma * sigma), (Real)(4.0 * eps), (Real)(24.0 * eps), (Real)ecut};
  const PairCommon pc = make_common(c, rcut);
  const Outputs o = make_outputs(c, epi);
  const unsigned g = pair_grid(c.n);
  const bool energy = flags & XS_ENERGY, vir = flags & XS_VIRIAL;
  ProfScope prof(c.prof, c.stream, XS_PROF_LJ);
#define XS_LJ(EN, VI, FU) k_lj<Real, EN, VI, FU><<<g, kBlock, 0, c.stream>>>(pc, lp, o)
  if (epi.fused) {
    if (energy && vir) XS_LJ(true, true, true);
    else if (energy) XS_LJ(true, false, true);
    else if (vir) XS_LJ(false, true, true);
    else XS_LJ(false, false, true);
  } else {
    if (energy && vir) XS_LJ(true, true, false);
    else if (energy) XS_LJ(true, false, false);
    else if (vir) XS_LJ(false, true, false);
    else XS_LJ(false, false, false);
  }
#undef XS_LJ
  XS_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------
// Newton-symmetric operators: SymRhoOp / SymForceOp with NewtonSym = true
// (eam_force_op_multimat.h:114-343) and lj_compute_force_symetric
// (pair_potential_singlemat_symetric.cpp:101-332).  Each pair is evaluated once,
// from the half list; the atom of the list adds its share in registers, the
// partner's share goes out as FP64 atomic adds (rho_j += rho_{type_i}(r);
// f_j -= fe; ep_j += phi/2; the partner gets the same virial term), and each atom
// adds its own sum atomically at the end.  Operator form only (accumulating into
// the force arrays); the MD loop keeps the full-list kernels.
// ------------------------------------------------------------------------------

// Half of the global list: entry (j, image) of atom i is kept when j > i, or, for
// a periodic image of i itself (j == i), when the image's first nonzero
// component is positive.  Each pair (i, j, image) / (j, i, -image) is then listed
// once, from one of its two atoms.
__device__ __forceinline__ bool half_keep(uint32_t i, uint32_t e)
{
  const uint32_t j = xsmd_nbr_index(e);
  if (j != i) return j > i;
  const int ix = xsmd_nbr_img_x(e), iy = xsmd_nbr_img_y(e), iz = xsmd_nbr_img_z(e);
  return ix > 0 || (ix == 0 && (iy > 0 || (iy == 0 && iz > 0)));
}

__global__ void __launch_bounds__(256) k_half_list(int64_t n, int64_t npad, const uint32_t* __restrict__ nbr,
                                                   const int* __restrict__ nnb, uint32_t* __restrict__ nbr_h,
                                                   int* __restrict__ nnb_h)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int cnt = nnb[i];
  int m = 0;
  for (int k = 0; k < cnt; k++) {
    const uint32_t e = nbr[(int64_t)k * npad + i];
    if (half_keep((uint32_t)i, e)) nbr_h[(int64_t)(m++) * npad + i] = e;
  }
  nnb_h[i] = m;
}

__global__ void __launch_bounds__(256) k_fill0(double* __restrict__ x, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// rho_i += rho_{type_j}(r), rho_j += rho_{type_i}(r) into rho_acc (zeroed before)
template <class Real, bool MULTI>
__global__ void __launch_bounds__(256) k_eam_rho_sym(PairCommon pc, EamTabDev<Real> t, const Real* __restrict__ pair,
                                                     double* __restrict__ rho_acc)
{
  const int64_t i = logical_block() * blockDim.x + threadIdx.x;
  if (i >= pc.n) return;
  const int ti = MULTI ? pc.type[i] : 0;
  const double* L = pc.box[pc.rep ? pc.rep[i] : 0].L;
  const Pos4 pi = load_pos(pc.pos, i);
  const int cnt = (!MULTI || eam_on(t, ti)) ? pc.nnb[i] : 0;
  double rho = 0.0;
  for (int k0 = 0; k0 < cnt; k0 += kBatch) {
    Nb nb[kBatch];
    fetch<kBatch>(pc, i, k0, cnt, pi, L, nb);
    int tj[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      tj[u] = MULTI ? pc.type[nb[u].j] : 0;
      if constexpr (MULTI) nb[u].in = nb[u].in && eam_on(t, tj[u]);
    }
    Real to_i[kBatch], to_j[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      to_i[u] = eam_rho_knot<Real, MULTI>(t, pair, ti, tj[u], nb[u].d2, nb[u].in);
      to_j[u] = to_i[u];
      if constexpr (MULTI) to_j[u] = eam_rho_knot<Real, MULTI>(t, pair, tj[u], ti, nb[u].d2, nb[u].in);
      if (nb[u].in) {
        rho += (double)to_i[u];
        atomicAdd(rho_acc + nb[u].j, (double)to_j[u]);
      }
    }
  }
  atomicAdd(rho_acc + i, rho);
}

// the partner's share of one pair: -fe, phi/2 and the pair's virial term
template <bool ENERGY, bool VIRIAL>
__device__ __forceinline__ void scatter_partner(const Outputs& o, uint32_t j, double fe0, double fe1, double fe2,
                                                double eh, const Nb& nb)
{
  atomicAdd(o.ax + j, -fe0);
  atomicAdd(o.ay + j, -fe1);
  atomicAdd(o.az + j, -fe2);
  if constexpr (ENERGY) atomicAdd(o.ep + j, eh);
  if constexpr (VIRIAL) {
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    add_virial(v, fe0, fe1, fe2, nb.x, nb.y, nb.z);
#pragma unroll
    for (int q = 0; q < 9; q++) atomicAdd(o.vir + q * o.npad + j, v[q]);
  }
}

template <bool ENERGY, bool VIRIAL>
__device__ __forceinline__ void add_own(const Outputs& o, int64_t i, double fx, double fy, double fz, double e,
                                        const double* v)
{
  atomicAdd(o.ax + i, fx);
  atomicAdd(o.ay + i, fy);
  atomicAdd(o.az + i, fz);
  if constexpr (ENERGY) atomicAdd(o.ep + i, e);
  if constexpr (VIRIAL) {
#pragma unroll
    for (int q = 0; q < 9; q++) atomicAdd(o.vir + q * o.npad + i, v[q]);
  }
}

template <class Real, bool MULTI, bool ENERGY, bool VIRIAL>
__global__ void __launch_bounds__(256) k_eam_force_sym(PairCommon pc, EamTabDev<Real> t, const Real* __restrict__ pair,
                                                       Outputs o)
{
  const int64_t i = logical_block() * blockDim.x + threadIdx.x;
  if (i >= pc.n) return;
  const int ti = MULTI ? pc.type[i] : 0;
  const double* L = pc.box[pc.rep ? pc.rep[i] : 0].L;
  const Pos4 pi = load_pos(pc.pos, i);
  const Real fpi = (Real)pi.w;
  double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int cnt = (!MULTI || eam_on(t, ti)) ? pc.nnb[i] : 0;
  for (int k0 = 0; k0 < cnt; k0 += kBatch) {
    Nb nb[kBatch];
    fetch<kBatch>(pc, i, k0, cnt, pi, L, nb);
    int tj[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      tj[u] = MULTI ? pc.type[nb[u].j] : 0;
      if constexpr (MULTI) nb[u].in = nb[u].in && eam_on(t, tj[u]);
    }
    Real fp_r[kBatch], phi[kBatch];
    bool slow[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) phi[u] = Real(0);
    eam_pair_batch<Real, MULTI, ENERGY, kBatch>(t, pair, ti, tj, fpi, nb, fp_r, phi, slow);
    if constexpr (kSTable<Real, MULTI>) {
#pragma unroll
      for (int u = 0; u < kBatch; u++)
        if (slow[u]) eam_pair_knot<Real, false>(t, pair, 0, 0, fpi, (Real)nb[u].w, nb[u].d2, true, fp_r[u], phi[u]);
    }
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      if (nb[u].in) {
        const double fpair = (double)fp_r[u];
        const double fe0 = nb[u].x * fpair, fe1 = nb[u].y * fpair, fe2 = nb[u].z * fpair;
        const double eh = 0.5 * (double)phi[u];
        fx += fe0;
        fy += fe1;
        fz += fe2;
        if constexpr (ENERGY) e += eh;
        if constexpr (VIRIAL) add_virial(v, fe0, fe1, fe2, nb[u].x, nb[u].y, nb[u].z);
        scatter_partner<ENERGY, VIRIAL>(o, nb[u].j, fe0, fe1, fe2, eh, nb[u]);
      }
    }
  }
  add_own<ENERGY, VIRIAL>(o, i, fx, fy, fz, e, v);
}

template <class Real, bool ENERGY, bool VIRIAL>
__global__ void __launch_bounds__(256) k_lj_sym(PairCommon pc, LJParams<Real> lp, Outputs o)
{
  const int64_t i = logical_block() * blockDim.x + threadIdx.x;
  if (i >= pc.n) return;
  const double* L = pc.box[pc.rep ? pc.rep[i] : 0].L;
  const Pos4 pi = load_pos(pc.pos, i);
  double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int cnt = pc.nnb[i];
  for (int k0 = 0; k0 < cnt; k0 += kBatch) {
    Nb nb[kBatch];
    fetch<kBatch>(pc, i, k0, cnt, pi, L, nb);
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      double de_r, eh;
      pm::lj_pair(lp, nb[u].d2, nb[u].in, de_r, eh);
      if (nb[u].in) {
        const double fe0 = de_r * nb[u].x, fe1 = de_r * nb[u].y, fe2 = de_r * nb[u].z;
        fx += fe0;
        fy += fe1;
        fz += fe2;
        if constexpr (ENERGY) e += eh;
        if constexpr (VIRIAL) add_virial(v, fe0, fe1, fe2, nb[u].x, nb[u].y, nb[u].z);
        scatter_partner<ENERGY, VIRIAL>(o, nb[u].j, fe0, fe1, fe2, eh, nb[u]);
      }
    }
  }
  add_own<ENERGY, VIRIAL>(o, i, fx, fy, fz, e, v);
}

template <class Real>
void lj_multi_dispatch(Ctx& c, unsigned flags)
{
  const PairCommon pc = make_common(c, c.ljm_rcmax);
  const Outputs o = make_outputs(c, ForceEpilogue{});
  const pm::LJMultiDev lm = ljm_dev<Real>(c);
  const unsigned g = pair_grid(c.n);
  const bool energy = flags & XS_ENERGY, vir = flags & XS_VIRIAL;
  const LJParams<Real> none{};
  ProfScope prof(c.prof, c.stream, XS_PROF_LJ);
  if (energy && vir) k_lj<Real, true, true, false, true><<<g, kBlock, 0, c.stream>>>(pc, none, o, lm);
  else if (energy) k_lj<Real, true, false, false, true><<<g, kBlock, 0, c.stream>>>(pc, none, o, lm);
  else if (vir) k_lj<Real, false, true, false, true><<<g, kBlock, 0, c.stream>>>(pc, none, o, lm);
  else k_lj<Real, false, false, false, true><<<g, kBlock, 0, c.stream>>>(pc, none, o, lm);
  XS_HIP(hipGetLastError());
}

PairCommon half_common(Ctx& c, double rcut)
{
  ensure_half_list(c);
  PairCommon pc = make_common(c, rcut);
  pc.nbr = c.nbr_half.p;
  pc.nnb = c.nnb_half.p;
  return pc;
}

template <class Real, bool MULTI>
void eam_sym_dispatch(Ctx& c, double rcut, bool do_rho, bool do_emb, bool do_force, unsigned flags)
{
  const EamTabDev<Real> t = make_eam_tab<Real>(c);
  const Real* pair = eam_pair_table<Real>(c);
  const PairCommon pc = half_common(c, rcut);
  const unsigned g = pair_grid(c.n);
  const bool energy = flags & XS_ENERGY;
  double* pos_w = reinterpret_cast<double*>(c.pos.p);
  if (do_rho || do_emb) {
    ProfScope prof(c.prof, c.stream, XS_PROF_EAM_RHO);
    if (do_rho) {
      k_fill0<<<grid_for(c.n), kBlock, 0, c.stream>>>(c.rho.p, c.n);
      k_eam_rho_sym<Real, MULTI><<<g, kBlock, 0, c.stream>>>(pc, t, pair, c.rho.p);
    }
    if (do_emb) {  // F(rho), F'(rho) of the summed densities
      if (energy) k_eam_rho<Real, MULTI, false, true, true, false><<<g, kBlock, 0, c.stream>>>(pc, t, pair, c.rho.p, pos_w, c.ep.p);
      else k_eam_rho<Real, MULTI, false, true, false, false><<<g, kBlock, 0, c.stream>>>(pc, t, pair, c.rho.p, pos_w, c.ep.p);
    }
    XS_HIP(hipGetLastError());
  }
  if (do_force) {
    const Outputs o = make_outputs(c, ForceEpilogue{});
    const bool vir = flags & XS_VIRIAL;
    ProfScope prof(c.prof, c.stream, XS_PROF_EAM_FORCE);
    if (energy && vir) k_eam_force_sym<Real, MULTI, true, true><<<g, kBlock, 0, c.stream>>>(pc, t, pair, o);
    else if (energy) k_eam_force_sym<Real, MULTI, true, false><<<g, kBlock, 0, c.stream>>>(pc, t, pair, o);
    else if (vir) k_eam_force_sym<Real, MULTI, false, true><<<g, kBlock, 0, c.stream>>>(pc, t, pair, o);
    else k_eam_force_sym<Real, MULTI, false, false><<<g, kBlock, 0, c.stream>>>(pc, t, pair, o);
    XS_HIP(hipGetLastError());
  }
}

template <class Real>
void lj_sym_dispatch(Ctx& c, double eps, double sigma, double rcut, unsigned flags)
{
  const LJParams<Real> lp{(Real)(sigma * sigma), (Real)(4.0 * eps), (Real)(24.0 * eps),
                          (Real)lj_energy_cutoff(eps, sigma, rcut)};
  const PairCommon pc = half_common(c, rcut);
  const Outputs o = make_outputs(c, ForceEpilogue{});
  const unsigned g = pair_grid(c.n);
  const bool energy = flags & XS_ENERGY, vir = flags & XS_VIRIAL;
  ProfScope prof(c.prof, c.stream, XS_PROF_LJ);
  if (energy && vir) k_lj_sym<Real, true, true><<<g, kBlock, 0, c.stream>>>(pc, lp, o);
  else if (energy) k_lj_sym<Real, true, false><<<g, kBlock, 0, c.stream>>>(pc, lp, o);
  else if (vir) k_lj_sym<Real, false, true><<<g, kBlock, 0, c.stream>>>(pc, lp, o);
  else k_lj_sym<Real, false, false><<<g, kBlock, 0, c.stream>>>(pc, lp, o);
  XS_HIP(hipGetLastError());
}

}  // namespace

// conversion_z2r = 1 eV.A and conversion_frho = 1 eV of eam_alloy.h:156-157 in
// exaStamp's internal units (kEV, xsmd_internal.hpp).

template <>
EamTabDev<double> make_eam_tab<double>(Ctx& c)
{
  EamTabDev<double> t;
  t.frho = c.eam.frho_d.p;
  t.rhor = c.eam.rhor_d.p;
  t.z2r = c.eam.z2r_d.p;
  t.nr = c.eam.nr;
  t.nrho = c.eam.nrho;
  t.rdr = c.eam.rdr;
  t.rdr_d = c.eam.rdr;
  t.rdrho = c.eam.rdrho;
  t.rhomax = c.eam.rhomax;
  t.conv = kEV;
  std::memcpy(t.t2rhor, c.eam.t2rhor, sizeof(t.t2rhor));
  std::memcpy(t.t2z2r, c.eam.t2z2r, sizeof(t.t2z2r));
  std::memcpy(t.t2frho, c.eam.t2frho, sizeof(t.t2frho));
  t.types = c.eam_types;
  return t;
}

template <>
EamTabDev<float> make_eam_tab<float>(Ctx& c)
{
  EamTabDev<float> t;
  t.frho = c.eam.frho_f.p;
  t.rhor = c.eam.rhor_f.p;
  t.z2r = c.eam.z2r_f.p;
  t.nr = c.eam.nr;
  t.nrho = c.eam.nrho;
  t.rdr = (float)c.eam.rdr;
  t.rdr_d = c.eam.rdr;
  t.rdrho = (float)c.eam.rdrho;
  t.rhomax = (float)c.eam.rhomax;
  t.conv = (float)kEV;
  // s = r^2 tables after the two knot planes of the FP32 pair table (setfl.cpp)
  const EamHost& e = c.eam;
  if (e.s_pieces > 0) {
    const int knot_planes = 8 * (e.nr + 1);
    t.sprows = e.s_pieces;
    t.s_force = (knot_planes + 7) / 8 * 8;
    t.s_phi = t.s_force + 8 * e.s_pieces;
    t.s_rds = e.s_pieces / (e.s_hi - e.s_lo);
    t.s_off = -e.s_lo * t.s_rds;
    t.s_xmax = e.s_pieces;
  } else {  // no s-table: x = -1 for every d^2, so every pair reads the knot rows
    t.sprows = 1;
    t.s_force = t.s_phi = 0;
    t.s_rds = 0.0;
    t.s_off = -1.0;
    t.s_xmax = 1.0;
  }
  std::memcpy(t.t2rhor, c.eam.t2rhor, sizeof(t.t2rhor));
  std::memcpy(t.t2z2r, c.eam.t2z2r, sizeof(t.t2z2r));
  std::memcpy(t.t2frho, c.eam.t2frho, sizeof(t.t2frho));
  t.types = c.eam_types;
  return t;
}

template <>
const double* eam_pair_table<double>(Ctx& c)
{
  return c.eam.pair_d.p;
}
template <>
const float* eam_pair_table<float>(Ctx& c)
{
  return c.eam.pair_f.p;
}

Outputs make_outputs(Ctx& c, const ForceEpilogue& ep)
{
  Outputs o;
  o.ax = c.ax.p;
  o.ay = c.ay.p;
  o.az = c.az.p;
  o.ep = c.ep.p;
  o.vir = c.vir.p;
  o.mass = c.mass_d.p;
  o.type = c.type.p;
  o.vx = c.vx.p;
  o.vy = c.vy.p;
  o.vz = c.vz.p;
  o.kick_dt = ep.kick_dt;
  o.npad = c.npad;
  return o;
}

void ensure_half_list(Ctx& c)
{
  if (c.half_valid) return;
  ensure_global_list(c);
  c.nbr_half.reserve((size_t)c.nmax * c.npad);
  c.nnb_half.reserve(c.npad);
  k_half_list<<<grid_for(c.n), kBlock, 0, c.stream>>>(c.n, c.npad, c.nbr.p, c.nnb.p, c.nbr_half.p, c.nnb_half.p);
  XS_HIP(hipGetLastError());
  c.half_valid = true;
}

void launch_lj(Ctx& c, double eps, double sigma, double rcut, unsigned flags, const ForceEpilogue& epi)
{
  c.drift_done = false;
  if (c.n == 0) return;
  if (flags & XS_SYMMETRIC) {
    if (epi.fused) throw Error(XS_E_ARG, "the Newton-symmetric form is an operator form");
    if (flags & XS_MIXED) lj_sym_dispatch<float>(c, eps, sigma, rcut, flags);
    else lj_sym_dispatch<double>(c, eps, sigma, rcut, flags);
    return;
  }
  if (c.list_kind == ListKind::Block) {
    launch_lj_block(c, eps, sigma, rcut, flags, epi);
    return;
  }
  if (flags & XS_MIXED) lj_dispatch<float>(c, eps, sigma, rcut, flags, epi);
  else lj_dispatch<double>(c, eps, sigma, rcut, flags, epi);
}

void launch_lj_multi(Ctx& c, unsigned flags)
{
  c.drift_done = false;
  if (c.n == 0) return;
  if (c.list_kind == ListKind::Block) {
    launch_lj_multi_block(c, flags);
    return;
  }
  if (flags & XS_MIXED) lj_multi_dispatch<float>(c, flags);
  else lj_multi_dispatch<double>(c, flags);
}

void launch_johnson(Ctx& c, const JohnsonParams& p, double rcut, bool do_emb, bool do_force, unsigned flags)
{
  c.drift_done = false;
  if (do_force && !do_emb && c.rho_state != RhoState::FpJohnson)
    throw Error(XS_E_STATE, "johnson_force_reuse_emb needs F'(rho) from johnson_emb at the current atom order");
  // F' of this potential replaces the position records' w: no eam_alloy state survives
  c.rho_ckpt = false;
  if (do_emb) c.rho_state = RhoState::FpJohnson;
  if (c.n == 0 || !(do_emb || do_force)) return;
  ensure_global_list(c);
  const PairCommon pc = make_common(c, rcut);
  const unsigned g = pair_grid(c.n);
  if (do_emb) {
    ProfScope prof(c.prof, c.stream, XS_PROF_EAM_RHO);
    k_johnson_rho<<<g, kBlock, 0, c.stream>>>(pc, p, c.rho.p, reinterpret_cast<double*>(c.pos.p), c.ep.p);
    XS_HIP(hipGetLastError());
  }
  if (!do_force) return;
  const Outputs o = make_outputs(c, ForceEpilogue{});
  ProfScope prof(c.prof, c.stream, XS_PROF_EAM_FORCE);
  if (flags & XS_VIRIAL) k_johnson_force<true><<<g, kBlock, 0, c.stream>>>(pc, p, o);
  else k_johnson_force<false><<<g, kBlock, 0, c.stream>>>(pc, p, o);
  XS_HIP(hipGetLastError());
}

void launch_eam(Ctx& c, double rcut, bool do_rho, bool do_emb, bool do_force, unsigned flags,
                const ForceEpilogue& epi)
{
  c.drift_done = false;
  if (c.n == 0) return;
  if (do_emb && !do_rho && c.rho_state != RhoState::Rho) throw Error(XS_E_STATE, "eam_rho2emb needs eam_rho first");
  const bool mixed = (flags & XS_MIXED) != 0;
  if (do_force && !(do_rho && do_emb)) {
    const RhoState want = mixed ? RhoState::FpFloat : RhoState::FpDouble;
    if (c.rho_state != want)
      throw Error(XS_E_STATE, "eam_force needs F'(rho) from eam_rho2emb in the same precision");
  }
  const bool multi = eam_multi(c);
  if (flags & XS_SYMMETRIC) {
    if (epi.fused) throw Error(XS_E_ARG, "the Newton-symmetric form is an operator form");
    if (mixed) {
      if (multi) eam_sym_dispatch<float, true>(c, rcut, do_rho, do_emb, do_force, flags);
      else eam_sym_dispatch<float, false>(c, rcut, do_rho, do_emb, do_force, flags);
    } else {
      if (multi) eam_sym_dispatch<double, true>(c, rcut, do_rho, do_emb, do_force, flags);
      else eam_sym_dispatch<double, false>(c, rcut, do_rho, do_emb, do_force, flags);
    }
  } else if (c.list_kind == ListKind::Block) {
    launch_eam_block(c, rcut, do_rho, do_emb, do_force, flags, epi);
  } else if (mixed) {
    if (multi) eam_dispatch<float, true>(c, rcut, do_rho, do_emb, do_force, flags, epi);
    else eam_dispatch<float, false>(c, rcut, do_rho, do_emb, do_force, flags, epi);
  } else {
    if (multi) eam_dispatch<double, true>(c, rcut, do_rho, do_emb, do_force, flags, epi);
    else eam_dispatch<double, false>(c, rcut, do_rho, do_emb, do_force, flags, epi);
  }
  if (do_rho || do_emb) {
    c.rho_state = do_emb ? (mixed ? RhoState::FpFloat : RhoState::FpDouble) : RhoState::Rho;
    if (do_rho) c.rho_ckpt = true;
  }
  if (c.drift_done) {  // positions moved on to the next step: rho and F' are stale
    c.rho_state = RhoState::None;
    c.rho_ckpt = false;
  }
}

}  // namespace xsmd
