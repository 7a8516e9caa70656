// xsmd_internal.hpp -- context, device buffers and kernel-launch declarations
// shared by the libxsmd.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#define XSMD_HD __host__ __device__ __forceinline__
#include "xsmd.h"
#include "xsmd_pair_geometry.h"

namespace xsmd {

// Internal error: carries the C-ABI status; converted at the C boundary.
struct Error : std::runtime_error {
  xs_status code;
  Error(xs_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what)
{
  if (e != hipSuccess) throw Error(XS_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define XS_HIP(call) ::xsmd::hip_check((call), #call)

// Device array owned by a context; grows, never shrinks.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void reserve(size_t count)
  {
    if (count <= n && p) return;
    release();
    if (count == 0) count = 1;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
      p = nullptr;
      n = 0;
      throw Error(XS_E_ALLOC, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    n = count;
  }
  void release()
  {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void swap(DBuf& o)
  {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
};

// Position record of one atom in HBM: x, y, z and a per-atom scalar w (the
// embedding derivative F'(rho) between the two EAM passes).  A neighbour gather is
// one 32-byte sector (two 16-byte loads) instead of several scattered words.
struct alignas(32) Pos4 {
  double x, y, z, w;
};

__device__ __forceinline__ Pos4 load_pos(const Pos4* __restrict__ p, int64_t j)
{
  const double2* q = reinterpret_cast<const double2*>(p + j);
  const double2 a = q[0];
  const double2 b = q[1];
  return {a.x, a.y, b.x, b.y};
}
__device__ __forceinline__ void store_pos(Pos4* __restrict__ p, int64_t j, const Pos4& v)
{
  double2* q = reinterpret_cast<double2*>(p + j);
  q[0] = make_double2(v.x, v.y);
  q[1] = make_double2(v.z, v.w);
}

// Box of one replica, as kernels read it: lo[3], L[3], n/L[3].
struct BoxDev {
  double lo[3];
  double L[3];
  double inv_cs[3];
};

// Domain xform of a deformed domain (exaNBody's Domain::xform): m maps grid-space
// displacements to physical ones, mi = m^-1 maps a physical step back to grid space.
// Positions stay in grid coordinates; forces, velocities and the virial are physical.
// Kernels take a device pointer to it, NULL for the identity (no transform at all), so
// the identity form keeps its argument size and code.
struct XformDev {
  double m[9];
  double mi[9];
  int on;
};

inline double xform_det(const XformDev& x)
{
  const double* m = x.m;
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// Spline tables on the device.  Each row of the reference's SplineCoeffs
// (eam_alloy.h:37-42: c0..c2 derivative quadratic, c3..c6 value cubic) is
// stored as 8 values reordered to [c3 c4 c5 c6 | c0 c1 c2 0] so the value
// cubic and the derivative quadratic are each one aligned 4-vector load.
template <class Real>
struct EamTabDev {
  const Real* frho = nullptr;  // nel tables of (nrho+1) rows
  const Real* rhor = nullptr;  // nel tables of (nr+1) rows
  const Real* z2r = nullptr;   // nel(nel+1)/2 tables of (nr+1) rows
  int nr = 0, nrho = 0;
  Real rdr = 0, rdrho = 0, rhomax = 0, conv = 0;
  double rdr_d = 0;  // 1/dr in FP64 (the knot of r is found in FP64 in both forms)
  int8_t t2rhor[8][8] = {};
  int8_t t2z2r[8][8] = {};
  int8_t t2frho[8] = {};
  // types slot of eam_alloy_force: bit t set = species t takes part (alloy kernels)
  uint32_t types = ~0u;
  // FP32 s = r^2 tables of the single-element pair table (pair_math.hpp): pieces per
  // plane (0 = none), plane offsets in the pair buffer (in Reals), and the piece
  // coordinate x = d^2 s_rds + s_off of d^2, valid in [0, s_xmax = sprows]
  int sprows = 0;
  int s_force = 0, s_phi = 0;
  double s_rds = 0, s_off = 0, s_xmax = 0;
};

struct EamHost {
  int nel = 0, nr = 0, nrho = 0;
  double dr = 0, drho = 0, rdr = 0, rdrho = 0, rc = 0, rhomax = 0;
  std::vector<double> frho_knots, rhor_knots, z2r_knots;
  int8_t t2rhor[8][8] = {};
  int8_t t2z2r[8][8] = {};
  int8_t t2frho[8] = {};
  DBuf<double> frho_d, rhor_d, z2r_d;
  DBuf<float> frho_f, rhor_f, z2r_f;
  // single-element pair table: plane 0 row m = rho [c3 c4 c5 c6], plane 1 row m =
  // z2r [c3 c4 c5 c6]; the derivative quadratics follow from the value cubics, so
  // the two planes serve both EAM passes
  DBuf<double> pair_d;
  DBuf<float> pair_f;
  // FP32 s = r^2 tables after the knot planes of pair_f (pair_math.hpp): pieces
  // (0 = none), their range [s_lo, s_hi], and their largest FP32 errors against the
  // FP64 spline in FP32 ulps of the largest |rho'/r|, |phi| and |phi'/r|
  int s_pieces = 0;
  double s_lo = 0, s_hi = 0;
  double s_err[3] = {0, 0, 0};
};

// What the per-atom EAM scalar currently holds: rho (c.rho) after the density
// pass alone, or F'(rho) (pos[i].w) after the embedding step.
enum class RhoState { None, Rho, FpDouble, FpFloat, FpJohnson };

// Which neighbour list is current: the global ELL list of (index, image)
// entries, or the cell-block list of 16-bit halo slots (k_block.hip).
enum class ListKind { Global, Block };

// Cell-block decomposition: bricks of bx*by*bz cells, nbx*nby*nbz per replica.
struct BlockGeom {
  int nx, ny, nz;
  int bx, by, bz;
  int nbx, nby, nbz;
  int px, py, pz;
  int nh_max;        // (bx+2)(by+2)(bz+2) halo cells
  int64_t ncell_rep;
  int64_t nblocks;   // over all replicas
};

// 1 eV in exaStamp's internal energy unit Da.A^2/ps^2: e / (amu * 1e4 J), exact SI
// charge, CODATA-2022 atomic mass constant (onika's).
constexpr double kEV = 1.602176634e-19 / (1.66053906892e-27 * 1.0e4);
constexpr double kBlockAtoms = 230.0;      // largest expected central atoms per brick
constexpr size_t kLdsPerCu = 160 * 1024;   // LDS of one CU
constexpr int kMaxHaloCells = 216;
constexpr size_t kBlockLdsLimit = 64 * 1024;  // default dynamic-LDS limit of a launch

// Event-pair timing per kernel class (see XS_PROF_* in xsmd.h).
struct Profiler {
  bool on = false;
  std::vector<hipEvent_t> pool;
  struct Pending {
    int cls;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  double ms[XS_PROF_NCLASSES] = {};
  int64_t count[XS_PROF_NCLASSES] = {};

  hipEvent_t take()
  {
    if (pool.empty()) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "hipEventCreate");
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  // Accumulate every recorded pair; the caller has synchronised the stream.
  void flush()
  {
    for (const Pending& p : pending) {
      float t = 0.f;
      hip_check(hipEventElapsedTime(&t, p.a, p.b), "hipEventElapsedTime");
      ms[p.cls] += t;
      count[p.cls]++;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  void reset()
  {
    pending.clear();
    for (int k = 0; k < XS_PROF_NCLASSES; k++) {
      ms[k] = 0;
      count[k] = 0;
    }
  }
  ~Profiler()
  {
    for (const Pending& p : pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

// Records a start event at construction and a stop event at destruction on the
// given stream when profiling is on.
struct ProfScope {
  Profiler& p;
  hipStream_t s;
  int cls;
  hipEvent_t a = nullptr;
  ProfScope(Profiler& prof, hipStream_t stream, int c) : p(prof), s(stream), cls(c)
  {
    if (p.on) {
      a = p.take();
      hip_check(hipEventRecord(a, s), "hipEventRecord");
    }
  }
  ~ProfScope()
  {
    if (a) {
      hipEvent_t b = p.take();
      if (hipEventRecord(b, s) == hipSuccess) p.pending.push_back({cls, a, b});
    }
  }
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;      // the stream every call is ordered on
  hipStream_t own_stream = nullptr;  // the context's own (stream unless xs_ctx_set_stream chose another)
  std::string err;

  int nspecies = 0;
  double mass[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  DBuf<double> inv_mass_d;  // unused slots 0
  DBuf<double> mass_d;

  int64_t n = 0;
  int64_t npad = 0;
  int nrep = 1;
  int periodic[3] = {1, 1, 1};
  std::vector<BoxDev> box_h;
  DBuf<BoxDev> box_d;
  XformDev xf = {{1, 0, 0, 0, 1, 0, 0, 0, 1}, {1, 0, 0, 0, 1, 0, 0, 0, 1}, 0};
  DBuf<XformDev> xf_d;  // device copy of xf; kernels get xf_dev(): NULL for the identity
  const XformDev* xf_dev() const { return xf.on ? xf_d.p : nullptr; }

  // per-atom state in cell-sorted order (and the permutation targets)
  DBuf<Pos4> pos, pos2;
  DBuf<double> vx, vy, vz, ax, ay, az;
  DBuf<double> vx2, vy2, vz2, ax2, ay2, az2;
  DBuf<double> rrx, rry, rrz;  // positions at the last list build
  DBuf<uint8_t> type, type2;
  DBuf<int64_t> id, id2;
  DBuf<uint16_t> rep, rep2;
  bool has_rep = false;
  DBuf<double> ep, vir, rho;
  RhoState rho_state = RhoState::None;
  bool rho_ckpt = false;  // rho (c.rho) and F' (pos.w) belong to the current positions

  // cell binning
  int nc[3] = {1, 1, 1};
  int64_t ncell_rep = 1, ncell_tot = 1;
  DBuf<uint32_t> keys, keys2, vals, vals2;
  DBuf<uint8_t> sort_tmp;

  // neighbour list, ELL column-major: nbr[k*npad + i], k < nnb[i] <= nmax
  DBuf<uint32_t> nbr;
  DBuf<int> nnb;
  int nmax = 0;
  bool list_valid = false;
  double list_rcut = 0, list_skin = 0;
  // cell-block list (same layout, 16-bit halo slots) and its geometry
  ListKind list_kind = ListKind::Global;
  bool global_from_block = false;  // nbr holds the block list converted to global form
  bool block_disabled = false;     // XSMD_LIST=global at context creation
  double block_atoms = kBlockAtoms;  // XSMD_BLOCK_ATOMS at context creation
  bool block_atoms_set = false;
  // half list for the Newton-symmetric operators (global encoding; built on demand
  // from the global form of the current list)
  DBuf<uint32_t> nbr_half;
  DBuf<int> nnb_half;
  bool half_valid = false;
  DBuf<uint16_t> nbr16;
  DBuf<uint8_t> bins8;  // distance keys of the block-list entries (for the sort)
  DBuf<int> cell_off;              // first atom of each cell, ncell_tot + 1
  BlockGeom bgeom{};
  int bcap = 0;                    // LDS slots per block (max halo atoms, rounded)
  int bcentral = 0;                // max central atoms per block
  DBuf<int> brick_rows;            // longest list of each block
  int num_cu = 0;           // CUs of the device (persistent grids)
  size_t lds_max = 65536;   // dynamic LDS one workgroup may use
  bool list_sort_disabled = false;  // XSMD_LIST_SORT=0: block lists stay in stencil order
  int sort_bits = 7;                // distance keys of the list sort: 2^sort_bits bins over [0, rl^2), <= 8
  bool s_table_disabled = false;    // XSMD_EAM_S_TABLE=0: the FP32 passes read the knot rows
  bool drift_done = false;  // the last force launch applied ForceEpilogue::drift

  // small device scalars + pinned host mirror
  DBuf<unsigned long long> dscalar;  // [0] max count, [1] max displacement^2 bits, [2..3] halo maxima,
                                     // [4] atom outside its cell, [5] fused-drift rebuild flag,
                                     // [6] block-pass capacity exceeded
  unsigned long long* hscalar = nullptr;
  hipEvent_t flag_event = nullptr;  // the deferred copy of dscalar[6] to hscalar[6] (check_or_defer_overflow)
  bool flag_pending = false;
  DBuf<double> red_part;  // thermo partial sums
  DBuf<double> red_out;

  bool has_lj = false;
  double lj_eps = 0, lj_sigma = 0, lj_rcut = 0;
  // lj_multi_force: per type pair (packed triangle) FP32 and FP64 rows
  // {sigma^2, 4 eps, 24 eps, E(rcut)} and the pairs' cutoffs
  bool has_ljm = false;
  int ljm_ntypes = 0;
  double ljm_rcmax = 0;
  DBuf<float> ljm_f;
  DBuf<double> ljm_d, ljm_rcut;
  bool has_eam = false;
  EamHost eam;
  // types slot of eam_alloy_force (eam_potential_multimat.cu:90,154-166) as enable
  // bits per species; ~0u = every species (the empty list)
  uint32_t eam_types = ~0u;

  // EAM adjoint scratch: coefficient gradients, lambda_i, dL/dF records, inputs
  DBuf<double> adj_g, adj_lam, adj_in;
  DBuf<double> adj_fc, adj_part;  // per-atom F-table contributions; per-workgroup partial rows
  DBuf<double> adj_grad;          // dL/dtheta on the device (host-output adjoint)
  DBuf<double> rep_part;          // per-replica partial energy sums (forces_energies_dev)
  DBuf<int> adj_fm;               // per-atom F-table row
  DBuf<Pos4> adj_gvec;

  Profiler prof;

  ~Ctx();
};

// The alloy kernels (per-element tables, type maps, the types mask) serve several
// elements, and a single element whose only species the types slot disables.
inline bool eam_multi(const Ctx& c) { return c.eam.nel > 1 || (c.eam_types & 1u) == 0; }

// ---- launches (k_neighbors.hip) ----
// prefer_block: build the cell-block list when its LDS footprint allows
void build_neighbor_list(Ctx& c, double rcut, double skin, bool prefer_block = false);

// ---- cell blocks (k_block.hip) ----
void cell_offsets(Ctx& c);
void block_setup(Ctx& c);
bool block_feasible(const Ctx& c);
size_t block_lds_bytes(const Ctx& c, bool with_type, bool with_w);
unsigned block_grid(const Ctx& c);
void block_build(Ctx& c, double rl2);
void block_to_global(Ctx& c);
void ensure_global_list(Ctx& c);
// the half list (nbr_half / nnb_half) of the current list
void ensure_half_list(Ctx& c);
double max_displacement2(Ctx& c);
// raises XS_E_STATE (and clears the flag) when a block pass found a brick beyond the
// capacities measured at list build (dscalar[6]); synchronises the stream
void check_pass_overflow(Ctx& c);
// The operators' form: on the context's own stream, check_pass_overflow; on a stream
// the caller set (xs_ctx_set_stream), the flag is copied to the host asynchronously
// and checked by poll_pass_overflow at the next call (no synchronisation mid-step).
void check_or_defer_overflow(Ctx& c);
// raises XS_E_STATE when a deferred flag copy has arrived (wait: wait for it) and
// found the flag set
void poll_pass_overflow(Ctx& c, bool wait);
// average ms of the gather-ceiling kernel of the density (force = false) or force pass
double gather_ceiling(Ctx& c, double rcut, bool force, int level, int reps);
void pair_count(Ctx& c, double rcut, int64_t& entries, int64_t& within);

// ---- launches (k_integrate.hip) ----
void launch_zero(Ctx& c);
void launch_force_to_accel(Ctx& c);
// inv_xform: xform_mode INV_XFORM (the step is mapped to grid space by the xform's
// inverse); false: IDENTITY
void launch_push_f_v_r(Ctx& c, double dt, bool inv_xform);
void launch_push_f_v(Ctx& c, double dt);
// fused first half of a step: r += v dt + a dt^2/2 ; v += a dt/2 ; max |r - r_ref|^2
void launch_drift_kick(Ctx& c, double dt);
void launch_thermo(Ctx& c, double out[27]);
// once per device, with that device current: address of the fixed partial-sum array
// of launch_thermo_noalloc on that device
void thermo_scratch_init(int device);
// xs_thermo into device memory (27 doubles at out27_dev) through the fixed array:
// no allocation, no copy, no synchronisation
void launch_thermo_noalloc(Ctx& c, double* out27_dev);
double stream_copy(Ctx& c, size_t bytes, int reps);

// ---- launches (k_pair.hip) ----
struct ForceEpilogue {
  bool fused = false;   // overwrite a with f/m and kick v by a*kick_dt (MD loop)
  double kick_dt = 0;
  // MD loop, fused form only: also apply the next step's push_f_v_r(drift_dt) and
  // push_f_v(drift_dt/2), writing the new positions to pos2 (then made current),
  // and set dscalar[5] to 1 if some atom moved more than sqrt(drift_thr2) since the
  // last build.  A launcher that can do this sets Ctx::drift_done; otherwise the
  // loop runs k_drift_kick itself.
  bool drift = false;
  double drift_dt = 0;
  double drift_thr2 = 0;
};
// energy_cutoff (pair_potential_impl.hxx:487-498): E_LJ(rcut) in FP64, 0 for rcut <= 0
inline double lj_energy_cutoff(double eps, double sigma, double rcut)
{
  if (rcut <= 0.0) return 0.0;
  const double ratio = sigma / rcut, r2 = ratio * ratio, r6 = r2 * r2 * r2, r12 = r6 * r6;
  return 4. * eps * (r12 - r6);
}
void launch_lj(Ctx& c, double eps, double sigma, double rcut, unsigned flags, const ForceEpilogue& ep);
// lj_multi_force (operator form) on the current list; tables set by xs_lj_multi_force
void launch_lj_multi(Ctx& c, unsigned flags);
// EamJohnsonParameters (johnson.h:30-51) in internal units, in the order of
// xs_johnson_force's parameter array
struct JohnsonParams {
  double re, fe, rhoe, alpha, beta, A, B, kappa, lambda, Fn0, Fn1, Fn2, Fn3, F0, F1, F2, F3, Fo, eta;
};
// johnson_force / johnson_emb / johnson_force_reuse_emb: the passes of the
// single-material EAM template, FP64, global list
void launch_johnson(Ctx& c, const JohnsonParams& p, double rcut, bool do_emb, bool do_force, unsigned flags);
void launch_lj_multi_block(Ctx& c, unsigned flags);
// LJ on the cell-block list (k_block.hip)
void launch_lj_block(Ctx& c, double eps, double sigma, double rcut, unsigned flags, const ForceEpilogue& ep);
void launch_eam(Ctx& c, double rcut, bool do_rho, bool do_rho2emb, bool do_force, unsigned flags,
                const ForceEpilogue& ep);
void launch_eam_block(Ctx& c, double rcut, bool do_rho, bool do_emb, bool do_force, unsigned flags,
                      const ForceEpilogue& ep);
template <class Real>
EamTabDev<Real> make_eam_tab(Ctx& c);
template <class Real>
const Real* eam_pair_table(Ctx& c);

// ---- EAM adjoint (k_adjoint.hip) ----
// dL/dtheta of eam_alloy_force for the incoming gradients dLdF_id (3n) and dLdE (nrep)
// times `scale`.  They and grad are device pointers when `device` is set (nothing is
// copied to or from the host and the call does not synchronise), host pointers
// otherwise.
void eam_adjoint(Ctx& c, double rcut, const double* dLdF_id, const double* dLdE, double scale, double* grad,
                 bool mixed, bool device);
void spline_interpolate_adjoint(int n, double delta, const double* G, double* out);
// dL/d(epsilon, sigma) of lj_compute_force's forces and shifted energies (k_adjoint.hip)
void lj_adjoint(Ctx& c, double eps, double sigma, double rcut, const double* dLdF_id, const double* dLdE,
                double grad[2]);

// ---- setfl + splines (setfl.cpp) ----
void spline_interpolate(int n, double delta, const double* f, double* coeffs);
void read_setfl(const char* path, int& nel, int& nrho, double& drho, int& nr, double& dr,
                double& rcut, std::vector<double>& frho, std::vector<double>& rhor,
                std::vector<double>& z2r, std::vector<double>& masses);
void eam_upload(Ctx& c);
// New knots theta (setfl order, device memory) for the loaded tables' grid: the spline
// rows of both precisions built on the device, bit for bit the host build; the FP32
// s = r^2 tables are dropped (the mixed force pass reads the knot rows).
void eam_theta_dev(Ctx& c, const double* theta);
// forces m a_i in id order (3n, x y z per atom) and per-replica energies (nrep), both
// times `scale`, into device memory, on the context's stream
void forces_energies_dev(Ctx& c, double* F, double* E, double scale);

// ---- helpers ----
constexpr int kBlock = 256;
inline unsigned grid_for(int64_t n, int block = kBlock)
{
  return (unsigned)((n + block - 1) / block);
}

}  // namespace xsmd
