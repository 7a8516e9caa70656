// capi.hip -- extern "C" entry points of libxsmd.so (declared in include/xsmd.h).
// Every entry point converts internal errors to an xs_status and records the
// message on the context; nothing aborts.
#include "xsmd_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace xsmd {

Ctx::~Ctx()
{
  if (hscalar) (void)hipHostFree(hscalar);
  if (flag_event) (void)hipEventDestroy(flag_event);
  if (own_stream) (void)hipStreamDestroy(own_stream);
}

namespace {

template <class F>
xs_status guarded(xs_ctx* h, F&& f)
{
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return XS_E_ARG;
  try {
    XS_HIP(hipSetDevice(c->device));
    poll_pass_overflow(*c, false);
    f(*c);
    c->err.clear();
    return XS_OK;
  } catch (const Error& e) {
    c->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c->err = e.what();
    return XS_E_STATE;
  }
}

void require(bool ok, xs_status code, const char* msg)
{
  if (!ok) throw Error(code, msg);
}

void require_system(Ctx& c) { require(c.n > 0 || c.box_h.size() > 0, XS_E_STATE, "no system loaded"); }
void require_list(Ctx& c) { require(c.list_valid, XS_E_STATE, "neighbour list not built (call xs_chunk_neighbors)"); }

template <class T>
void to_device(Ctx& c, DBuf<T>& d, const T* h, int64_t n)
{
  XS_HIP(hipMemcpyAsync(d.p, h, sizeof(T) * n, hipMemcpyHostToDevice, c.stream));
}

template <class T>
std::vector<T> from_device(Ctx& c, const DBuf<T>& d, int64_t n)
{
  std::vector<T> h(n);
  if (n) XS_HIP(hipMemcpyAsync(h.data(), d.p, sizeof(T) * n, hipMemcpyDeviceToHost, c.stream));
  XS_HIP(hipStreamSynchronize(c.stream));
  return h;
}

// gather a device array in cell order back to id order
void to_id_order(Ctx& c, const DBuf<double>& d, double* out, int64_t stride = 0)
{
  if (!out) return;
  const std::vector<int64_t> id = from_device(c, c.id, c.n);
  if (stride == 0) {
    const std::vector<double> v = from_device(c, d, c.n);
    for (int64_t k = 0; k < c.n; k++) out[id[k]] = v[k];
  } else {
    // 9 components stored as vir[q*npad + k]
    const std::vector<double> v = from_device(c, d, 9 * c.npad);
    for (int64_t k = 0; k < c.n; k++)
      for (int q = 0; q < 9; q++) out[9 * id[k] + q] = v[q * c.npad + k];
  }
}

void allocate_atoms(Ctx& c, int64_t n)
{
  c.npad = std::max<int64_t>(64, (n + 63) / 64 * 64);
  const int64_t np = c.npad;
  for (DBuf<double>* b : {&c.vx, &c.vy, &c.vz, &c.ax, &c.ay, &c.az, &c.vx2, &c.vy2, &c.vz2, &c.ax2, &c.ay2,
                          &c.az2, &c.rrx, &c.rry, &c.rrz, &c.ep, &c.rho})
    b->reserve(np);
  c.pos.reserve(np);
  c.pos2.reserve(np);
  c.vir.reserve(9 * np);
  c.type.reserve(np);
  c.type2.reserve(np);
  c.id.reserve(np);
  c.id2.reserve(np);
  c.keys.reserve(np);
  c.keys2.reserve(np);
  c.vals.reserve(np);
  c.vals2.reserve(np);
  if (c.has_rep) {
    c.rep.reserve(np);
    c.rep2.reserve(np);
  }
}

void set_lj(Ctx& c, double eps, double sigma, double rcut)
{
  require(eps >= 0 && sigma > 0 && rcut > 0, XS_E_ARG, "lj: need epsilon >= 0, sigma > 0, rcut > 0");
  c.has_lj = true;
  c.lj_eps = eps;
  c.lj_sigma = sigma;
  c.lj_rcut = rcut;
}

void set_eam(Ctx& c, int nel, int nrho, double drho, int nr, double dr, double rcut, const double* frho,
             const double* rhor, const double* z2r)
{
  require(nel >= 1 && nel <= 7, XS_E_ARG, "eam: 1..7 elements");
  require(nrho >= 5 && nr >= 5 && drho > 0 && dr > 0 && rcut > 0, XS_E_ARG, "eam: bad table grid");
  require(frho && rhor && z2r, XS_E_ARG, "eam: null table");
  require(c.nspecies == 0 || c.nspecies <= nel, XS_E_ARG, "eam: more species than table elements");
  EamHost& e = c.eam;
  e.nel = nel;
  e.nrho = nrho;
  e.nr = nr;
  e.drho = drho;
  e.dr = dr;
  e.rc = rcut;
  e.rdr = 1.0 / dr;
  e.rdrho = 1.0 / drho;
  e.rhomax = (nrho - 1) * drho;
  const int nz = nel * (nel + 1) / 2;
  e.frho_knots.assign(frho, frho + (size_t)nel * nrho);
  e.rhor_knots.assign(rhor, rhor + (size_t)nel * nr);
  e.z2r_knots.assign(z2r, z2r + (size_t)nz * nr);
  eam_upload(c);
  c.has_eam = true;
  c.rho_state = RhoState::None;
}

void compute_forces(Ctx& c, int potential, double rcut, unsigned flags, const ForceEpilogue& epi)
{
  if (potential == XS_POT_LJ) {
    require(c.has_lj, XS_E_STATE, "lj parameters not set");
    launch_lj(c, c.lj_eps, c.lj_sigma, rcut, flags, epi);
  } else {
    require(c.has_eam, XS_E_STATE, "eam tables not set");
    launch_eam(c, rcut, true, true, true, flags, epi);
  }
}

}  // namespace
}  // namespace xsmd

using namespace xsmd;

extern "C" {

int xs_version(void) { return 1; }

xs_status xs_ctx_create(int device, xs_ctx** out)
{
  if (!out) return XS_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return XS_E_HIP;
  Ctx* c = new (std::nothrow) Ctx();
  if (!c) return XS_E_ALLOC;
  c->device = device;
  try {
    XS_HIP(hipSetDevice(device));
    thermo_scratch_init(device);
    XS_HIP(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
    c->stream = c->own_stream;
    XS_HIP(hipEventCreateWithFlags(&c->flag_event, hipEventDisableTiming));
    XS_HIP(hipHostMalloc((void**)&c->hscalar, 8 * sizeof(unsigned long long), hipHostMallocDefault));
    c->dscalar.reserve(8);
    XS_HIP(hipMemsetAsync(c->dscalar.p, 0, 8 * sizeof(unsigned long long), c->stream));  // flags start clear
    XS_HIP(hipStreamSynchronize(c->stream));
    c->mass_d.reserve(8);
    c->box_d.reserve(1);
    const char* kind = std::getenv("XSMD_LIST");
    c->block_disabled = kind && std::strcmp(kind, "global") == 0;
    hipDeviceProp_t prop;
    XS_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cu = prop.multiProcessorCount;
    c->lds_max = prop.sharedMemPerBlockOptin > prop.sharedMemPerBlock ? prop.sharedMemPerBlockOptin
                                                                       : prop.sharedMemPerBlock;
    if (const char* ls = std::getenv("XSMD_LIST_SORT")) c->list_sort_disabled = std::strcmp(ls, "0") == 0;
    if (const char* st = std::getenv("XSMD_EAM_S_TABLE")) c->s_table_disabled = std::strcmp(st, "0") == 0;
    if (const char* sb = std::getenv("XSMD_SORT_BITS")) {
      const int v = std::atoi(sb);
      if (v >= 1 && v <= 8) c->sort_bits = v;
    }
    if (const char* ba = std::getenv("XSMD_BLOCK_ATOMS")) {
      const double v = std::atof(ba);
      if (v >= 1.0 && v <= 4096.0) {
        c->block_atoms = v;
        c->block_atoms_set = true;
      }
    }
  } catch (const Error&) {
    delete c;
    return XS_E_HIP;
  }
  *out = reinterpret_cast<xs_ctx*>(c);
  return XS_OK;
}

xs_status xs_ctx_destroy(xs_ctx* h)
{
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c) return XS_E_ARG;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->own_stream && c->own_stream != c->stream) (void)hipStreamSynchronize(c->own_stream);
  delete c;
  return XS_OK;
}

const char* xs_last_error(const xs_ctx* h)
{
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  return c ? c->err.c_str() : "null context";
}

namespace {
void switch_stream(Ctx& c, hipStream_t next)
{
  if (next == c.stream) return;
  // work queued so far stays ahead of what comes next
  XS_HIP(hipStreamSynchronize(c.stream));
  c.stream = next;
  poll_pass_overflow(c, true);
}
}  // namespace

xs_status xs_ctx_set_stream(xs_ctx* h, void* stream)
{
  return guarded(h, [&](Ctx& c) { switch_stream(c, reinterpret_cast<hipStream_t>(stream)); });
}

xs_status xs_ctx_own_stream(xs_ctx* h)
{
  return guarded(h, [&](Ctx& c) { switch_stream(c, c.own_stream); });
}

xs_status xs_synchronize(xs_ctx* h)
{
  return guarded(h, [&](Ctx& c) {
    XS_HIP(hipStreamSynchronize(c.stream));
    poll_pass_overflow(c, true);
  });
}

xs_status xs_device_memory(xs_ctx* h, size_t* free_bytes, size_t* total_bytes)
{
  return guarded(h, [&](Ctx&) {
    size_t f = 0, t = 0;
    XS_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
  });
}

xs_status xs_species_set(xs_ctx* h, int nspecies, const double* masses)
{
  return guarded(h, [&](Ctx& c) {
    require(nspecies >= 1 && nspecies <= 7 && masses, XS_E_ARG, "species: 1..7 masses required");
    double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = 0; s < nspecies; s++) {
      require(masses[s] > 0, XS_E_ARG, "species: masses must be positive");
      m[s] = masses[s];
    }
    for (int s = nspecies; s < 8; s++) m[s] = 1.0;  // unused slots stay finite
    std::memcpy(c.mass, m, sizeof(m));
    c.nspecies = nspecies;
    XS_HIP(hipMemcpyAsync(c.mass_d.p, c.mass, sizeof(c.mass), hipMemcpyHostToDevice, c.stream));
    XS_HIP(hipStreamSynchronize(c.stream));
  });
}

xs_status xs_system_set(xs_ctx* h, int64_t n, const double* rx, const double* ry, const double* rz,
                        const double* vx, const double* vy, const double* vz, const uint8_t* type,
                        const uint16_t* rep, int nrep, const double* lo, const double* hi,
                        const int periodic[3])
{
  return guarded(h, [&](Ctx& c) {
    require(c.nspecies > 0, XS_E_STATE, "set species masses before the system");
    require(n >= 0 && n < XSMD_NBR_MAX_ATOMS, XS_E_LIMIT, "atom count must be below 2^26");
    require(nrep >= 1 && nrep <= 65535 && lo && hi && periodic, XS_E_ARG, "bad box arguments");
    require(nrep == 1 || rep, XS_E_ARG, "rep array required when nrep > 1");
    require(n == 0 || (rx && ry && rz && vx && vy && vz && type), XS_E_ARG, "null atom array");
    for (int64_t i = 0; i < n; i++) {
      require(type[i] < c.nspecies, XS_E_ARG, "atom type out of range of the species table");
      if (rep) require(rep[i] < nrep, XS_E_ARG, "atom replica out of range");
    }
    c.box_h.assign(nrep, BoxDev{});
    for (int r = 0; r < nrep; r++)
      for (int a = 0; a < 3; a++) {
        const double L = hi[3 * r + a] - lo[3 * r + a];
        require(L > 0 && std::isfinite(L), XS_E_ARG, "box must have positive extent");
        c.box_h[r].lo[a] = lo[3 * r + a];
        c.box_h[r].L[a] = L;
        c.box_h[r].inv_cs[a] = 1.0 / L;
      }
    for (int a = 0; a < 3; a++) c.periodic[a] = periodic[a] ? 1 : 0;
    c.box_d.reserve(nrep);
    XS_HIP(hipMemcpyAsync(c.box_d.p, c.box_h.data(), sizeof(BoxDev) * nrep, hipMemcpyHostToDevice, c.stream));
    c.nrep = nrep;
    c.has_rep = nrep > 1;
    c.n = n;
    allocate_atoms(c, n);
    c.list_valid = false;
    c.half_valid = false;
    c.rho_state = RhoState::None;
    c.rho_ckpt = false;
    if (n > 0) {
      std::vector<Pos4> p4(n);
      for (int64_t i = 0; i < n; i++) p4[i] = Pos4{rx[i], ry[i], rz[i], 0.0};
      to_device(c, c.pos, p4.data(), n);
      to_device(c, c.rrx, rx, n);
      to_device(c, c.rry, ry, n);
      to_device(c, c.rrz, rz, n);
      to_device(c, c.vx, vx, n);
      to_device(c, c.vy, vy, n);
      to_device(c, c.vz, vz, n);
      to_device(c, c.type, type, n);
      std::vector<int64_t> id(n);
      for (int64_t i = 0; i < n; i++) id[i] = i;
      to_device(c, c.id, id.data(), n);
      if (c.has_rep) to_device(c, c.rep, rep, n);
      launch_zero(c);
    }
    XS_HIP(hipStreamSynchronize(c.stream));
  });
}

int64_t xs_system_size(const xs_ctx* h)
{
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  return c ? c->n : -1;
}

xs_status xs_system_get(xs_ctx* h, double* rx, double* ry, double* rz, double* vx, double* vy, double* vz)
{
  return guarded(h, [&](Ctx& c) {
    if (rx || ry || rz) {
      const std::vector<int64_t> id = from_device(c, c.id, c.n);
      const std::vector<Pos4> p = from_device(c, c.pos, c.n);
      for (int64_t k = 0; k < c.n; k++) {
        if (rx) rx[id[k]] = p[k].x;
        if (ry) ry[id[k]] = p[k].y;
        if (rz) rz[id[k]] = p[k].z;
      }
    }
    to_id_order(c, c.vx, vx);
    to_id_order(c, c.vy, vy);
    to_id_order(c, c.vz, vz);
  });
}

xs_status xs_velocity_set(xs_ctx* h, const double* vx, const double* vy, const double* vz)
{
  return guarded(h, [&](Ctx& c) {
    require(vx && vy && vz, XS_E_ARG, "null velocity array");
    const std::vector<int64_t> id = from_device(c, c.id, c.n);
    std::vector<double> t(c.n);
    for (auto [src, dst] : {std::pair<const double*, DBuf<double>*>{vx, &c.vx}, {vy, &c.vy}, {vz, &c.vz}}) {
      for (int64_t k = 0; k < c.n; k++) t[k] = src[id[k]];
      to_device(c, *dst, t.data(), c.n);
      XS_HIP(hipStreamSynchronize(c.stream));
    }
  });
}

xs_status xs_field_get(xs_ctx* h, int field, double* out)
{
  return guarded(h, [&](Ctx& c) {
    require(out != nullptr, XS_E_ARG, "null output");
    switch (field) {
      case XS_FIELD_EP: to_id_order(c, c.ep, out); break;
      case XS_FIELD_VIRIAL: to_id_order(c, c.vir, out, 9); break;
      case XS_FIELD_RHO_DEMB:
        if (c.rho_state == RhoState::FpFloat || c.rho_state == RhoState::FpDouble ||
            c.rho_state == RhoState::FpJohnson) {
          const std::vector<int64_t> id = from_device(c, c.id, c.n);
          const std::vector<Pos4> p = from_device(c, c.pos, c.n);
          for (int64_t k = 0; k < c.n; k++) out[id[k]] = p[k].w;
        } else {
          to_id_order(c, c.rho, out);
        }
        break;
      case XS_FIELD_ACCEL:
        to_id_order(c, c.ax, out);
        to_id_order(c, c.ay, out + c.n);
        to_id_order(c, c.az, out + 2 * c.n);
        break;
      default: throw Error(XS_E_ARG, "unknown field");
    }
  });
}

xs_status xs_chunk_neighbors(xs_ctx* h, double rcut_max, double rcut_inc, int chunk_size)
{
  return guarded(h, [&](Ctx& c) {
    require(chunk_size >= 1 && (chunk_size & (chunk_size - 1)) == 0, XS_E_ARG,
            "chunk_size must be a power of two");
    require(rcut_max > 0 && rcut_inc >= 0, XS_E_ARG, "need rcut_max > 0 and rcut_inc >= 0");
    require(!c.box_h.empty(), XS_E_STATE, "no system loaded (call xs_system_set)");
    build_neighbor_list(c, rcut_max, rcut_inc, true);
  });
}

xs_status xs_particle_displ_over(xs_ctx* h, double threshold, int* result)
{
  return guarded(h, [&](Ctx& c) {
    require(result != nullptr, XS_E_ARG, "null result");
    require_list(c);
    *result = max_displacement2(c) > threshold * threshold ? 1 : 0;
  });
}

xs_status xs_neighbors_get(xs_ctx* h, int64_t* counts, int64_t* total, int64_t* ids, int32_t* imgs)
{
  return guarded(h, [&](Ctx& c) {
    require_list(c);
    require(counts && total, XS_E_ARG, "null counts/total");
    ensure_global_list(c);
    const std::vector<int64_t> id = from_device(c, c.id, c.n);
    const std::vector<int> cnt = from_device(c, c.nnb, c.n);
    int64_t tot = 0;
    for (int64_t k = 0; k < c.n; k++) {
      counts[id[k]] = cnt[k];
      tot += cnt[k];
    }
    *total = tot;
    if (!ids && !imgs) return;
    const std::vector<uint32_t> nb = from_device(c, c.nbr, (int64_t)c.nmax * c.npad);
    std::vector<int64_t> off(c.n + 1, 0);
    for (int64_t i = 0; i < c.n; i++) off[i + 1] = off[i] + counts[i];
    for (int64_t k = 0; k < c.n; k++) {
      int64_t o = off[id[k]];
      for (int q = 0; q < cnt[k]; q++, o++) {
        const uint32_t e = nb[(size_t)q * c.npad + k];
        if (ids) ids[o] = id[xsmd_nbr_index(e)];
        if (imgs) {
          imgs[3 * o] = xsmd_nbr_img_x(e);
          imgs[3 * o + 1] = xsmd_nbr_img_y(e);
          imgs[3 * o + 2] = xsmd_nbr_img_z(e);
        }
      }
    }
  });
}

xs_status xs_order_get(xs_ctx* h, int64_t* perm)
{
  return guarded(h, [&](Ctx& c) {
    require(perm != nullptr, XS_E_ARG, "null perm");
    const std::vector<int64_t> id = from_device(c, c.id, c.n);
    std::copy(id.begin(), id.end(), perm);
  });
}

xs_status xs_pair_count(xs_ctx* h, double rcut, int64_t* entries, int64_t* within_rcut)
{
  return guarded(h, [&](Ctx& c) {
    require_list(c);
    int64_t e, w;
    pair_count(c, rcut, e, w);
    if (entries) *entries = e;
    if (within_rcut) *within_rcut = w;
  });
}

xs_status xs_list_info(xs_ctx* h, int* kind, int64_t* blocks, int* halo_capacity, int* central_max, int* brick)
{
  return guarded(h, [&](Ctx& c) {
    require_list(c);
    const bool blk = c.list_kind == ListKind::Block;
    if (kind) *kind = blk ? 1 : 0;
    if (blocks) *blocks = blk ? c.bgeom.nblocks : 0;
    if (halo_capacity) *halo_capacity = blk ? c.bcap : 0;
    if (central_max) *central_max = blk ? c.bcentral : 0;
    if (brick) {
      brick[0] = blk ? c.bgeom.bx : 0;
      brick[1] = blk ? c.bgeom.by : 0;
      brick[2] = blk ? c.bgeom.bz : 0;
    }
  });
}

xs_status xs_zero_force_energy(xs_ctx* h)
{
  return guarded(h, [&](Ctx& c) { launch_zero(c); });
}

xs_status xs_force_to_accel(xs_ctx* h)
{
  return guarded(h, [&](Ctx& c) { launch_force_to_accel(c); });
}

xs_status xs_lj_compute_force(xs_ctx* h, double epsilon, double sigma, double rcut, unsigned flags)
{
  return guarded(h, [&](Ctx& c) {
    set_lj(c, epsilon, sigma, rcut);
    require_list(c);
    require(rcut <= c.list_rcut + c.list_skin, XS_E_ARG, "rcut exceeds the neighbour-list radius");
    launch_lj(c, epsilon, sigma, rcut, flags, ForceEpilogue{});
    check_or_defer_overflow(c);
  });
}

xs_status xs_lj_parameters(xs_ctx* h, double epsilon, double sigma, double rcut)
{
  return guarded(h, [&](Ctx& c) { set_lj(c, epsilon, sigma, rcut); });
}

xs_status xs_lj_multi_force(xs_ctx* h, int ntypes, const double* epsilon, const double* sigma, const double* rcut,
                            unsigned flags)
{
  return guarded(h, [&](Ctx& c) {
    require(ntypes >= 1 && ntypes <= 7 && ntypes == c.nspecies, XS_E_ARG,
            "lj_multi_force takes one parameter set per pair of the context's species");
    require(epsilon && sigma && rcut, XS_E_ARG, "null parameter array");
    require(!(flags & XS_SYMMETRIC), XS_E_ARG, "lj_multi_force has no Newton-symmetric form");
    const int np = ntypes * (ntypes + 1) / 2;
    double rcmax = 0.0;
    for (int k = 0; k < np; k++) {
      require(std::isfinite(epsilon[k]) && std::isfinite(sigma[k]) && std::isfinite(rcut[k]) && rcut[k] >= 0.0,
              XS_E_ARG, "bad lj_multi_force parameters");
      rcmax = std::max(rcmax, rcut[k]);
    }
    require(rcmax > 0.0, XS_E_ARG, "lj_multi_force needs a positive cutoff");
    require_list(c);
    require(rcmax <= c.list_rcut + c.list_skin, XS_E_ARG, "rcut exceeds the neighbour-list radius");
    std::vector<double> d(4 * (size_t)np);
    std::vector<float> f(4 * (size_t)np);
    for (int k = 0; k < np; k++) {
      const double row[4] = {sigma[k] * sigma[k], 4.0 * epsilon[k], 24.0 * epsilon[k],
                             lj_energy_cutoff(epsilon[k], sigma[k], rcut[k])};
      for (int q = 0; q < 4; q++) {
        d[4 * k + q] = row[q];
        f[4 * k + q] = (float)row[q];
      }
    }
    c.ljm_d.reserve(d.size());
    c.ljm_f.reserve(f.size());
    c.ljm_rcut.reserve(np);
    to_device(c, c.ljm_d, d.data(), (int64_t)d.size());
    to_device(c, c.ljm_f, f.data(), (int64_t)f.size());
    to_device(c, c.ljm_rcut, rcut, np);
    XS_HIP(hipStreamSynchronize(c.stream));  // the host rows go out of scope
    c.has_ljm = true;
    c.ljm_ntypes = ntypes;
    c.ljm_rcmax = rcmax;
    launch_lj_multi(c, flags);
    check_or_defer_overflow(c);
  });
}

xs_status xs_eam_alloy_tables(xs_ctx* h, int nel, int nrho, double drho, int nr, double dr, double rcut,
                              const double* frho, const double* rhor, const double* z2r)
{
  return guarded(h, [&](Ctx& c) { set_eam(c, nel, nrho, drho, nr, dr, rcut, frho, rhor, z2r); });
}

xs_status xs_eam_alloy_load(xs_ctx* h, const char* path, double* masses, int* nel_out)
{
  return guarded(h, [&](Ctx& c) {
    require(path != nullptr, XS_E_ARG, "null path");
    int nel, nrho, nr;
    double drho, dr, rc;
    std::vector<double> f, r, z, m;
    read_setfl(path, nel, nrho, drho, nr, dr, rc, f, r, z, m);
    set_eam(c, nel, nrho, drho, nr, dr, rc, f.data(), r.data(), z.data());
    if (masses) std::copy(m.begin(), m.end(), masses);
    if (nel_out) *nel_out = nel;
  });
}

xs_status xs_eam_alloy_force(xs_ctx* h, double rcut, int eam_rho, int eam_rho2emb, int eam_force, unsigned flags)
{
  return guarded(h, [&](Ctx& c) {
    require(c.has_eam, XS_E_STATE, "eam tables not set (xs_eam_alloy_tables / xs_eam_alloy_load)");
    require(rcut > 0, XS_E_ARG, "rcut must be positive");
    if (eam_rho || eam_force) {
      require_list(c);
      require(rcut <= c.list_rcut + c.list_skin, XS_E_ARG, "rcut exceeds the neighbour-list radius");
    }
    launch_eam(c, rcut, eam_rho != 0, eam_rho2emb != 0, eam_force != 0, flags, ForceEpilogue{});
    check_or_defer_overflow(c);
  });
}

xs_status xs_eam_alloy_types(xs_ctx* h, unsigned enabled)
{
  return guarded(h, [&](Ctx& c) {
    if (c.eam_types != enabled) {
      c.eam_types = enabled;
      c.rho_state = RhoState::None;  // rho and F' of the previous mask no longer hold
      c.rho_ckpt = false;
    }
  });
}

xs_status xs_eam_table_info(xs_ctx* h, int* s_pieces, double* r_lo, double err[3])
{
  return guarded(h, [&](Ctx& c) {
    require(c.has_eam, XS_E_STATE, "eam tables not set");
    if (s_pieces) *s_pieces = c.eam.s_pieces;
    if (r_lo) *r_lo = std::sqrt(c.eam.s_lo);
    if (err)
      for (int k = 0; k < 3; k++) err[k] = c.eam.s_err[k];
  });
}

xs_status xs_johnson_force(xs_ctx* h, const double params[19], double rcut, int johnson_emb, int johnson_force,
                           unsigned flags)
{
  return guarded(h, [&](Ctx& c) {
    require(params != nullptr, XS_E_ARG, "null Johnson parameters");
    require(rcut > 0, XS_E_ARG, "rcut must be positive");
    require(!(flags & (XS_MIXED | XS_SYMMETRIC)), XS_E_ARG,
            "johnson_force runs in FP64 on the full list (no XS_MIXED or XS_SYMMETRIC form)");
    JohnsonParams p;
    static_assert(sizeof(JohnsonParams) == 19 * sizeof(double), "19 packed doubles");
    std::memcpy(&p, params, sizeof p);
    require(p.re > 0 && p.rhoe > 0, XS_E_ARG, "re and rhoe must be positive");
    require_list(c);
    require(rcut <= c.list_rcut + c.list_skin, XS_E_ARG, "rcut exceeds the neighbour-list radius");
    launch_johnson(c, p, rcut, johnson_emb != 0, johnson_force != 0, flags);
  });
}

xs_status xs_eam_adjoint(xs_ctx* h, double rcut, const double* dL_dF, const double* dL_dE, double* grad,
                         unsigned flags)
{
  return guarded(h, [&](Ctx& c) {
    require(dL_dF && dL_dE && grad, XS_E_ARG, "null adjoint argument");
    require(c.has_eam, XS_E_STATE, "eam tables not set");
    require_list(c);
    const bool mixed = (flags & XS_MIXED) != 0;
    require((c.rho_state == RhoState::FpFloat || c.rho_state == RhoState::FpDouble) && c.rho_ckpt, XS_E_STATE,
            "adjoint needs the rho/F' checkpoint of an eam_alloy_force call at the current positions");
    eam_adjoint(c, rcut, dL_dF, dL_dE, 1.0, grad, mixed, false);
  });
}

xs_status xs_eam_adjoint_dev(xs_ctx* h, double rcut, const double* dL_dF_dev, const double* dL_dE_dev,
                             double scale, double* grad_dev, unsigned flags)
{
  return guarded(h, [&](Ctx& c) {
    require(dL_dF_dev && dL_dE_dev && grad_dev, XS_E_ARG, "null adjoint argument");
    require(c.has_eam, XS_E_STATE, "eam tables not set");
    require_list(c);
    require((c.rho_state == RhoState::FpFloat || c.rho_state == RhoState::FpDouble) && c.rho_ckpt, XS_E_STATE,
            "adjoint needs the rho/F' checkpoint of an eam_alloy_force call at the current positions");
    eam_adjoint(c, rcut, dL_dF_dev, dL_dE_dev, scale, grad_dev, (flags & XS_MIXED) != 0, true);
  });
}

xs_status xs_eam_theta_dev(xs_ctx* h, const double* theta_dev)
{
  return guarded(h, [&](Ctx& c) {
    require(theta_dev != nullptr, XS_E_ARG, "null theta");
    require(c.has_eam, XS_E_STATE, "eam tables not set (the knot grid comes from xs_eam_alloy_tables)");
    eam_theta_dev(c, theta_dev);
  });
}

xs_status xs_forces_energies_dev(xs_ctx* h, double* F_dev, double* E_dev, double scale)
{
  return guarded(h, [&](Ctx& c) {
    require(F_dev && E_dev, XS_E_ARG, "null output");
    require(c.n == 0 || c.list_valid, XS_E_STATE, "no neighbour list (the replica ranges come from its cell order)");
    forces_energies_dev(c, F_dev, E_dev, scale);
  });
}

xs_status xs_lj_adjoint(xs_ctx* h, double epsilon, double sigma, double rcut, const double* dL_dF,
                        const double* dL_dE, double grad[2], unsigned flags)
{
  return guarded(h, [&](Ctx& c) {
    require(dL_dF && dL_dE && grad, XS_E_ARG, "null adjoint argument");
    require(epsilon >= 0 && sigma > 0 && rcut > 0, XS_E_ARG, "lj: need epsilon >= 0, sigma > 0, rcut > 0");
    require(flags == 0, XS_E_ARG, "xs_lj_adjoint: flags must be 0 (the sweep is FP64)");
    require_list(c);
    require(rcut <= c.list_rcut + c.list_skin, XS_E_ARG, "rcut exceeds the neighbour-list radius");
    lj_adjoint(c, epsilon, sigma, rcut, dL_dF, dL_dE, grad);
  });
}

xs_status xs_spline_adjoint(int n, double delta, const double* G, double* out)
{
  if (n < 5 || !G || !out || !(delta > 0)) return XS_E_ARG;
  try {
    spline_interpolate_adjoint(n, delta, G, out);
  } catch (...) {
    return XS_E_ARG;
  }
  return XS_OK;
}

xs_status xs_push_f_v_r(xs_ctx* h, double dt, double dt_scale, int xform_mode)
{
  return guarded(h, [&](Ctx& c) {
    if (xform_mode != XS_XFORM_IDENTITY && xform_mode != XS_XFORM_INV)
      throw Error(XS_E_ARG, "push_f_v_r: xform_mode must be XS_XFORM_IDENTITY or XS_XFORM_INV");
    launch_push_f_v_r(c, dt * dt_scale, xform_mode == XS_XFORM_INV);
  });
}

xs_status xs_domain_xform(xs_ctx* h, const double m[9])
{
  return guarded(h, [&](Ctx& c) {
    static const double id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    XformDev x{};
    const double* a = m ? m : id;
    bool ident = true;
    for (int q = 0; q < 9; q++) {
      if (!std::isfinite(a[q])) throw Error(XS_E_ARG, "domain xform: non-finite entry");
      x.m[q] = a[q];
      ident = ident && a[q] == id[q];
    }
    const double det = xform_det(x);
    if (!(det > 0.0)) throw Error(XS_E_ARG, "domain xform: det(m) must be positive");
    // inverse by cofactors, as the oracle computes it
    const double* M = x.m;
    const double inv[9] = {(M[4] * M[8] - M[5] * M[7]) / det, (M[2] * M[7] - M[1] * M[8]) / det,
                           (M[1] * M[5] - M[2] * M[4]) / det, (M[5] * M[6] - M[3] * M[8]) / det,
                           (M[0] * M[8] - M[2] * M[6]) / det, (M[2] * M[3] - M[0] * M[5]) / det,
                           (M[3] * M[7] - M[4] * M[6]) / det, (M[1] * M[6] - M[0] * M[7]) / det,
                           (M[0] * M[4] - M[1] * M[3]) / det};
    for (int q = 0; q < 9; q++) x.mi[q] = inv[q];
    x.on = ident ? 0 : 1;
    c.xf = x;
    c.xf_d.reserve(1);
    XS_HIP(hipMemcpyAsync(c.xf_d.p, &c.xf, sizeof(XformDev), hipMemcpyHostToDevice, c.stream));
    XS_HIP(hipStreamSynchronize(c.stream));
    c.list_valid = false;
    c.rho_state = RhoState::None;
    c.rho_ckpt = false;
  });
}

xs_status xs_push_f_v(xs_ctx* h, double dt, double dt_scale)
{
  return guarded(h, [&](Ctx& c) { launch_push_f_v(c, dt * dt_scale); });
}

xs_status xs_thermo(xs_ctx* h, double out27[27])
{
  return guarded(h, [&](Ctx& c) {
    require(out27 != nullptr, XS_E_ARG, "null output");
    launch_thermo(c, out27);
  });
}

xs_status xs_thermo_noalloc(xs_ctx* h, double* out27_dev)
{
  return guarded(h, [&](Ctx& c) {
    require(out27_dev != nullptr, XS_E_ARG, "null output");
    launch_thermo_noalloc(c, out27_dev);
  });
}

xs_status xs_md_run(xs_ctx* h, int potential, int nsteps, double dt, double rcut, double skin, unsigned flags,
                    int thermo_every, double* thermo_out, xs_run_stats* stats)
{
  return guarded(h, [&](Ctx& c) {
    require(potential == XS_POT_LJ || potential == XS_POT_EAM, XS_E_ARG, "unknown potential");
    require(nsteps >= 0 && dt > 0 && rcut > 0 && skin >= 0, XS_E_ARG, "bad run parameters");
    require(thermo_every <= 0 || thermo_out, XS_E_ARG, "thermo_out required when thermo_every > 0");
    require(potential == XS_POT_LJ ? c.has_lj : c.has_eam, XS_E_STATE, "potential not set");
    require(!(flags & XS_SYMMETRIC), XS_E_ARG, "xs_md_run evaluates full lists (XS_SYMMETRIC is for the operators)");
    const unsigned prec = flags & XS_MIXED;
    const bool cont = (flags & XS_CONTINUE) != 0;
    if (cont) {
      require_list(c);
      require(std::abs(c.list_rcut - rcut) < 1e-12 && std::abs(c.list_skin - skin) < 1e-12, XS_E_ARG,
              "continued run must keep rcut and skin of the current list");
    }
    const auto t0 = std::chrono::steady_clock::now();
    int64_t rebuilds = 0;
    const double thr2 = (0.5 * skin) * (0.5 * skin);

    if (!cont) {
      build_neighbor_list(c, rcut, skin, true);
      rebuilds++;
    }
    auto is_thermo = [&](int step) { return thermo_every > 0 && step % thermo_every == 0; };
    // forces of `step` and its closing kick; with want_drift the force launch may
    // also apply the next step's drift/kick (returns whether it did)
    auto forces = [&](int step, double kick, bool want_drift) {
      const bool th = is_thermo(step);
      const unsigned f = prec | (th ? (XS_ENERGY | XS_VIRIAL) : 0u);
      if (th && potential == XS_POT_LJ) launch_zero(c);
      ForceEpilogue epi;
      epi.fused = true;
      epi.kick_dt = kick;
      epi.drift = want_drift && !th;
      epi.drift_dt = dt;
      epi.drift_thr2 = thr2;
      compute_forces(c, potential, rcut, f, epi);
      if (th) launch_thermo(c, thermo_out + 27 * (size_t)(step / thermo_every));
      return c.drift_done;
    };
    // first_iteration: forces at t = 0, no kick
    bool drifted = false;
    if (!cont) drifted = forces(0, 0.0, nsteps >= 1);
    for (int step = 1; step <= nsteps; step++) {
      bool over;
      if (!drifted) {
        XS_HIP(hipMemsetAsync(c.dscalar.p + 1, 0, sizeof(unsigned long long), c.stream));
        launch_drift_kick(c, dt);
        XS_HIP(hipMemcpyAsync(c.hscalar + 1, c.dscalar.p + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              c.stream));
        XS_HIP(hipStreamSynchronize(c.stream));
        double d2;
        std::memcpy(&d2, &c.hscalar[1], sizeof(double));
        over = d2 > thr2;
      } else {  // the force pass of the previous step moved the atoms and tested them
        // the rebuild flag and the block passes' capacity flag in one copy
        XS_HIP(hipMemcpyAsync(c.hscalar + 5, c.dscalar.p + 5, 2 * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, c.stream));
        XS_HIP(hipStreamSynchronize(c.stream));
        over = c.hscalar[5] != 0;
        if (c.hscalar[6] != 0) check_pass_overflow(c);
      }
      if (c.prof.on) c.prof.flush();
      if (over) {
        build_neighbor_list(c, rcut, skin, true);
        rebuilds++;
      }
      drifted = forces(step, 0.5 * dt, step < nsteps);
    }
    XS_HIP(hipStreamSynchronize(c.stream));
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    check_pass_overflow(c);
    if (c.prof.on) c.prof.flush();
    if (stats) {
      stats->steps = nsteps;
      stats->rebuilds = rebuilds;
      stats->max_neighbors = c.nmax;
      stats->seconds = sec;
    }
  });
}

xs_status xs_gather_ceiling(xs_ctx* h, double rcut, int force_pass, int level, int reps, double* ms)
{
  return guarded(h, [&](Ctx& c) {
    require(ms != nullptr, XS_E_ARG, "null output");
    require_list(c);
    *ms = gather_ceiling(c, rcut, force_pass != 0, level, reps);
  });
}

xs_status xs_stream_copy(xs_ctx* h, size_t bytes, int reps, double* gbs)
{
  return guarded(h, [&](Ctx& c) {
    require(gbs != nullptr, XS_E_ARG, "null output");
    *gbs = stream_copy(c, bytes, reps);
  });
}

xs_status xs_profile_enable(xs_ctx* h, int on)
{
  return guarded(h, [&](Ctx& c) {
    XS_HIP(hipStreamSynchronize(c.stream));
    c.prof.flush();
    c.prof.reset();
    c.prof.on = on != 0;
  });
}

xs_status xs_profile_get(xs_ctx* h, double* ms_total, int64_t* launches)
{
  return guarded(h, [&](Ctx& c) {
    XS_HIP(hipStreamSynchronize(c.stream));
    c.prof.flush();
    for (int k = 0; k < XS_PROF_NCLASSES; k++) {
      if (ms_total) ms_total[k] = c.prof.ms[k];
      if (launches) launches[k] = c.prof.count[k];
    }
  });
}

}  // extern "C"
