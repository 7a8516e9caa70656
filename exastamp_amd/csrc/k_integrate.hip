// k_integrate.hip -- per-atom operators (zero, force -> acceleration, velocity-
// Verlet pushes) and the deterministic FP64 thermodynamic reduction.
#include "xsmd_internal.hpp"

#include <algorithm>
#include <mutex>

namespace xsmd {
namespace {

__global__ void k_zero(int64_t n, int64_t npad, double* ax, double* ay, double* az, double* ep,
                       double* vir)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ax[i] = 0.0;
  ay[i] = 0.0;
  az[i] = 0.0;
  ep[i] = 0.0;
#pragma unroll
  for (int q = 0; q < 9; q++) vir[q * npad + i] = 0.0;
}

__global__ void k_force_to_accel(int64_t n, double* ax, double* ay, double* az,
                                 const uint8_t* __restrict__ type, const double* __restrict__ mass)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m = mass[type[i]];
  ax[i] /= m;
  ay[i] /= m;
  az[i] /= m;
}

// x += v*dt + a*dt2 with dt2 = dt*dt/2, evaluated as x + fma(v, dt, a*dt2); with
// xf (xform_mode INV_XFORM under a non-identity xform) the physical step is mapped
// to grid space first, x += mi * (v*dt + a*dt2)
// (push_vec3_2nd_order_xform_minimization.cpp:34).
__global__ void k_push_f_v_r(int64_t n, double dt, double dt2, Pos4* __restrict__ pos,
                             const double* __restrict__ vx, const double* __restrict__ vy,
                             const double* __restrict__ vz, const double* __restrict__ ax,
                             const double* __restrict__ ay, const double* __restrict__ az,
                             const XformDev* __restrict__ xf)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Pos4 p = load_pos(pos, i);
  double sx = fma(vx[i], dt, ax[i] * dt2), sy = fma(vy[i], dt, ay[i] * dt2), sz = fma(vz[i], dt, az[i] * dt2);
  if (xf) xsmd_xform_apply(xf->mi, &sx, &sy, &sz);
  p.x += sx;
  p.y += sy;
  p.z += sz;
  store_pos(pos, i, p);
}

__global__ void k_push_f_v(int64_t n, double dt, double* vx, double* vy, double* vz,
                           const double* __restrict__ ax, const double* __restrict__ ay,
                           const double* __restrict__ az)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vx[i] = fma(ax[i], dt, vx[i]);
  vy[i] = fma(ay[i], dt, vy[i]);
  vz[i] = fma(az[i], dt, vz[i]);
}

// push_f_v_r(dt) then push_f_v(dt/2), plus the largest squared displacement
// since the last list build (as bits of a non-negative double, which order
// like the doubles themselves): reduced per block, one atomic per block.  Under a
// non-identity xform the step is mapped by mi (INV_XFORM) and the displacement is
// measured in physical space.
__global__ void __launch_bounds__(256) k_drift_kick(
    int64_t n, double dt, double dt2, double hdt, Pos4* __restrict__ pos, double* vx, double* vy,
    double* vz, const double* __restrict__ ax, const double* __restrict__ ay,
    const double* __restrict__ az, const double* __restrict__ rrx, const double* __restrict__ rry,
    const double* __restrict__ rrz, unsigned long long* __restrict__ maxd2, const XformDev* __restrict__ xf)
{
  double d2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double a0 = ax[i], a1 = ay[i], a2 = az[i];
    const double v0 = vx[i], v1 = vy[i], v2 = vz[i];
    Pos4 p = load_pos(pos, i);
    double sx = fma(v0, dt, a0 * dt2), sy = fma(v1, dt, a1 * dt2), sz = fma(v2, dt, a2 * dt2);
    if (xf) xsmd_xform_apply(xf->mi, &sx, &sy, &sz);
    p.x += sx;
    p.y += sy;
    p.z += sz;
    store_pos(pos, i, p);
    vx[i] = fma(a0, hdt, v0);
    vy[i] = fma(a1, hdt, v1);
    vz[i] = fma(a2, hdt, v2);
    double ux = p.x - rrx[i], uy = p.y - rry[i], uz = p.z - rrz[i];
    if (xf) xsmd_xform_apply(xf->m, &ux, &uy, &uz);
    d2 = fmax(d2, xsmd_d2(ux, uy, uz));
  }
  for (int off = 32; off > 0; off >>= 1) d2 = fmax(d2, __shfl_xor(d2, off));
  __shared__ double sh[kBlock / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = d2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sh[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, sh[w]);
    if (m > 0.0) atomicMax(maxd2, (unsigned long long)__double_as_longlong(m));
  }
}

// Plain streaming copy, 16 bytes per lane per iteration.
__global__ void __launch_bounds__(256) k_stream_copy(int64_t n16, const double2* __restrict__ src,
                                                     double2* __restrict__ dst)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

constexpr int kThermoBlocks = 1024;
constexpr int kThermoQ = 26;

template <class T>
__device__ __forceinline__ T wave_sum(T v)
{
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Per-block partial sums of the 26 quantities of simulation_thermodynamic_state
// (virial 9, v(x)v m 9, m v 3, m v^2 3, ep, m) over a fixed grid, so the
// summation order -- and the result -- does not depend on timing.
__global__ void __launch_bounds__(256) k_thermo_partial(
    int64_t n, int64_t npad, const double* __restrict__ vx, const double* __restrict__ vy,
    const double* __restrict__ vz, const double* __restrict__ ep, const double* __restrict__ vir,
    const uint8_t* __restrict__ type, const double* __restrict__ mass, double* __restrict__ part)
{
  double acc[kThermoQ];
#pragma unroll
  for (int q = 0; q < kThermoQ; q++) acc[q] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double m = mass[type[i]];
    const double v[3] = {vx[i], vy[i], vz[i]};
#pragma unroll
    for (int q = 0; q < 9; q++) acc[q] += vir[q * npad + i];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) acc[9 + 3 * a + b] += (v[a] * v[b]) * m;
#pragma unroll
    for (int a = 0; a < 3; a++) acc[18 + a] += v[a] * m;
#pragma unroll
    for (int a = 0; a < 3; a++) acc[21 + a] += v[a] * v[a] * m;
    acc[24] += ep[i];
    acc[25] += m;
  }
  __shared__ double sh[4][kThermoQ];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < kThermoQ; q++) {
    const double s = wave_sum(acc[q]);
    if (lane == 0) sh[w][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < kThermoQ) {
    const int q = threadIdx.x;
    part[(int64_t)blockIdx.x * kThermoQ + q] = ((sh[0][q] + sh[1][q]) + sh[2][q]) + sh[3][q];
  }
}

__global__ void k_thermo_final(int nblocks, const double* __restrict__ part, double* __restrict__ out)
{
  const int q = threadIdx.x;
  if (q >= kThermoQ) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; b++) s += part[(int64_t)b * kThermoQ + q];
  out[q] = s;
}

// The same sums in the same order as k_thermo_final, finished on the device into the
// 27-double layout of xs_thermo (kinetic tensor and kinetic energy halved, particle
// count last), so the result never leaves device memory.
__global__ void k_thermo_final27(int nblocks, const double* __restrict__ part, int64_t n,
                                 double* __restrict__ out27)
{
  const int q = threadIdx.x;
  if (q >= 27) return;
  if (q == 26) {
    out27[26] = (double)n;
    return;
  }
  double s = 0.0;
  for (int b = 0; b < nblocks; b++) s += part[(int64_t)b * kThermoQ + q];
  const bool half = (q >= 9 && q < 18) || (q >= 21 && q < 24);
  out27[q] = half ? 0.5 * s : s;
}

// Partial sums of xs_thermo_noalloc: kThermoBlocks rows of kThermoQ doubles
// (208 KiB) in a fixed device array rather than context buffers, so that the call
// allocates nothing.  Each device holds its own copy of the array for the life of
// the process; host code must use the address thermo_scratch_init obtained for that
// device, never the array's name (which on the host is only a stand-in symbol).
constexpr size_t kScratchDoubles = (size_t)kThermoBlocks * kThermoQ;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

void launch_zero(Ctx& c)
{
  if (c.n == 0) return;
  k_zero<<<grid_for(c.n), kBlock, 0, c.stream>>>(c.n, c.npad, c.ax.p, c.ay.p, c.az.p, c.ep.p, c.vir.p);
  XS_HIP(hipGetLastError());
}

void launch_force_to_accel(Ctx& c)
{
  if (c.n == 0) return;
  k_force_to_accel<<<grid_for(c.n), kBlock, 0, c.stream>>>(c.n, c.ax.p, c.ay.p, c.az.p, c.type.p,
                                                           c.mass_d.p);
  XS_HIP(hipGetLastError());
}

void launch_push_f_v_r(Ctx& c, double dt, bool inv_xform)
{
  if (c.n == 0) return;
  c.rho_ckpt = false;
  const double dt2 = dt * dt * 0.5;
  k_push_f_v_r<<<grid_for(c.n), kBlock, 0, c.stream>>>(c.n, dt, dt2, c.pos.p, c.vx.p, c.vy.p, c.vz.p,
                                                       c.ax.p, c.ay.p, c.az.p,
                                                       inv_xform ? c.xf_dev() : nullptr);
  XS_HIP(hipGetLastError());
}

void launch_push_f_v(Ctx& c, double dt)
{
  if (c.n == 0) return;
  k_push_f_v<<<grid_for(c.n), kBlock, 0, c.stream>>>(c.n, dt, c.vx.p, c.vy.p, c.vz.p, c.ax.p, c.ay.p,
                                                     c.az.p);
  XS_HIP(hipGetLastError());
}

void launch_drift_kick(Ctx& c, double dt)
{
  if (c.n == 0) return;
  c.rho_ckpt = false;
  const double dt2 = dt * dt * 0.5;
  ProfScope prof(c.prof, c.stream, XS_PROF_DRIFT);
  const unsigned g = std::min<unsigned>(grid_for(c.n), 4096);
  k_drift_kick<<<g, kBlock, 0, c.stream>>>(c.n, dt, dt2, dt * 0.5, c.pos.p, c.vx.p, c.vy.p, c.vz.p, c.ax.p,
                                           c.ay.p, c.az.p, c.rrx.p, c.rry.p, c.rrz.p, c.dscalar.p + 1, c.xf_dev());
  XS_HIP(hipGetLastError());
}

void launch_thermo(Ctx& c, double out[27])
{
  c.red_part.reserve((size_t)kThermoBlocks * kThermoQ);
  c.red_out.reserve(kThermoQ);
  double s[kThermoQ] = {0};
  if (c.n > 0) {
    ProfScope prof(c.prof, c.stream, XS_PROF_THERMO);
    const int nb = (int)std::min<int64_t>(kThermoBlocks, grid_for(c.n));
    k_thermo_partial<<<nb, kBlock, 0, c.stream>>>(c.n, c.npad, c.vx.p, c.vy.p, c.vz.p, c.ep.p, c.vir.p,
                                                  c.type.p, c.mass_d.p, c.red_part.p);
    XS_HIP(hipGetLastError());
    k_thermo_final<<<1, 64, 0, c.stream>>>(nb, c.red_part.p, c.red_out.p);
    XS_HIP(hipGetLastError());
    XS_HIP(hipMemcpyAsync(s, c.red_out.p, sizeof(s), hipMemcpyDeviceToHost, c.stream));
    XS_HIP(hipStreamSynchronize(c.stream));
  }
  for (int q = 0; q < 9; q++) out[q] = s[q];
  for (int q = 9; q < 18; q++) out[q] = 0.5 * s[q];
  for (int q = 18; q < 21; q++) out[q] = s[q];
  for (int q = 21; q < 24; q++) out[q] = 0.5 * s[q];
  out[24] = s[24];
  out[25] = s[25];
  out[26] = (double)c.n;
}

void thermo_scratch_init(int device)
{
  if (device < 0 || device >= kMaxDevices) throw Error(XS_E_LIMIT, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  XS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = static_cast<double*>(p);
}

void launch_thermo_noalloc(Ctx& c, double* out27_dev)
{
  double* part = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    part = (c.device >= 0 && c.device < kMaxDevices) ? g_scratch_ptr[c.device] : nullptr;
  }
  if (!part) throw Error(XS_E_STATE, "thermo scratch of this device not initialised");
  const int nb = c.n > 0 ? (int)std::min<int64_t>(kThermoBlocks, grid_for(c.n)) : 0;
  if ((size_t)nb * kThermoQ > kScratchDoubles)
    throw Error(XS_E_LIMIT, "thermo partial sums exceed the fixed scratch (" + std::to_string(kScratchDoubles) +
                                " doubles)");
  if (nb > 0) {
    k_thermo_partial<<<nb, kBlock, 0, c.stream>>>(c.n, c.npad, c.vx.p, c.vy.p, c.vz.p, c.ep.p, c.vir.p, c.type.p,
                                                  c.mass_d.p, part);
    XS_HIP(hipGetLastError());
  }
  k_thermo_final27<<<1, 64, 0, c.stream>>>(nb, part, c.n, out27_dev);
  XS_HIP(hipGetLastError());
}

}  // namespace xsmd

namespace xsmd {
namespace {

// F (id order) = m_type a_i, the forces behind the accelerations force_to_accel left
__global__ void k_forces_id(int64_t n, const int64_t* __restrict__ id, const uint8_t* __restrict__ type,
                            const double* __restrict__ mass, const double* __restrict__ ax,
                            const double* __restrict__ ay, const double* __restrict__ az, double scale,
                            double* __restrict__ F)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m = mass[type[i]];
  double* f = F + 3 * id[i];
  f[0] = ax[i] * m * scale;
  f[1] = ay[i] * m * scale;
  f[2] = az[i] * m * scale;
}

// Per-replica energy: the atoms of replica r are the cell-sorted range
// [cell_off[r ncell_rep], cell_off[(r + 1) ncell_rep]); kRepChunks workgroups per
// replica each sum a fixed slice (thread-strided, then a fixed tree), and one thread
// per replica adds the slices in order, so the sums do not depend on scheduling.
constexpr int kRepChunks = 32;
constexpr int kRepThreads = 256;

__global__ void __launch_bounds__(kRepThreads) k_replica_partial(const int* __restrict__ cell_off,
                                                                 int64_t ncell_rep, const double* __restrict__ ep,
                                                                 double* __restrict__ part)
{
  __shared__ double red[kRepThreads / 64];
  const int r = blockIdx.x / kRepChunks, b = blockIdx.x % kRepChunks;
  const int64_t lo = cell_off[(int64_t)r * ncell_rep], hi = cell_off[(int64_t)(r + 1) * ncell_rep];
  const int64_t len = hi - lo;
  const int64_t s0 = lo + len * b / kRepChunks, s1 = lo + len * (b + 1) / kRepChunks;
  double v = 0.0;
  for (int64_t i = s0 + threadIdx.x; i < s1; i += kRepThreads) v += ep[i];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kRepThreads / 64; w++) t += red[w];
    part[blockIdx.x] = t;
  }
}

__global__ void k_replica_final(int nrep, const double* __restrict__ part, double scale, double* __restrict__ E)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  double t = 0.0;
  for (int b = 0; b < kRepChunks; b++) t += part[r * kRepChunks + b];
  E[r] = t * scale;
}

}  // namespace

void forces_energies_dev(Ctx& c, double* F, double* E, double scale)
{
  if (c.n == 0) {
    XS_HIP(hipMemsetAsync(E, 0, sizeof(double) * c.nrep, c.stream));
    return;
  }
  k_forces_id<<<grid_for(c.n), kBlock, 0, c.stream>>>(c.n, c.id.p, c.type.p, c.mass_d.p, c.ax.p, c.ay.p, c.az.p,
                                                     scale, F);
  XS_HIP(hipGetLastError());
  c.rep_part.reserve((size_t)c.nrep * kRepChunks);
  k_replica_partial<<<(unsigned)(c.nrep * kRepChunks), kRepThreads, 0, c.stream>>>(c.cell_off.p, c.ncell_rep,
                                                                                   c.ep.p, c.rep_part.p);
  XS_HIP(hipGetLastError());
  k_replica_final<<<grid_for(c.nrep), kBlock, 0, c.stream>>>(c.nrep, c.rep_part.p, scale, E);
  XS_HIP(hipGetLastError());
}

double stream_copy(Ctx& c, size_t bytes, int reps)
{
  const int64_t n16 = (int64_t)(bytes / 16);
  if (n16 <= 0 || reps <= 0) throw Error(XS_E_ARG, "stream copy needs bytes >= 16 and reps >= 1");
  DBuf<double2> a, b;
  a.reserve(n16);
  b.reserve(n16);
  XS_HIP(hipMemsetAsync(a.p, 0, n16 * 16, c.stream));
  hipEvent_t e0, e1;
  XS_HIP(hipEventCreate(&e0));
  XS_HIP(hipEventCreate(&e1));
  const unsigned g = std::min<unsigned>(grid_for(n16), 256u * 64u);
  float best = 1e30f;
  for (int r = 0; r < reps; r++) {
    XS_HIP(hipEventRecord(e0, c.stream));
    k_stream_copy<<<g, kBlock, 0, c.stream>>>(n16, a.p, b.p);
    XS_HIP(hipEventRecord(e1, c.stream));
    XS_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    XS_HIP(hipEventElapsedTime(&ms, e0, e1));
    best = std::min(best, ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 2.0 * (double)n16 * 16.0 / (best * 1e-3) / 1e9;
}

}  // namespace xsmd
