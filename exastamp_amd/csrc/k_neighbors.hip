// k_neighbors.hip -- cell binning, atom re-sort and full neighbour-list build.
//
// A rebuild is: wrap atoms into their periodic box, compute each atom's cell,
// stable radix sort by cell (so the in-cell order is deterministic and equal to
// a stable CPU sort), gather every per-atom array into the new order, find the
// cell ranges, then one thread per atom scans the 27 surrounding cells and
// writes the (atom, image) entries within rcut+skin into an ELL list stored
// column-major (entry k of atom i at nbr[k*npad + i]) so the force kernels read
// it with coalesced loads.
#include "xsmd_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

namespace xsmd {
namespace {

__global__ void k_wrap_cellkey(int64_t n, double* __restrict__ rx, double* __restrict__ ry,
                               double* __restrict__ rz, const uint16_t* __restrict__ rep,
                               const BoxDev* __restrict__ box, int px, int py, int pz, int nx,
                               int ny, int nz, int64_t ncell_rep, uint32_t* __restrict__ keys,
                               uint32_t* __restrict__ vals)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = rep ? rep[i] : 0;
  const BoxDev& b = box[r];
  double x = rx[i], y = ry[i], z = rz[i];
  if (px) { x = xsmd_wrap(x, b.lo[0], b.L[0]); rx[i] = x; }
  if (py) { y = xsmd_wrap(y, b.lo[1], b.L[1]); ry[i] = y; }
  if (pz) { z = xsmd_wrap(z, b.lo[2], b.L[2]); rz[i] = z; }
  const int cx = xsmd_cell_coord(x, b.lo[0], b.inv_cs[0], nx);
  const int cy = xsmd_cell_coord(y, b.lo[1], b.inv_cs[1], ny);
  const int cz = xsmd_cell_coord(z, b.lo[2], b.inv_cs[2], nz);
  keys[i] = (uint32_t)((int64_t)r * ncell_rep + ((int64_t)cz * ny + cy) * nx + cx);
  vals[i] = (uint32_t)i;
}

struct PermuteArgs {
  const double* src[12];
  double* dst[12];
  const uint8_t* type_src;
  uint8_t* type_dst;
  const int64_t* id_src;
  int64_t* id_dst;
  const uint16_t* rep_src;
  uint16_t* rep_dst;
};

__global__ void k_permute(int64_t n, const uint32_t* __restrict__ perm, PermuteArgs a)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = perm[i];
#pragma unroll
  for (int d = 0; d < 12; d++) a.dst[d][i] = a.src[d][p];
  a.type_dst[i] = a.type_src[p];
  a.id_dst[i] = a.id_src[p];
  if (a.rep_src) a.rep_dst[i] = a.rep_src[p];
}

__global__ void k_cell_bounds(int64_t n, const uint32_t* __restrict__ keys, int* __restrict__ cstart,
                              int* __restrict__ cend)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  if (i == 0 || keys[i - 1] != k) cstart[k] = (int)i;
  if (i == n - 1 || keys[i + 1] != k) cend[k] = (int)(i + 1);
}

// One thread per atom; stencil order dz, dy, dx (outer to inner), in-cell order
// ascending -- the same visit order as the oracle, so lists match entry by entry.
__global__ void __launch_bounds__(256) k_build(
    int64_t n, const double* __restrict__ rx, const double* __restrict__ ry,
    const double* __restrict__ rz, const uint32_t* __restrict__ keys,
    const BoxDev* __restrict__ box, int px, int py, int pz, int nx, int ny, int nz,
    int64_t ncell_rep, const int* __restrict__ cstart, const int* __restrict__ cend, double rl2,
    uint32_t* __restrict__ nbr, int* __restrict__ nnb, int nmax, int64_t npad,
    unsigned long long* __restrict__ maxcnt)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t key = keys[i];
  const int64_t r = key / ncell_rep;
  const int64_t cl = key - r * ncell_rep;
  const int cx = (int)(cl % nx);
  const int cy = (int)((cl / nx) % ny);
  const int cz = (int)(cl / ((int64_t)nx * ny));
  const BoxDev& b = box[r];
  const double xi = rx[i], yi = ry[i], zi = rz[i];
  int cnt = 0;
  for (int dz = -1; dz <= 1; dz++) {
    int z = cz + dz, kz = 0;
    if (z < 0) { if (!pz) continue; z += nz; kz = -1; }
    else if (z >= nz) { if (!pz) continue; z -= nz; kz = 1; }
    for (int dy = -1; dy <= 1; dy++) {
      int y = cy + dy, ky = 0;
      if (y < 0) { if (!py) continue; y += ny; ky = -1; }
      else if (y >= ny) { if (!py) continue; y -= ny; ky = 1; }
      for (int dx = -1; dx <= 1; dx++) {
        int x = cx + dx, kx = 0;
        if (x < 0) { if (!px) continue; x += nx; kx = -1; }
        else if (x >= nx) { if (!px) continue; x -= nx; kx = 1; }
        const int64_t cell = r * ncell_rep + ((int64_t)z * ny + y) * nx + x;
        const int j0 = cstart[cell], j1 = cend[cell];
        const bool self_img = (kx == 0 && ky == 0 && kz == 0);
        for (int j = j0; j < j1; j++) {
          if (self_img && j == (int)i) continue;
          const double ddx = xsmd_image(rx[j], kx, b.L[0]) - xi;
          const double ddy = xsmd_image(ry[j], ky, b.L[1]) - yi;
          const double ddz = xsmd_image(rz[j], kz, b.L[2]) - zi;
          if (xsmd_d2(ddx, ddy, ddz) < rl2) {
            if (cnt < nmax) nbr[(int64_t)cnt * npad + i] = xsmd_nbr_encode((uint32_t)j, kx, ky, kz);
            cnt++;
          }
        }
      }
    }
  }
  nnb[i] = cnt;
  if (cnt > nmax) atomicMax(maxcnt, (unsigned long long)cnt);
}

}  // namespace

void build_neighbor_list(Ctx& c, double rcut, double skin)
{
  const int64_t n = c.n;
  if (n == 0) {
    c.list_valid = true;
    return;
  }
  const double rl = rcut + skin;
  if (!(rl > 0)) throw Error(XS_E_ARG, "rcut + skin must be positive");

  // cell grid: common dims for all replicas (from the smallest box), each
  // replica with its own n/L so every cell is at least rl wide
  for (int a = 0; a < 3; a++) {
    int nmin = 1 << 30;
    for (const BoxDev& b : c.box_h) nmin = std::min(nmin, xsmd_cells_along(b.L[a], rl));
    c.nc[a] = nmin;
  }
  for (BoxDev& b : c.box_h)
    for (int a = 0; a < 3; a++) b.inv_cs[a] = c.nc[a] / b.L[a];
  XS_HIP(hipMemcpyAsync(c.box_d.p, c.box_h.data(), sizeof(BoxDev) * c.box_h.size(),
                        hipMemcpyHostToDevice, c.stream));
  c.ncell_rep = (int64_t)c.nc[0] * c.nc[1] * c.nc[2];
  c.ncell_tot = c.ncell_rep * c.nrep;
  if (c.ncell_tot >= (int64_t)1 << 32) throw Error(XS_E_LIMIT, "too many cells");

  const unsigned g = grid_for(n);
  k_wrap_cellkey<<<g, kBlock, 0, c.stream>>>(n, c.rx.p, c.ry.p, c.rz.p, c.has_rep ? c.rep.p : nullptr,
                                             c.box_d.p, c.periodic[0], c.periodic[1], c.periodic[2],
                                             c.nc[0], c.nc[1], c.nc[2], c.ncell_rep, c.keys.p, c.vals.p);
  XS_HIP(hipGetLastError());

  int nbits = 1;
  while (((int64_t)1 << nbits) < c.ncell_tot) nbits++;
  size_t tmp_bytes = 0;
  XS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, c.keys.p, c.keys2.p, c.vals.p,
                                            c.vals2.p, (int)n, 0, nbits, c.stream));
  c.sort_tmp.reserve(tmp_bytes);
  XS_HIP(hipcub::DeviceRadixSort::SortPairs(c.sort_tmp.p, tmp_bytes, c.keys.p, c.keys2.p, c.vals.p,
                                            c.vals2.p, (int)n, 0, nbits, c.stream));

  PermuteArgs pa{};
  const double* src[12] = {c.rx.p, c.ry.p, c.rz.p, c.vx.p, c.vy.p, c.vz.p,
                           c.ax.p, c.ay.p, c.az.p, c.rx.p, c.ry.p, c.rz.p};
  double* dst[12] = {c.rx2.p, c.ry2.p, c.rz2.p, c.vx2.p, c.vy2.p, c.vz2.p,
                     c.ax2.p, c.ay2.p, c.az2.p, c.rrx.p, c.rry.p, c.rrz.p};
  for (int d = 0; d < 12; d++) {
    pa.src[d] = src[d];
    pa.dst[d] = dst[d];
  }
  pa.type_src = c.type.p;
  pa.type_dst = c.type2.p;
  pa.id_src = c.id.p;
  pa.id_dst = c.id2.p;
  pa.rep_src = c.has_rep ? c.rep.p : nullptr;
  pa.rep_dst = c.has_rep ? c.rep2.p : nullptr;
  k_permute<<<g, kBlock, 0, c.stream>>>(n, c.vals2.p, pa);
  XS_HIP(hipGetLastError());
  c.rx.swap(c.rx2);
  c.ry.swap(c.ry2);
  c.rz.swap(c.rz2);
  c.vx.swap(c.vx2);
  c.vy.swap(c.vy2);
  c.vz.swap(c.vz2);
  c.ax.swap(c.ax2);
  c.ay.swap(c.ay2);
  c.az.swap(c.az2);
  c.type.swap(c.type2);
  c.id.swap(c.id2);
  if (c.has_rep) c.rep.swap(c.rep2);
  c.rho_state = RhoState::None;  // per-atom EAM state is in the old order

  c.cstart.reserve(c.ncell_tot);
  c.cend.reserve(c.ncell_tot);
  XS_HIP(hipMemsetAsync(c.cstart.p, 0, sizeof(int) * c.ncell_tot, c.stream));
  XS_HIP(hipMemsetAsync(c.cend.p, 0, sizeof(int) * c.ncell_tot, c.stream));
  k_cell_bounds<<<g, kBlock, 0, c.stream>>>(n, c.keys2.p, c.cstart.p, c.cend.p);
  XS_HIP(hipGetLastError());

  // list capacity: estimate from the density, grow to the measured maximum on overflow
  if (c.nmax == 0) {
    double vol = 0;
    for (const BoxDev& b : c.box_h) vol += b.L[0] * b.L[1] * b.L[2];
    const double dens = (double)n / vol;
    const double est = dens * 4.18879020478639 * rl * rl * rl;  // 4/3 pi rl^3
    c.nmax = std::max(16, (int)std::ceil(est * 1.25 / 8.0) * 8);
  }
  const double rl2 = rl * rl;
  for (int attempt = 0; attempt < 2; attempt++) {
    c.nbr.reserve((size_t)c.nmax * c.npad);
    c.nnb.reserve(c.npad);
    XS_HIP(hipMemsetAsync(c.dscalar.p, 0, sizeof(unsigned long long), c.stream));
    k_build<<<g, kBlock, 0, c.stream>>>(n, c.rx.p, c.ry.p, c.rz.p, c.keys2.p, c.box_d.p, c.periodic[0],
                                        c.periodic[1], c.periodic[2], c.nc[0], c.nc[1], c.nc[2],
                                        c.ncell_rep, c.cstart.p, c.cend.p, rl2, c.nbr.p, c.nnb.p,
                                        c.nmax, c.npad, c.dscalar.p);
    XS_HIP(hipGetLastError());
    XS_HIP(hipMemcpyAsync(c.hscalar, c.dscalar.p, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          c.stream));
    XS_HIP(hipStreamSynchronize(c.stream));
    const unsigned long long need = c.hscalar[0];
    if (need == 0) break;
    if (attempt == 1) throw Error(XS_E_STATE, "neighbour list overflow after resize");
    c.nmax = (int)((need + 7) / 8 * 8 + 8);
  }
  c.list_valid = true;
  c.list_rcut = rcut;
  c.list_skin = skin;
}

namespace {

__global__ void k_max_displ2(int64_t n, const double* __restrict__ rx, const double* __restrict__ ry,
                             const double* __restrict__ rz, const double* __restrict__ rrx,
                             const double* __restrict__ rry, const double* __restrict__ rrz,
                             unsigned long long* __restrict__ out)
{
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double d2 = xsmd_d2(rx[i] - rrx[i], ry[i] - rry[i], rz[i] - rrz[i]);
    m = fmax(m, d2);
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

}  // namespace

double max_displacement2(Ctx& c)
{
  if (c.n == 0) return 0.0;
  XS_HIP(hipMemsetAsync(c.dscalar.p + 1, 0, sizeof(unsigned long long), c.stream));
  const unsigned g = std::min<unsigned>(grid_for(c.n), 2048);
  k_max_displ2<<<g, kBlock, 0, c.stream>>>(c.n, c.rx.p, c.ry.p, c.rz.p, c.rrx.p, c.rry.p, c.rrz.p,
                                          c.dscalar.p + 1);
  XS_HIP(hipGetLastError());
  XS_HIP(hipMemcpyAsync(c.hscalar + 1, c.dscalar.p + 1, sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, c.stream));
  XS_HIP(hipStreamSynchronize(c.stream));
  double d;
  std::memcpy(&d, &c.hscalar[1], sizeof(double));
  return d;
}

}  // namespace xsmd
