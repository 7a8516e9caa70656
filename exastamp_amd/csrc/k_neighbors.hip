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

__global__ void k_wrap_cellkey(int64_t n, Pos4* __restrict__ pos, const uint16_t* __restrict__ rep,
                               const BoxDev* __restrict__ box, int px, int py, int pz, int nx, int ny,
                               int nz, int64_t ncell_rep, uint32_t* __restrict__ keys,
                               uint32_t* __restrict__ vals, unsigned long long* __restrict__ outside)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = rep ? rep[i] : 0;
  const BoxDev& b = box[r];
  Pos4 p = load_pos(pos, i);
  if (px || py || pz) {
    if (px) p.x = xsmd_wrap(p.x, b.lo[0], b.L[0]);
    if (py) p.y = xsmd_wrap(p.y, b.lo[1], b.L[1]);
    if (pz) p.z = xsmd_wrap(p.z, b.lo[2], b.L[2]);
    store_pos(pos, i, p);
  }
  const int cx = xsmd_cell_coord(p.x, b.lo[0], b.inv_cs[0], nx);
  const int cy = xsmd_cell_coord(p.y, b.lo[1], b.inv_cs[1], ny);
  const int cz = xsmd_cell_coord(p.z, b.lo[2], b.inv_cs[2], nz);
  keys[i] = (uint32_t)((int64_t)r * ncell_rep + ((int64_t)cz * ny + cy) * nx + cx);
  vals[i] = (uint32_t)i;
  // an atom outside its own cell's slab (more than one box length away along a
  // periodic axis) voids the cell-box culling of the block build
  if (xsmd_cell_gap(p.x, b.lo[0], b.L[0], nx, cx, px) > 0.0 || xsmd_cell_gap(p.y, b.lo[1], b.L[1], ny, cy, py) > 0.0 ||
      xsmd_cell_gap(p.z, b.lo[2], b.L[2], nz, cz, pz) > 0.0)
    atomicOr(outside, 1ull);
}

struct PermuteArgs {
  const Pos4* pos_src;
  Pos4* pos_dst;
  double* rrx;
  double* rry;
  double* rrz;
  const double* src[6];
  double* dst[6];
  const uint8_t* type_src;
  uint8_t* type_dst;
  const int64_t* id_src;
  int64_t* id_dst;
  const uint16_t* rep_src;
  uint16_t* rep_dst;
};

// Gather every per-atom array into the new order; the new positions also become
// the reference positions of the displacement check.
__global__ void k_permute(int64_t n, const uint32_t* __restrict__ perm, PermuteArgs a)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = perm[i];
  const Pos4 q = load_pos(a.pos_src, p);
  store_pos(a.pos_dst, i, q);
  a.rrx[i] = q.x;
  a.rry[i] = q.y;
  a.rrz[i] = q.z;
#pragma unroll
  for (int d = 0; d < 6; d++) a.dst[d][i] = a.src[d][p];
  a.type_dst[i] = a.type_src[p];
  a.id_dst[i] = a.id_src[p];
  if (a.rep_src) a.rep_dst[i] = a.rep_src[p];
}

// One thread per atom; stencil order dz, dy, dx (outer to inner), in-cell order
// ascending -- the same visit order as the oracle, so lists match entry by entry.
__global__ void __launch_bounds__(256) k_build(
    int64_t n, const Pos4* __restrict__ pos, const uint32_t* __restrict__ keys,
    const BoxDev* __restrict__ box, int px, int py, int pz, int nx, int ny, int nz,
    int64_t ncell_rep, const int* __restrict__ cell_off, double rl2,
    uint32_t* __restrict__ nbr, int* __restrict__ nnb, int nmax, int64_t npad,
    unsigned long long* __restrict__ maxcnt, const XformDev* __restrict__ xf)
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
  const Pos4 pi = load_pos(pos, i);
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
        const int j0 = cell_off[cell], j1 = cell_off[cell + 1];
        const bool self_img = (kx == 0 && ky == 0 && kz == 0);
        for (int j = j0; j < j1; j++) {
          if (self_img && j == (int)i) continue;
          const Pos4 pj = load_pos(pos, j);
          double ddx = xsmd_image(pj.x, kx, b.L[0]) - pi.x;
          double ddy = xsmd_image(pj.y, ky, b.L[1]) - pi.y;
          double ddz = xsmd_image(pj.z, kz, b.L[2]) - pi.z;
          if (xf) xsmd_xform_apply(xf->m, &ddx, &ddy, &ddz);
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

// block-wide max of a non-negative double, one atomic per block
__device__ __forceinline__ void block_max_to(double v, unsigned long long* out)
{
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  __shared__ double sh[kBlock / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sh[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, sh[w]);
    if (m > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
  }
}

__global__ void __launch_bounds__(256) k_max_displ2(int64_t n, const Pos4* __restrict__ pos,
                                                    const double* __restrict__ rrx,
                                                    const double* __restrict__ rry,
                                                    const double* __restrict__ rrz,
                                                    unsigned long long* __restrict__ out,
                                                    const XformDev* __restrict__ xf)
{
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const Pos4 p = load_pos(pos, i);
    double ux = p.x - rrx[i], uy = p.y - rry[i], uz = p.z - rrz[i];
    if (xf) xsmd_xform_apply(xf->m, &ux, &uy, &uz);  // the threshold is a physical distance
    m = fmax(m, xsmd_d2(ux, uy, uz));
  }
  block_max_to(m, out);
}

__global__ void k_pair_count(int64_t n, int64_t npad, const Pos4* __restrict__ pos,
                             const uint16_t* __restrict__ rep, const BoxDev* __restrict__ box,
                             const uint32_t* __restrict__ nbr, const int* __restrict__ nnb, double rc2,
                             unsigned long long* __restrict__ out, const XformDev* __restrict__ xf)
{
  unsigned long long e = 0, w = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double* L = box[rep ? rep[i] : 0].L;
    const Pos4 pi = load_pos(pos, i);
    const int cnt = nnb[i];
    e += cnt;
    for (int k = 0; k < cnt; k++) {
      const uint32_t en = nbr[(int64_t)k * npad + i];
      const Pos4 pj = load_pos(pos, xsmd_nbr_index(en));
      double dx = xsmd_image(pj.x, xsmd_nbr_img_x(en), L[0]) - pi.x;
      double dy = xsmd_image(pj.y, xsmd_nbr_img_y(en), L[1]) - pi.y;
      double dz = xsmd_image(pj.z, xsmd_nbr_img_z(en), L[2]) - pi.z;
      if (xf) xsmd_xform_apply(xf->m, &dx, &dy, &dz);
      if (xsmd_d2(dx, dy, dz) <= rc2) w++;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off);
    w += __shfl_xor(w, off);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(out, e);
    atomicAdd(out + 1, w);
  }
}

}  // namespace

void build_neighbor_list(Ctx& c, double rcut, double skin, bool prefer_block)
{
  ProfScope prof(c.prof, c.stream, XS_PROF_NEIGHBORS);
  const int64_t n = c.n;
  c.half_valid = false;
  if (n == 0) {
    c.list_valid = true;
    c.list_rcut = rcut;
    c.list_skin = skin;
    c.list_kind = ListKind::Global;
    c.global_from_block = false;
    return;
  }
  const double rl = rcut + skin;
  if (!(rl > 0)) throw Error(XS_E_ARG, "rcut + skin must be positive");

  // cell grid: common dims for all replicas (from the smallest box), each
  // replica with its own n/L so every cell is at least rl wide (in grid space,
  // rl times the xform's reach along the axis)
  for (int a = 0; a < 3; a++) {
    const double rla = c.xf.on ? rl * xsmd_xform_reach(c.xf.mi, a) : rl;
    int nmin = 1 << 30;
    for (const BoxDev& b : c.box_h) nmin = std::min(nmin, xsmd_cells_along(b.L[a], rla));
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
  XS_HIP(hipMemsetAsync(c.dscalar.p + 4, 0, sizeof(unsigned long long), c.stream));
  k_wrap_cellkey<<<g, kBlock, 0, c.stream>>>(n, c.pos.p, c.has_rep ? c.rep.p : nullptr, c.box_d.p,
                                             c.periodic[0], c.periodic[1], c.periodic[2], c.nc[0], c.nc[1],
                                             c.nc[2], c.ncell_rep, c.keys.p, c.vals.p, c.dscalar.p + 4);
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
  pa.pos_src = c.pos.p;
  pa.pos_dst = c.pos2.p;
  pa.rrx = c.rrx.p;
  pa.rry = c.rry.p;
  pa.rrz = c.rrz.p;
  const double* src[6] = {c.vx.p, c.vy.p, c.vz.p, c.ax.p, c.ay.p, c.az.p};
  double* dst[6] = {c.vx2.p, c.vy2.p, c.vz2.p, c.ax2.p, c.ay2.p, c.az2.p};
  for (int d = 0; d < 6; d++) {
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
  c.pos.swap(c.pos2);
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
  c.rho_ckpt = false;

  cell_offsets(c);

  // list capacity: estimate from the density, grow to the measured maximum on overflow
  if (c.nmax == 0) {
    double vol = 0;
    for (const BoxDev& b : c.box_h) vol += b.L[0] * b.L[1] * b.L[2] * xform_det(c.xf);
    const double dens = (double)n / vol;
    const double est = dens * 4.18879020478639 * rl * rl * rl;  // 4/3 pi rl^3
    c.nmax = std::max(16, (int)std::ceil(est * 1.25 / 8.0) * 8);
  }
  const double rl2 = rl * rl;
  c.global_from_block = false;
  c.list_kind = ListKind::Global;
  if (prefer_block && !c.block_disabled) {
    block_setup(c);
    if (block_feasible(c)) {
      block_build(c, rl2);
      c.list_kind = ListKind::Block;
    }
  }
  for (int attempt = 0; c.list_kind == ListKind::Global && attempt < 2; attempt++) {
    c.nbr.reserve((size_t)c.nmax * c.npad);
    c.nnb.reserve(c.npad);
    XS_HIP(hipMemsetAsync(c.dscalar.p, 0, sizeof(unsigned long long), c.stream));
    k_build<<<g, kBlock, 0, c.stream>>>(n, c.pos.p, c.keys2.p, c.box_d.p, c.periodic[0], c.periodic[1],
                                        c.periodic[2], c.nc[0], c.nc[1], c.nc[2], c.ncell_rep, c.cell_off.p, rl2,
                                        c.nbr.p, c.nnb.p, c.nmax, c.npad, c.dscalar.p, c.xf_dev());
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

void pair_count(Ctx& c, double rcut, int64_t& entries, int64_t& within)
{
  entries = within = 0;
  if (c.n == 0) return;
  ensure_global_list(c);
  XS_HIP(hipMemsetAsync(c.dscalar.p + 2, 0, 2 * sizeof(unsigned long long), c.stream));
  const unsigned g = std::min<unsigned>(grid_for(c.n), 4096);
  k_pair_count<<<g, kBlock, 0, c.stream>>>(c.n, c.npad, c.pos.p, c.has_rep ? c.rep.p : nullptr, c.box_d.p,
                                           c.nbr.p, c.nnb.p, rcut * rcut, c.dscalar.p + 2, c.xf_dev());
  XS_HIP(hipGetLastError());
  XS_HIP(hipMemcpyAsync(c.hscalar + 2, c.dscalar.p + 2, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        c.stream));
  XS_HIP(hipStreamSynchronize(c.stream));
  entries = (int64_t)c.hscalar[2];
  within = (int64_t)c.hscalar[3];
}

double max_displacement2(Ctx& c)
{
  if (c.n == 0) return 0.0;
  XS_HIP(hipMemsetAsync(c.dscalar.p + 1, 0, sizeof(unsigned long long), c.stream));
  const unsigned g = std::min<unsigned>(grid_for(c.n), 2048);
  k_max_displ2<<<g, kBlock, 0, c.stream>>>(c.n, c.pos.p, c.rrx.p, c.rry.p, c.rrz.p, c.dscalar.p + 1, c.xf_dev());
  XS_HIP(hipGetLastError());
  XS_HIP(hipMemcpyAsync(c.hscalar + 1, c.dscalar.p + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        c.stream));
  XS_HIP(hipStreamSynchronize(c.stream));
  double d;
  std::memcpy(&d, &c.hscalar[1], sizeof(double));
  return d;
}

}  // namespace xsmd
