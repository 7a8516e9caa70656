// k_block.hip -- cell-block (LDS-staged) neighbour list and EAM passes.
//
// The cell grid is cut into bricks of bx*by*bz cells.  One workgroup owns one
// brick: it stages the brick's halo -- the brick plus one cell on every side,
// periodic images applied once at staging -- into LDS as 32-byte records
// {x, y, z, w}, then each thread walks the list of one central atom.  List
// entries are 16-bit halo-local slots, so a pair costs one LDS record read
// instead of a global gather, and the list is half the size of the global one.
//
// The halo of a brick is an (ex+2)x(ey+2)x(ez+2) box of cells enumerated z-major,
// x fastest.  Atoms of one cell are contiguous in HBM (cell-sorted order) and are
// staged contiguously, cells in enumeration order, so the local slot order equals
// the visit order of the global build: a block list maps back to exactly the
// global list (same entries, same order).
#include "pair_math.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace xsmd {
namespace {

struct HaloCell {
  int gstart;  // first atom of the cell in HBM
  int lstart;  // first slot in LDS
  int count;
  int img;     // periodic image, 2 bits per axis as in list entries (bits 26..31 >> 26)
};

struct BlockShape {
  int r;                 // replica
  int cx0, cy0, cz0;     // first central cell
  int ex, ey, ez;        // central cells along each axis (clipped at the grid edge)
  int hx, hy, hz;        // halo dims = e + 2
};

__device__ __forceinline__ BlockShape block_shape(const BlockGeom& g, int64_t b)
{
  BlockShape s;
  const int64_t per = (int64_t)g.nbx * g.nby * g.nbz;
  s.r = (int)(b / per);
  int64_t q = b - (int64_t)s.r * per;
  const int ibx = (int)(q % g.nbx);
  q /= g.nbx;
  const int iby = (int)(q % g.nby);
  const int ibz = (int)(q / g.nby);
  s.cx0 = ibx * g.bx;
  s.cy0 = iby * g.by;
  s.cz0 = ibz * g.bz;
  s.ex = min(g.bx, g.nx - s.cx0);
  s.ey = min(g.by, g.ny - s.cy0);
  s.ez = min(g.bz, g.nz - s.cz0);
  s.hx = s.ex + 2;
  s.hy = s.ey + 2;
  s.hz = s.ez + 2;
  return s;
}

__device__ __forceinline__ bool wrap_axis(int c, int n, int periodic, int& w, int& k)
{
  k = 0;
  w = c;
  if (c < 0) {
    if (!periodic) return false;
    w = c + n;
    k = -1;
  } else if (c >= n) {
    if (!periodic) return false;
    w = c - n;
    k = 1;
  }
  return true;
}

// Staged halo records in LDS as two 16-byte arrays, {x, y} and {z, w}: a random
// 16-byte read then starts on any of 16 bank groups instead of the 8 that 32-byte
// records allow, which halves the expected bank-conflict depth of a gather.
struct LRec {
  double2* xy;
  double2* zw;
  __device__ __forceinline__ Pos4 get(int c) const
  {
    const double2 a = xy[c];
    const double2 b = zw[c];
    return {a.x, a.y, b.x, b.y};
  }
  __device__ __forceinline__ void put(int c, const Pos4& p) const
  {
    xy[c] = make_double2(p.x, p.y);
    zw[c] = make_double2(p.z, p.w);
  }
};

// Halo table of block b in LDS (hc[0..nh)), filled by all threads; returns the
// number of halo atoms.  Ends with __syncthreads().
__device__ int build_halo(const BlockGeom& g, const BlockShape& s, const int* __restrict__ cell_off,
                          HaloCell* hc, int* total_out)
{
  const int nh = s.hx * s.hy * s.hz;
  for (int h = threadIdx.x; h < nh; h += blockDim.x) {
    const int ix = h % s.hx, iy = (h / s.hx) % s.hy, iz = h / (s.hx * s.hy);
    int wx, wy, wz, kx, ky, kz;
    HaloCell e{0, 0, 0, 0};
    if (wrap_axis(s.cx0 - 1 + ix, g.nx, g.px, wx, kx) && wrap_axis(s.cy0 - 1 + iy, g.ny, g.py, wy, ky) &&
        wrap_axis(s.cz0 - 1 + iz, g.nz, g.pz, wz, kz)) {
      const int64_t c = (int64_t)s.r * g.ncell_rep + ((int64_t)wz * g.ny + wy) * g.nx + wx;
      e.gstart = cell_off[c];
      e.count = cell_off[c + 1] - e.gstart;
      e.img = (int)(xsmd_nbr_encode(0u, kx, ky, kz) >> XSMD_NBR_INDEX_BITS);
    }
    hc[h] = e;
  }
  __syncthreads();
  // exclusive prefix sum of the counts in the first wave, 64 cells at a time
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    int carry = 0;
    for (int base = 0; base < nh; base += 64) {
      const int h = base + lane;
      const int cnt = (h < nh) ? hc[h].count : 0;
      int v = cnt;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(v, off);
        if (lane >= off) v += t;
      }
      if (h < nh) hc[h].lstart = carry + v - cnt;
      carry += __shfl(v, 63);
    }
    if (lane == 0) *total_out = carry;
  }
  __syncthreads();
  return *total_out;
}

// halo cell holding local slot c (binary search over lstart)
__device__ __forceinline__ int find_cell(const HaloCell* hc, int nh, int c)
{
  int lo = 0, hi = nh - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (hc[mid].lstart <= c) lo = mid;
    else hi = mid - 1;
  }
  // the last cell starting at or before c: among cells sharing one lstart only the
  // last can be non-empty, so this is the cell that holds slot c
  return lo;
}

// Stage halo atoms: record c = position (+ image shift) and w; type when requested.
__device__ void stage(const BlockShape& s, const HaloCell* hc, int total, const Pos4* __restrict__ pos,
                      const double* L, const uint8_t* __restrict__ type, const LRec& rec, uint8_t* ltype)
{
  const int nh = s.hx * s.hy * s.hz;
  for (int c = threadIdx.x; c < total; c += blockDim.x) {
    const int h = find_cell(hc, nh, c);
    const HaloCell e = hc[h];
    const int j = e.gstart + (c - e.lstart);
    Pos4 p = load_pos(pos, j);
    if (e.img) {
      const uint32_t code = (uint32_t)e.img << XSMD_NBR_INDEX_BITS;
      p.x = xsmd_image(p.x, xsmd_nbr_img_x(code), L[0]);
      p.y = xsmd_image(p.y, xsmd_nbr_img_y(code), L[1]);
      p.z = xsmd_image(p.z, xsmd_nbr_img_z(code), L[2]);
    }
    rec.put(c, p);
    if (ltype) ltype[c] = type[j];
  }
  __syncthreads();
}

// Central atom q of the block: its halo cell h, local slot and global index.
struct Central {
  int h, local;
  int64_t global;
};

__device__ __forceinline__ bool central_atom(const BlockShape& s, const HaloCell* hc, int q, Central& out)
{
  // central cells are ix in [1, ex], iy in [1, ey], iz in [1, ez]; walk rows
  for (int iz = 1; iz <= s.ez; iz++)
    for (int iy = 1; iy <= s.ey; iy++) {
      const int h0 = (iz * s.hy + iy) * s.hx + 1;
      const int h1 = h0 + s.ex - 1;
      const int n = hc[h1].lstart + hc[h1].count - hc[h0].lstart;
      if (q < n) {
        const int local = hc[h0].lstart + q;
        int h = h0;
        while (hc[h].lstart + hc[h].count <= local) h++;
        out.h = h;
        out.local = local;
        out.global = (int64_t)hc[h0].gstart + q;
        return true;
      }
      q -= n;
    }
  return false;
}

__device__ __forceinline__ int64_t lblock()
{
  const unsigned b = blockIdx.x, g = gridDim.x;
  return (int64_t)(b & 7u) * (g >> 3) + (b >> 3);
}

struct LdsLayout {
  LRec rec;
  HaloCell* hc;
  uint8_t* ltype;
  int* total;
};

__device__ __forceinline__ LdsLayout lds_layout(int cap, int nh_max, bool with_type)
{
  extern __shared__ __align__(32) unsigned char smem[];
  LdsLayout l;
  l.rec.xy = reinterpret_cast<double2*>(smem);
  l.rec.zw = reinterpret_cast<double2*>(smem + (size_t)cap * sizeof(double2));
  l.hc = reinterpret_cast<HaloCell*>(smem + (size_t)cap * sizeof(Pos4));
  l.total = reinterpret_cast<int*>(l.hc + nh_max);
  l.ltype = with_type ? reinterpret_cast<uint8_t*>(l.total + 4) : nullptr;
  return l;
}

// Two lanes per central atom: the two halves of a wave take the same 32 atoms
// (list build: alternate candidates; EAM passes: even and odd list entries).
constexpr int kPassThreads = 512;
constexpr int kPassAtoms = kPassThreads / 2;

// sum of v over the two halves of the wave (lanes l and l ^ 32)
__device__ __forceinline__ double half_sum(double v) { return v + __shfl_xor(v, 32); }

// this lane's atom slot and which half of the wave it is in
__device__ __forceinline__ int pass_slot() { return (threadIdx.x >> 6) * 32 + (threadIdx.x & 31); }
__device__ __forceinline__ int pass_half() { return (threadIdx.x >> 5) & 1; }

// ---------------------------------------------------------------------------
// list build
// ---------------------------------------------------------------------------
// Squared distance from p to the box of the halo cell at unwrapped grid coordinate
// (gx, gy, gz), widened by a margin that covers the rounding of cell assignment.
// Along a non-periodic axis the edge cells also hold the atoms clamped into them,
// so their boxes extend to infinity there.
__device__ __forceinline__ double cell_box_d2(const BlockGeom& g, const BoxDev& b, int gx, int gy, int gz,
                                              const Pos4& p)
{
  const double ux = xsmd_cell_gap(p.x, b.lo[0], b.L[0], g.nx, gx, g.px);
  const double uy = xsmd_cell_gap(p.y, b.lo[1], b.L[1], g.ny, gy, g.py);
  const double uz = xsmd_cell_gap(p.z, b.lo[2], b.L[2], g.nz, gz, g.pz);
  return xsmd_d2(ux, uy, uz);
}

// Block list: stencil order dz, dy, dx, in-cell order ascending (the order of the
// global build).  Two lanes per atom test alternate candidates of a cell and place
// their hits in candidate order; a neighbour cell whose box lies beyond rl is
// skipped (its atoms would all fail the exact test), unless some atom sits outside
// its cell's box (nocull, set by the cell-key kernel).
__global__ void __launch_bounds__(kPassThreads) k_block_build(BlockGeom g, const int* __restrict__ cell_off,
                                                              const Pos4* __restrict__ pos,
                                                              const BoxDev* __restrict__ box, double rl2,
                                                              uint16_t* __restrict__ nbr16, int* __restrict__ nnb,
                                                              int nmax, int64_t npad, int cap, int nh_max,
                                                              const unsigned long long* __restrict__ nocull,
                                                              int* __restrict__ brick_rows,
                                                              unsigned long long* __restrict__ flags)
{
  const int64_t b = lblock();
  if (b >= g.nblocks) return;
  LdsLayout l = lds_layout(cap, nh_max, false);
  const BlockShape s = block_shape(g, b);
  const int total = build_halo(g, s, cell_off, l.hc, l.total);
  if (total > cap) {  // cannot happen when cap = measured maximum; never write past LDS
    if (threadIdx.x == 0) atomicMax(flags + 1, (unsigned long long)total);
    return;
  }
  const BoxDev bx = box[s.r];
  const bool cull = *nocull == 0ull;
  if (threadIdx.x == 0) l.total[1] = 0;  // longest list of the brick (made visible by stage's barrier)
  stage(s, l.hc, total, pos, bx.L, nullptr, l.rec, nullptr);
  const int half = pass_half();
  for (int q = pass_slot();; q += kPassAtoms) {
    Central a;
    if (!central_atom(s, l.hc, q, a)) break;
    const Pos4 pi = l.rec.get(a.local);
    const int ix = a.h % s.hx, iy = (a.h / s.hx) % s.hy, iz = a.h / (s.hx * s.hy);
    int cnt = 0;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const HaloCell e = l.hc[((iz + dz) * s.hy + (iy + dy)) * s.hx + (ix + dx)];
          if (e.count == 0) continue;
          if (cull && cell_box_d2(g, bx, s.cx0 - 1 + ix + dx, s.cy0 - 1 + iy + dy, s.cz0 - 1 + iz + dz, pi) >= rl2)
            continue;
          for (int off = 0; off < e.count; off += 2) {
            const int k = off + half;
            const int c = e.lstart + k;
            bool hit = false;
            if (k < e.count && c != a.local) {
              const Pos4 pj = l.rec.get(c);
              hit = xsmd_d2(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z) < rl2;
            }
            const int other = __shfl_xor((int)hit, 32);
            const int slot = cnt + (half ? other : 0);
            if (hit && slot < nmax) nbr16[(int64_t)slot * npad + a.global] = (uint16_t)c;
            cnt += (int)hit + other;
          }
        }
    if (half == 0) {
      nnb[a.global] = cnt;
      atomicMax(l.total + 1, cnt);
      if (cnt > nmax) atomicMax(flags, (unsigned long long)cnt);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) brick_rows[b] = l.total[1];
}

// block-local entries -> global (index, image) entries of the same list
__global__ void __launch_bounds__(256) k_block_to_global(BlockGeom g, const int* __restrict__ cell_off,
                                                         const uint16_t* __restrict__ nbr16,
                                                         const int* __restrict__ nnb, uint32_t* __restrict__ nbr,
                                                         int64_t npad, int nh_max)
{
  const int64_t b = lblock();
  if (b >= g.nblocks) return;
  LdsLayout l = lds_layout(0, nh_max, false);
  const BlockShape s = block_shape(g, b);
  const int nh = s.hx * s.hy * s.hz;
  build_halo(g, s, cell_off, l.hc, l.total);
  for (int q = threadIdx.x;; q += blockDim.x) {
    Central a;
    if (!central_atom(s, l.hc, q, a)) break;
    const int cnt = nnb[a.global];
    for (int k = 0; k < cnt; k++) {
      const int c = nbr16[(int64_t)k * npad + a.global];
      const HaloCell e = l.hc[find_cell(l.hc, nh, c)];
      nbr[(int64_t)k * npad + a.global] =
          (uint32_t)(e.gstart + (c - e.lstart)) | ((uint32_t)e.img << XSMD_NBR_INDEX_BITS);
    }
  }
}

// per-block halo atom count (sizes the LDS of the block kernels)
__global__ void k_block_halo_max(BlockGeom g, const int* __restrict__ cell_off,
                                 unsigned long long* __restrict__ out)
{
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.nblocks) return;
  const BlockShape s = block_shape(g, b);
  int total = 0, central = 0;
  for (int iz = 0; iz < s.hz; iz++)
    for (int iy = 0; iy < s.hy; iy++)
      for (int ix = 0; ix < s.hx; ix++) {
        int wx, wy, wz, kx, ky, kz;
        if (!wrap_axis(s.cx0 - 1 + ix, g.nx, g.px, wx, kx) || !wrap_axis(s.cy0 - 1 + iy, g.ny, g.py, wy, ky) ||
            !wrap_axis(s.cz0 - 1 + iz, g.nz, g.pz, wz, kz))
          continue;
        const int64_t c = (int64_t)s.r * g.ncell_rep + ((int64_t)wz * g.ny + wy) * g.nx + wx;
        const int n = cell_off[c + 1] - cell_off[c];
        total += n;
        if (ix >= 1 && ix <= s.ex && iy >= 1 && iy <= s.ey && iz >= 1 && iz <= s.ez) central += n;
      }
  atomicMax(out, (unsigned long long)total);
  atomicMax(out + 1, (unsigned long long)central);
}

__global__ void k_cell_offsets(int64_t n, int64_t ncell, const uint32_t* __restrict__ keys, int* __restrict__ off)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const int64_t hi = (i == n) ? ncell : (int64_t)keys[i];
  const int64_t lo = (i == 0) ? -1 : (int64_t)keys[i - 1];
  for (int64_t c = lo + 1; c <= hi; c++) off[c] = (int)i;
}

// ---------------------------------------------------------------------------
// EAM passes on the block list
// ---------------------------------------------------------------------------
struct BlockArgs {
  BlockGeom g;
  const int* __restrict__ cell_off;
  const Pos4* __restrict__ pos;
  const uint8_t* __restrict__ type;
  const BoxDev* __restrict__ box;
  const uint16_t* __restrict__ nbr16;
  const int* __restrict__ nnb;
  int64_t npad;
  int cap;
  double rc2;
};

constexpr int kBBatch = 4;
// EAM passes run two lanes per central atom: the two halves of a wave take the
// even and odd list entries of the same 32 atoms and combine their sums with one
// cross-half exchange, doubling the waves that hide the table-gather latency for
// the LDS a brick needs.
// Minimum waves per SIMD the register allocation must allow: three 8-wave
// workgroups per CU (what the LDS of a ~200-atom brick permits) need 6; the
// virial form of the force pass carries 9 more sums and is held to 4.
template <bool VIRIAL>
constexpr int kPassWaves = VIRIAL ? 4 : 6;

// FP64 displacement to the staged record of slot c (image already applied)
struct LNb {
  double x, y, z, d2, w;
  int c;
  bool in;
};

// entries k0, k0+2, ..., k0+2(U-1) of atom gi (this lane's parity of the list)
template <int U>
__device__ __forceinline__ void lfetch(const BlockArgs& a, const LRec& rec, int64_t gi, int self, int k0, int cnt,
                                       const Pos4& pi, LNb (&nb)[U])
{
  // entries past this atom's count become the atom itself (a masked pair at d2 = 0)
  int c[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int k = k0 + 2 * u;
    c[u] = (k < cnt) ? (int)a.nbr16[(int64_t)k * a.npad + gi] : self;
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    const Pos4 pj = rec.get(c[u]);
    nb[u].x = pj.x - pi.x;
    nb[u].y = pj.y - pi.y;
    nb[u].z = pj.z - pi.z;
    nb[u].d2 = xsmd_d2(nb[u].x, nb[u].y, nb[u].z);
    nb[u].w = pj.w;
    nb[u].c = c[u];
    nb[u].in = (k0 + 2 * u < cnt) && nb[u].d2 <= a.rc2;
  }
}


// rho_i, then (EMB) F(rho_i) and F'(rho_i) for the central atoms of one brick
template <class Real, bool MULTI, bool EMB, bool ENERGY, bool SET_EP>
__global__ void __launch_bounds__(kPassThreads, kPassWaves<false>) k_block_eam_rho(BlockArgs a, EamTabDev<Real> t,
                                                                const Real* __restrict__ pair,
                                                                double* __restrict__ rho_io,
                                                                double* __restrict__ pos_w, double* __restrict__ ep)
{
  const int64_t b = lblock();
  if (b >= a.g.nblocks) return;
  LdsLayout l = lds_layout(a.cap, a.g.nh_max, MULTI);
  const BlockShape s = block_shape(a.g, b);
  const int total = build_halo(a.g, s, a.cell_off, l.hc, l.total);
  if (total > a.cap) return;  // excluded by the host (cap = measured maximum)
  stage(s, l.hc, total, a.pos, a.box[s.r].L, a.type, l.rec, l.ltype);
  const int half = pass_half();
  for (int q = pass_slot();; q += kPassAtoms) {
    Central ca;
    if (!central_atom(s, l.hc, q, ca)) break;
    const Pos4 pi = l.rec.get(ca.local);
    const int ti = MULTI ? l.ltype[ca.local] : 0;
    const int cnt = a.nnb[ca.global];
    double rho = 0.0;
    for (int k0 = half; k0 < cnt; k0 += 2 * kBBatch) {
      LNb nb[kBBatch];
      lfetch<kBBatch>(a, l.rec, ca.global, ca.local, k0, cnt, pi, nb);
#pragma unroll
      for (int u = 0; u < kBBatch; u++) {
        const int tj = MULTI ? l.ltype[nb[u].c] : 0;
        const Real val = pm::eam_rho<Real, MULTI>(t, pair, ti, tj, nb[u].in ? nb[u].d2 : 1.0);
        rho += nb[u].in ? (double)val : 0.0;
      }
    }
    rho = half_sum(rho);
    if (half == 0) {
      rho_io[ca.global] = rho;
      if constexpr (EMB) {
        Real F, Fp;
        pm::embed(t, ti, rho, F, Fp);
        pos_w[4 * ca.global + 3] = (double)Fp;
        if constexpr (ENERGY) {
          if constexpr (SET_EP) ep[ca.global] = (double)F;
          else ep[ca.global] += (double)F;
        }
      }
    }
  }
}

template <class Real, bool MULTI, bool ENERGY, bool VIRIAL, bool FUSED, bool DRIFT>
__global__ void __launch_bounds__(kPassThreads, kPassWaves<VIRIAL>) k_block_eam_force(BlockArgs a, EamTabDev<Real> t,
                                                                  const Real* __restrict__ pair, pm::Outputs o,
                                                                  pm::Drift dr)
{
  static_assert(!DRIFT || (FUSED && !ENERGY && !VIRIAL), "the drift epilogue is the plain MD-loop step");
  const int64_t b = lblock();
  if (b >= a.g.nblocks) return;
  LdsLayout l = lds_layout(a.cap, a.g.nh_max, MULTI);
  const BlockShape s = block_shape(a.g, b);
  const int total = build_halo(a.g, s, a.cell_off, l.hc, l.total);
  if (total > a.cap) return;  // excluded by the host (cap = measured maximum)
  stage(s, l.hc, total, a.pos, a.box[s.r].L, a.type, l.rec, l.ltype);
  const int half = pass_half();
  for (int q = pass_slot();; q += kPassAtoms) {
    Central ca;
    if (!central_atom(s, l.hc, q, ca)) break;
    const Pos4 pi = l.rec.get(ca.local);
    const int ti = MULTI ? l.ltype[ca.local] : 0;
    const Real fpi = (Real)pi.w;
    const int cnt = a.nnb[ca.global];
    double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k0 = half; k0 < cnt; k0 += 2 * kBBatch) {
      LNb nb[kBBatch];
      lfetch<kBBatch>(a, l.rec, ca.global, ca.local, k0, cnt, pi, nb);
#pragma unroll
      for (int u = 0; u < kBBatch; u++) {
        const int tj = MULTI ? l.ltype[nb[u].c] : 0;
        Real fp_r, phi;
        pm::eam_pair<Real, MULTI>(t, pair, ti, tj, fpi, (Real)nb[u].w, nb[u].in ? nb[u].d2 : 1.0, fp_r, phi);
        const double fpair = nb[u].in ? (double)fp_r : 0.0;
        const double fe0 = nb[u].x * fpair, fe1 = nb[u].y * fpair, fe2 = nb[u].z * fpair;
        fx += fe0;
        fy += fe1;
        fz += fe2;
        if constexpr (ENERGY) e += nb[u].in ? 0.5 * (double)phi : 0.0;
        if constexpr (VIRIAL) pm::add_virial(v, fe0, fe1, fe2, nb[u].x, nb[u].y, nb[u].z);
      }
    }
    fx = half_sum(fx);
    fy = half_sum(fy);
    fz = half_sum(fz);
    if constexpr (ENERGY) e = half_sum(e);
    if constexpr (VIRIAL) {
#pragma unroll
      for (int k = 0; k < 9; k++) v[k] = half_sum(v[k]);
    }
    if (half == 0) {
      if constexpr (DRIFT) pm::write_outputs_drift(o, dr, ca.global, pi, fx, fy, fz);
      else pm::write_outputs<ENERGY, VIRIAL, FUSED>(o, ca.global, fx, fy, fz, e, v);
    }
  }
}

// ---------------------------------------------------------------------------
// Persistent passes with the spline table in LDS (single element, FP32 tables)
// ---------------------------------------------------------------------------
// A wave's table gathers touch one cache line per distinct row, so rows read from
// HBM/L2 keep the texture data path busy for ~50 cycles per gather.  Here one
// 1024-thread workgroup per CU copies rows [m_lo, m_lo + rows) of the combined
// table into LDS once per launch and then walks a contiguous run of bricks, so
// table rows and neighbour records both come from LDS.  A pair whose knot index
// lies below m_lo (an unusually short pair) reads its row from global memory.
// Four lanes per central atom take list entries k = h, h + 4, ...; the lanes of
// one parity h are adjacent within the wave, so list reads stay contiguous.
constexpr int kTabThreads = 1024;
constexpr int kTabLanes = 4;
constexpr int kTabGroup = 64 / kTabLanes;           // atoms per wave
constexpr int kTabAtoms = kTabThreads / kTabLanes;  // atoms per pass over a brick
constexpr int kTBatch = 2;

struct TabArgs {
  BlockArgs a;
  int m_lo, rows;     // LDS holds table rows m_lo .. m_lo + rows - 1 (the last is nr - 1)
  int64_t per_wg;     // bricks per workgroup
  const int* __restrict__ brick_rows;  // longest list of each brick
  int nmax;           // rows of the list
};

__device__ __forceinline__ int tab_slot() { return (threadIdx.x >> 6) * kTabGroup + (threadIdx.x & (kTabGroup - 1)); }
__device__ __forceinline__ int tab_h() { return (threadIdx.x & 63) / kTabGroup; }

// sum over the four lanes of one atom (lanes l, l ^ 16, l ^ 32, l ^ 48)
__device__ __forceinline__ double group_sum(double v)
{
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}

struct TabLayout {
  float* tab;
  LRec rec;
  HaloCell* hc;
  int* total;
};

__device__ __forceinline__ TabLayout tab_layout(int rows, int row_floats, int cap, int nh_max)
{
  extern __shared__ __align__(32) unsigned char smem[];
  TabLayout l;
  l.tab = reinterpret_cast<float*>(smem);
  unsigned char* p = smem + (size_t)rows * row_floats * sizeof(float);
  l.rec.xy = reinterpret_cast<double2*>(p);
  l.rec.zw = reinterpret_cast<double2*>(p + (size_t)cap * sizeof(double2));
  l.hc = reinterpret_cast<HaloCell*>(p + (size_t)cap * sizeof(Pos4));
  l.total = reinterpret_cast<int*>(l.hc + nh_max);
  return l;
}

// list slots of entries k0, k0 + 4, ... of atom gi (the atom itself past the count).
// The row read is clamped into [0, cnt - 1], so every load is unconditional and in
// bounds, and the loads of the next batch can stay in flight while this one is used.
template <int U>
__device__ __forceinline__ void tlist(const BlockArgs& a, int64_t gi, int self, int k0, int cnt, int (&c)[U])
{
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int k = k0 + kTabLanes * u;
    const int kk = max(0, min(k, cnt - 1));
    const int raw = a.nbr16[(int64_t)kk * a.npad + gi];
    c[u] = (k < cnt) ? raw : self;
  }
}

// neighbour records of list slots c (entries k0, k0 + 4, ...); entries past the
// count are masked and evaluated at d2 = rc^2, whose knot is in LDS
template <int U>
__device__ __forceinline__ void tgather(const BlockArgs& a, const LRec& rec, const int (&c)[U], int k0, int cnt,
                                        const Pos4& pi, LNb (&nb)[U])
{
#pragma unroll
  for (int u = 0; u < U; u++) {
    const Pos4 pj = rec.get(c[u]);
    nb[u].x = pj.x - pi.x;
    nb[u].y = pj.y - pi.y;
    nb[u].z = pj.z - pi.z;
    nb[u].d2 = xsmd_d2(nb[u].x, nb[u].y, nb[u].z);
    nb[u].w = pj.w;
    nb[u].c = c[u];
    nb[u].in = (k0 + kTabLanes * u < cnt) && nb[u].d2 <= a.rc2;
  }
}

// value cubic of rho at knot m: LDS row when m >= m_lo, else the HBM row
__device__ __forceinline__ pm::Q4<float> tab_rho_row(const float* ltab, int m_lo, const float* __restrict__ g, int m)
{
  if (m >= m_lo) {
    const float4 v = reinterpret_cast<const float4*>(ltab)[m - m_lo];
    return {v.x, v.y, v.z, v.w};
  }
  return pm::load4(g + (int64_t)m * 8);
}

__device__ __forceinline__ void tab_pair_rows(const float* ltab, int m_lo, const float* __restrict__ g, int m,
                                              pm::Q4<float>& rc, pm::Q4<float>& zc)
{
  if (m >= m_lo) {
    const float4* q = reinterpret_cast<const float4*>(ltab) + (int64_t)(m - m_lo) * 2;
    const float4 u = q[0];
    const float4 w = q[1];
    rc = {u.x, u.y, u.z, u.w};
    zc = {w.x, w.y, w.z, w.w};
  } else {
    rc = pm::load4(g + (int64_t)m * 8);
    zc = pm::load4(g + (int64_t)m * 8 + 4);
  }
}

// ---- per-brick pipeline of the persistent passes ----
// Global loads of a brick form one round trip: the halo cells of brick b + 1
// (cell offsets and the brick's longest list) are fetched while brick b computes;
// at the start of a brick every lane issues its atom's count and list entries and
// then the halo record loads, all before the first wait; the pair loop itself
// reads only LDS and registers.
constexpr int kTabPL = 16;  // list entries per lane held in registers (rows < 4 * kTabPL), two per VGPR

// this thread's halo cell of a brick (threads >= the cell count hold nothing)
struct HaloNext {
  int gstart, count, img, rows;
};

__device__ __forceinline__ void halo_fetch(const BlockGeom& g, const int* __restrict__ cell_off,
                                           const int* __restrict__ brick_rows, int64_t b, HaloNext& hn)
{
  const BlockShape s = block_shape(g, b);
  const int nh = s.hx * s.hy * s.hz;
  hn = {0, 0, 0, brick_rows[b]};
  const int h = threadIdx.x;
  if (h < nh) {
    const int ix = h % s.hx, iy = (h / s.hx) % s.hy, iz = h / (s.hx * s.hy);
    int wx, wy, wz, kx, ky, kz;
    if (wrap_axis(s.cx0 - 1 + ix, g.nx, g.px, wx, kx) && wrap_axis(s.cy0 - 1 + iy, g.ny, g.py, wy, ky) &&
        wrap_axis(s.cz0 - 1 + iz, g.nz, g.pz, wz, kz)) {
      const int64_t c = (int64_t)s.r * g.ncell_rep + ((int64_t)wz * g.ny + wy) * g.nx + wx;
      hn.gstart = cell_off[c];
      hn.count = cell_off[c + 1] - hn.gstart;
      hn.img = (int)(xsmd_nbr_encode(0u, kx, ky, kz) >> XSMD_NBR_INDEX_BITS);
    }
  }
}

// halo table of the brick from the fetched cells, with the exclusive prefix of the
// counts (as build_halo); ends with __syncthreads()
__device__ int halo_commit(const BlockShape& s, const HaloNext& hn, HaloCell* hc, int* total_out)
{
  const int nh = s.hx * s.hy * s.hz;
  if ((int)threadIdx.x < nh) hc[threadIdx.x] = HaloCell{hn.gstart, 0, hn.count, hn.img};
  __syncthreads();
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    int carry = 0;
    for (int base = 0; base < nh; base += 64) {
      const int h = base + lane;
      const int cnt = (h < nh) ? hc[h].count : 0;
      int v = cnt;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(v, off);
        if (lane >= off) v += t;
      }
      if (h < nh) hc[h].lstart = carry + v - cnt;
      carry += __shfl(v, 63);
    }
    if (lane == 0) *total_out = carry;
  }
  __syncthreads();
  return *total_out;
}

// Halo records, at most two per thread (cap <= 2 * kTabThreads): all loads are
// issued before any record is written.  Ends with __syncthreads().
__device__ __forceinline__ void stage2(const BlockShape& s, const HaloCell* hc, int total,
                                       const Pos4* __restrict__ pos, const double* L, const LRec& rec)
{
  const int nh = s.hx * s.hy * s.hz;
  Pos4 p[2];
  int img[2];
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int c = threadIdx.x + r * kTabThreads;
    img[r] = 0;
    if (c < total) {
      const HaloCell e = hc[find_cell(hc, nh, c)];
      p[r] = load_pos(pos, e.gstart + (c - e.lstart));
      img[r] = e.img;
    }
  }
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int c = threadIdx.x + r * kTabThreads;
    if (c < total) {
      Pos4 q = p[r];
      if (img[r]) {
        const uint32_t code = (uint32_t)img[r] << XSMD_NBR_INDEX_BITS;
        q.x = xsmd_image(q.x, xsmd_nbr_img_x(code), L[0]);
        q.y = xsmd_image(q.y, xsmd_nbr_img_y(code), L[1]);
        q.z = xsmd_image(q.z, xsmd_nbr_img_z(code), L[2]);
      }
      rec.put(c, q);
    }
  }
  __syncthreads();
}

// list entries h, h + 4, ... of atom gi into registers (entry u in half u & 1 of
// E[u / 2]), rows < 4 * min(kTabPL, units) (units = the brick's longest list / 4,
// rounded up); rows past this atom's count hold another atom's data and are masked
// by the count later
__device__ __forceinline__ void list_fetch(const BlockArgs& a, int nmax, int64_t gi, int h, int units,
                                           uint32_t (&E)[kTabPL / 2])
{
#pragma unroll
  for (int u = 0; u < kTabPL; u += 2) {
    uint32_t lo = 0, hi = 0;
    if (u < units) lo = a.nbr16[(int64_t)min(h + kTabLanes * u, nmax - 1) * a.npad + gi];
    if (u + 1 < units) hi = a.nbr16[(int64_t)min(h + kTabLanes * (u + 1), nmax - 1) * a.npad + gi];
    E[u / 2] = lo | (hi << 16);
  }
}

__device__ __forceinline__ int list_entry(const uint32_t (&E)[kTabPL / 2], int u)
{
  return (int)((E[u / 2] >> (16 * (u & 1))) & 0xffffu);
}

// EAM density of central atom ca: entries from E (and, past 4 * kTabPL entries, from HBM)
template <int U>
__device__ __forceinline__ double tab_rho_entries(const TabArgs& ta, const EamTabDev<float>& t,
                                                  const float* __restrict__ pair, const float* ltab, const LRec& rec,
                                                  const Central& ca, const Pos4& pi, int h, int cnt,
                                                  const uint32_t (&E)[kTabPL / 2])
{
  const BlockArgs& a = ta.a;
  double rho = 0.0;
#pragma unroll 1
  for (int u0 = 0; u0 < kTabPL; u0 += U) {
    if (h + kTabLanes * u0 >= cnt) break;
    int c[U];
#pragma unroll
    for (int u = 0; u < U; u++) c[u] = (h + kTabLanes * (u0 + u) < cnt) ? list_entry(E, u0 + u) : ca.local;
    LNb nb[U];
    tgather<U>(a, rec, c, h + kTabLanes * u0, cnt, pi, nb);
#pragma unroll
    for (int u = 0; u < U; u++) {
      float r, rinv, p;
      int m;
      pm::dist((float)(nb[u].in ? nb[u].d2 : a.rc2), r, rinv);
      pm::knot_r(t, r, m, p);
      const float val = pm::cubic(tab_rho_row(ltab, ta.m_lo, pair, m), p);
      rho += nb[u].in ? (double)val : 0.0;
    }
  }
  for (int k0 = h + kTabLanes * kTabPL; k0 < cnt; k0 += kTabLanes * U) {
    int c[U];
    tlist<U>(a, ca.global, ca.local, k0, cnt, c);
    LNb nb[U];
    tgather<U>(a, rec, c, k0, cnt, pi, nb);
#pragma unroll
    for (int u = 0; u < U; u++) {
      float r, rinv, p;
      int m;
      pm::dist((float)(nb[u].in ? nb[u].d2 : a.rc2), r, rinv);
      pm::knot_r(t, r, m, p);
      const float val = pm::cubic(tab_rho_row(ltab, ta.m_lo, pair, m), p);
      rho += nb[u].in ? (double)val : 0.0;
    }
  }
  return rho;
}

struct ForceSums {
  double fx, fy, fz, e;
  double v[9];
};

template <int U, bool ENERGY, bool VIRIAL>
__device__ __forceinline__ void tab_force_batch(const TabArgs& ta, const EamTabDev<float>& t,
                                                const float* __restrict__ pair, const float* ltab, float fpi,
                                                const LNb (&nb)[U], ForceSums& fs)
{
#pragma unroll
  for (int u = 0; u < U; u++) {
    float r, rinv, p, fp_r, phi;
    int m;
    pm::dist((float)(nb[u].in ? nb[u].d2 : ta.a.rc2), r, rinv);
    pm::knot_r(t, r, m, p);
    pm::Q4<float> rc, zc;
    tab_pair_rows(ltab, ta.m_lo, pair, m, rc, zc);
    pm::eam_pair_rows(t, rc, zc, p, rinv, fpi, (float)nb[u].w, fp_r, phi);
    const double fpair = nb[u].in ? (double)fp_r : 0.0;
    const double fe0 = nb[u].x * fpair, fe1 = nb[u].y * fpair, fe2 = nb[u].z * fpair;
    fs.fx += fe0;
    fs.fy += fe1;
    fs.fz += fe2;
    if constexpr (ENERGY) fs.e += nb[u].in ? 0.5 * (double)phi : 0.0;
    if constexpr (VIRIAL) pm::add_virial(fs.v, fe0, fe1, fe2, nb[u].x, nb[u].y, nb[u].z);
  }
}

template <int U, bool ENERGY, bool VIRIAL>
__device__ __forceinline__ void tab_force_entries(const TabArgs& ta, const EamTabDev<float>& t,
                                                  const float* __restrict__ pair, const float* ltab,
                                                  const LRec& rec, const Central& ca, const Pos4& pi, int h,
                                                  int cnt, const uint32_t (&E)[kTabPL / 2], ForceSums& fs)
{
  const BlockArgs& a = ta.a;
  const float fpi = (float)pi.w;
#pragma unroll 1
  for (int u0 = 0; u0 < kTabPL; u0 += U) {
    if (h + kTabLanes * u0 >= cnt) break;
    int c[U];
#pragma unroll
    for (int u = 0; u < U; u++) c[u] = (h + kTabLanes * (u0 + u) < cnt) ? list_entry(E, u0 + u) : ca.local;
    LNb nb[U];
    tgather<U>(a, rec, c, h + kTabLanes * u0, cnt, pi, nb);
    tab_force_batch<U, ENERGY, VIRIAL>(ta, t, pair, ltab, fpi, nb, fs);
  }
  for (int k0 = h + kTabLanes * kTabPL; k0 < cnt; k0 += kTabLanes * U) {
    int c[U];
    tlist<U>(a, ca.global, ca.local, k0, cnt, c);
    LNb nb[U];
    tgather<U>(a, rec, c, k0, cnt, pi, nb);
    tab_force_batch<U, ENERGY, VIRIAL>(ta, t, pair, ltab, fpi, nb, fs);
  }
}

template <bool EMB, bool ENERGY, bool SET_EP>
__global__ void __launch_bounds__(kTabThreads) k_tab_eam_rho(TabArgs ta, EamTabDev<float> t,
                                                             const float* __restrict__ pair,
                                                             double* __restrict__ rho_io, double* __restrict__ pos_w,
                                                             double* __restrict__ ep)
{
  const BlockArgs& a = ta.a;
  const TabLayout l = tab_layout(ta.rows, 4, a.cap, a.g.nh_max);
  // rho value cubics: the first half of each combined row
  for (int i = threadIdx.x; i < ta.rows; i += blockDim.x)
    reinterpret_cast<float4*>(l.tab)[i] = reinterpret_cast<const float4*>(pair)[(int64_t)(ta.m_lo + i) * 2];
  const int64_t b0 = lblock() * ta.per_wg;
  const int64_t b1 = min(b0 + ta.per_wg, a.g.nblocks);
  const int h = tab_h();
  HaloNext hn;
  if (b0 < b1) halo_fetch(a.g, a.cell_off, ta.brick_rows, b0, hn);
  for (int64_t b = b0; b < b1; b++) {
    const BlockShape s = block_shape(a.g, b);
    const int total = halo_commit(s, hn, l.hc, l.total);
    const int units = (hn.rows + kTabLanes - 1) / kTabLanes;
    Central ca;
    const bool have = central_atom(s, l.hc, tab_slot(), ca);
    int cnt = 0;
    uint32_t E[kTabPL / 2];
    if (have) {
      cnt = a.nnb[ca.global];
      list_fetch(a, ta.nmax, ca.global, h, units, E);
    }
    if (total <= a.cap) stage2(s, l.hc, total, a.pos, a.box[s.r].L, l.rec);  // always (cap = measured maximum)
    else __syncthreads();
    if (b + 1 < b1) halo_fetch(a.g, a.cell_off, ta.brick_rows, b + 1, hn);
    if (total <= a.cap) {
      for (int q = tab_slot(); have; q += kTabAtoms) {
        if (q != tab_slot()) {  // atoms past the first pass (bricks of more than kTabAtoms atoms)
          if (!central_atom(s, l.hc, q, ca)) break;
          cnt = a.nnb[ca.global];
          list_fetch(a, ta.nmax, ca.global, h, units, E);
        }
        const Pos4 pi = l.rec.get(ca.local);
        double rho = tab_rho_entries<kTBatch>(ta, t, pair, l.tab, l.rec, ca, pi, h, cnt, E);
        rho = group_sum(rho);
        if (h == 0) {
          rho_io[ca.global] = rho;
          if constexpr (EMB) {
            float F, Fp;
            pm::embed(t, 0, rho, F, Fp);
            pos_w[4 * ca.global + 3] = (double)Fp;
            if constexpr (ENERGY) {
              if constexpr (SET_EP) ep[ca.global] = (double)F;
              else ep[ca.global] += (double)F;
            }
          }
        }
      }
    }
    __syncthreads();  // this brick's LDS is read by every lane before the next is staged
  }
}

template <bool ENERGY, bool VIRIAL, bool FUSED, bool DRIFT>
__global__ void __launch_bounds__(kTabThreads) k_tab_eam_force(TabArgs ta, EamTabDev<float> t,
                                                               const float* __restrict__ pair, pm::Outputs o,
                                                               pm::Drift dr)
{
  static_assert(!DRIFT || (FUSED && !ENERGY && !VIRIAL), "the drift epilogue is the plain MD-loop step");
  const BlockArgs& a = ta.a;
  const TabLayout l = tab_layout(ta.rows, 8, a.cap, a.g.nh_max);
  for (int i = threadIdx.x; i < 2 * ta.rows; i += blockDim.x)
    reinterpret_cast<float4*>(l.tab)[i] = reinterpret_cast<const float4*>(pair)[(int64_t)ta.m_lo * 2 + i];
  const int64_t b0 = lblock() * ta.per_wg;
  const int64_t b1 = min(b0 + ta.per_wg, a.g.nblocks);
  const int h = tab_h();
  HaloNext hn;
  if (b0 < b1) halo_fetch(a.g, a.cell_off, ta.brick_rows, b0, hn);
  for (int64_t b = b0; b < b1; b++) {
    const BlockShape s = block_shape(a.g, b);
    const int total = halo_commit(s, hn, l.hc, l.total);
    const int units = (hn.rows + kTabLanes - 1) / kTabLanes;
    Central ca;
    const bool have = central_atom(s, l.hc, tab_slot(), ca);
    int cnt = 0;
    uint32_t E[kTabPL / 2];
    if (have) {
      cnt = a.nnb[ca.global];
      list_fetch(a, ta.nmax, ca.global, h, units, E);
    }
    if (total <= a.cap) stage2(s, l.hc, total, a.pos, a.box[s.r].L, l.rec);  // always (cap = measured maximum)
    else __syncthreads();
    if (b + 1 < b1) halo_fetch(a.g, a.cell_off, ta.brick_rows, b + 1, hn);
    if (total <= a.cap) {
      for (int q = tab_slot(); have; q += kTabAtoms) {
        if (q != tab_slot()) {
          if (!central_atom(s, l.hc, q, ca)) break;
          cnt = a.nnb[ca.global];
          list_fetch(a, ta.nmax, ca.global, h, units, E);
        }
        const Pos4 pi = l.rec.get(ca.local);
        ForceSums fs{0.0, 0.0, 0.0, 0.0, {0, 0, 0, 0, 0, 0, 0, 0, 0}};
        tab_force_entries<kTBatch, ENERGY, VIRIAL>(ta, t, pair, l.tab, l.rec, ca, pi, h, cnt, E, fs);
        fs.fx = group_sum(fs.fx);
        fs.fy = group_sum(fs.fy);
        fs.fz = group_sum(fs.fz);
        if constexpr (ENERGY) fs.e = group_sum(fs.e);
        if constexpr (VIRIAL) {
#pragma unroll
          for (int k = 0; k < 9; k++) fs.v[k] = group_sum(fs.v[k]);
        }
        if (h == 0) {
          if constexpr (DRIFT) pm::write_outputs_drift(o, dr, ca.global, pi, fs.fx, fs.fy, fs.fz);
          else pm::write_outputs<ENERGY, VIRIAL, FUSED>(o, ca.global, fs.fx, fs.fy, fs.fz, fs.e, fs.v);
        }
      }
    }
    __syncthreads();
  }
}

// Fused next-step drift of the MD loop (ForceEpilogue::drift) for a force launch
// with the plain MD-loop epilogue; false when it does not apply.
bool drift_setup(Ctx& c, const ForceEpilogue& epi, unsigned flags, pm::Drift& dr)
{
  if (!(epi.fused && epi.drift) || (flags & (XS_ENERGY | XS_VIRIAL))) return false;
  c.pos2.reserve(c.npad);
  dr.pos_out = c.pos2.p;
  dr.rrx = c.rrx.p;
  dr.rry = c.rry.p;
  dr.rrz = c.rrz.p;
  dr.dt = epi.drift_dt;
  dr.dt2 = epi.drift_dt * epi.drift_dt * 0.5;
  dr.hdt = epi.drift_dt * 0.5;
  dr.thr2 = epi.drift_thr2;
  dr.over = c.dscalar.p + 5;
  XS_HIP(hipMemsetAsync(c.dscalar.p + 5, 0, sizeof(unsigned long long), c.stream));
  return true;
}

// after the launch: the next step's positions (pos2) become current
void drift_commit(Ctx& c)
{
  c.pos.swap(c.pos2);
  c.rho_ckpt = false;
  c.drift_done = true;
}

// LDS plan of the persistent passes: rows of the table that fit beside the halo.
struct TabPlan {
  bool ok;
  int m_lo, rows;
  size_t bytes;
};

TabPlan tab_plan(const Ctx& c, int row_floats)
{
  TabPlan tp{false, 0, 0, 0};
  if (!c.lds_table_enabled || c.eam.nel != 1 || c.num_cu <= 0 || c.bcap > 2 * kTabThreads) return tp;
  const size_t halo = (size_t)c.bcap * sizeof(Pos4) + (size_t)c.bgeom.nh_max * sizeof(HaloCell) + 4 * sizeof(int);
  if (halo >= c.lds_max) return tp;
  const size_t row_bytes = (size_t)row_floats * sizeof(float);
  const int nr = c.eam.nr;
  int rows = (int)std::min<size_t>((size_t)nr, (c.lds_max - halo) / row_bytes);
  if (c.lds_table_rows > 0) rows = std::min(rows, c.lds_table_rows);
  // short of a quarter of the table, the HBM fallback would carry too many pairs
  if (rows < std::max(1, nr / 4)) return tp;
  tp.ok = true;
  tp.rows = rows;
  tp.m_lo = nr - rows;
  tp.bytes = (size_t)rows * row_bytes + halo;
  return tp;
}

template <class K>
void set_lds_limit(K kernel, size_t bytes)
{
  XS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes));
}

// Persistent LDS-table passes; false when the table does not fit (caller falls back).
bool tab_eam_dispatch(Ctx& c, const BlockArgs& a, bool do_rho, bool do_emb, bool do_force, unsigned flags,
                      const ForceEpilogue& epi)
{
  const TabPlan pr = tab_plan(c, 4), pf = tab_plan(c, 8);
  if ((do_rho && !pr.ok) || (do_force && !pf.ok)) return false;
  if (do_rho) c.last_rho_rows = pr.rows;
  if (do_force) c.last_force_rows = pf.rows;
  const EamTabDev<float> t = make_eam_tab<float>(c);
  const float* pair = eam_pair_table<float>(c);
  const unsigned g = (unsigned)std::max<int64_t>(8, std::min<int64_t>((c.num_cu + 7) / 8 * 8,
                                                                         (a.g.nblocks + 7) / 8 * 8));
  TabArgs ta;
  ta.a = a;
  ta.brick_rows = c.brick_rows.p;
  ta.nmax = c.nmax;
  ta.per_wg = (a.g.nblocks + g - 1) / g;
  const bool energy = flags & XS_ENERGY;
  double* pos_w = reinterpret_cast<double*>(c.pos.p);
  if (do_rho) {
    ta.m_lo = pr.m_lo;
    ta.rows = pr.rows;
    ProfScope prof(c.prof, c.stream, XS_PROF_EAM_RHO);
#define XS_TRHO(EMB, EN, SET)                                                                               \
  do {                                                                                                      \
    set_lds_limit(k_tab_eam_rho<EMB, EN, SET>, pr.bytes);                                                   \
    k_tab_eam_rho<EMB, EN, SET><<<g, kTabThreads, pr.bytes, c.stream>>>(ta, t, pair, c.rho.p, pos_w, c.ep.p); \
  } while (0)
    if (!do_emb) XS_TRHO(false, false, false);
    else if (energy && epi.fused) XS_TRHO(true, true, true);
    else if (energy) XS_TRHO(true, true, false);
    else XS_TRHO(true, false, false);
#undef XS_TRHO
    XS_HIP(hipGetLastError());
  }
  if (do_force) {
    ta.m_lo = pf.m_lo;
    ta.rows = pf.rows;
    const pm::Outputs o = make_outputs(c, epi);
    const bool vir = flags & XS_VIRIAL;
    pm::Drift dr{};
    const bool drift = drift_setup(c, epi, flags, dr);
    ProfScope prof(c.prof, c.stream, XS_PROF_EAM_FORCE);
#define XS_TF(EN, VI, FU, DR)                                                                    \
  do {                                                                                           \
    set_lds_limit(k_tab_eam_force<EN, VI, FU, DR>, pf.bytes);                                    \
    k_tab_eam_force<EN, VI, FU, DR><<<g, kTabThreads, pf.bytes, c.stream>>>(ta, t, pair, o, dr); \
  } while (0)
    if (drift) {
      XS_TF(false, false, true, true);
    } else if (epi.fused) {
      if (energy && vir) XS_TF(true, true, true, false);
      else if (energy) XS_TF(true, false, true, false);
      else if (vir) XS_TF(false, true, true, false);
      else XS_TF(false, false, true, false);
    } else {
      if (energy && vir) XS_TF(true, true, false, false);
      else if (energy) XS_TF(true, false, false, false);
      else if (vir) XS_TF(false, true, false, false);
      else XS_TF(false, false, false, false);
    }
#undef XS_TF
    XS_HIP(hipGetLastError());
    if (drift) drift_commit(c);
  }
  return true;
}

template <class Real, bool MULTI>
void block_eam_dispatch(Ctx& c, double rcut, bool do_rho, bool do_emb, bool do_force, unsigned flags,
                        const ForceEpilogue& epi)
{
  const EamTabDev<Real> t = make_eam_tab<Real>(c);
  const Real* pair = eam_pair_table<Real>(c);
  BlockArgs a;
  a.g = c.bgeom;
  a.cell_off = c.cell_off.p;
  a.pos = c.pos.p;
  a.type = c.type.p;
  a.box = c.box_d.p;
  a.nbr16 = c.nbr16.p;
  a.nnb = c.nnb.p;
  a.npad = c.npad;
  a.cap = c.bcap;
  a.rc2 = rcut * rcut;
  if constexpr (std::is_same_v<Real, float> && !MULTI) {
    if (tab_eam_dispatch(c, a, do_rho, do_emb, do_force, flags, epi)) return;
  }
  if (do_rho) c.last_rho_rows = 0;
  if (do_force) c.last_force_rows = 0;
  const size_t lds = block_lds_bytes(c, MULTI);
  const unsigned g = block_grid(c);
  const bool energy = flags & XS_ENERGY;
  double* pos_w = reinterpret_cast<double*>(c.pos.p);
  if (do_rho) {
    ProfScope prof(c.prof, c.stream, XS_PROF_EAM_RHO);
#define XS_BRHO(EMB, EN, SET) \
  k_block_eam_rho<Real, MULTI, EMB, EN, SET><<<g, kPassThreads, lds, c.stream>>>(a, t, pair, c.rho.p, pos_w, c.ep.p)
    if (!do_emb) XS_BRHO(false, false, false);
    else if (energy && epi.fused) XS_BRHO(true, true, true);
    else if (energy) XS_BRHO(true, true, false);
    else XS_BRHO(true, false, false);
#undef XS_BRHO
    XS_HIP(hipGetLastError());
  }
  if (do_force) {
    const pm::Outputs o = make_outputs(c, epi);
    const bool vir = flags & XS_VIRIAL;
    pm::Drift dr{};
    const bool drift = drift_setup(c, epi, flags, dr);
    ProfScope prof(c.prof, c.stream, XS_PROF_EAM_FORCE);
#define XS_BF(EN, VI, FU, DR) \
  k_block_eam_force<Real, MULTI, EN, VI, FU, DR><<<g, kPassThreads, lds, c.stream>>>(a, t, pair, o, dr)
    if (drift) {
      XS_BF(false, false, true, true);
    } else if (epi.fused) {
      if (energy && vir) XS_BF(true, true, true, false);
      else if (energy) XS_BF(true, false, true, false);
      else if (vir) XS_BF(false, true, true, false);
      else XS_BF(false, false, true, false);
    } else {
      if (energy && vir) XS_BF(true, true, false, false);
      else if (energy) XS_BF(true, false, false, false);
      else if (vir) XS_BF(false, true, false, false);
      else XS_BF(false, false, false, false);
    }
#undef XS_BF
    XS_HIP(hipGetLastError());
    if (drift) drift_commit(c);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// cell offsets: off[k] = first atom with cell key >= k, off[ncell] = n
void cell_offsets(Ctx& c)
{
  c.cell_off.reserve(c.ncell_tot + 1);
  k_cell_offsets<<<grid_for(c.n + 1), kBlock, 0, c.stream>>>(c.n, c.ncell_tot, c.keys2.p, c.cell_off.p);
  XS_HIP(hipGetLastError());
}

void block_setup(Ctx& c)
{
  // brick of about kBlockAtoms atoms, as cubic as the grid allows
  const double per_cell = (double)c.n / (double)c.ncell_tot;
  const int want = std::max(1, (int)std::lround(c.block_atoms / std::max(per_cell, 0.5)));
  BlockGeom& g = c.bgeom;
  g.nx = c.nc[0];
  g.ny = c.nc[1];
  g.nz = c.nc[2];
  g.px = c.periodic[0];
  g.py = c.periodic[1];
  g.pz = c.periodic[2];
  g.ncell_rep = c.ncell_rep;
  int bz = std::max(1, std::min(g.nz, (int)std::floor(std::cbrt((double)want))));
  int by = std::max(1, std::min(g.ny, (int)std::floor(std::sqrt((double)want / bz))));
  int bx = std::max(1, std::min(g.nx, (int)std::lround((double)want / (by * bz))));
  while ((bx + 2) * (by + 2) * (bz + 2) > kMaxHaloCells && bx > 1) bx--;
  g.bx = bx;
  g.by = by;
  g.bz = bz;
  g.nbx = (g.nx + bx - 1) / bx;
  g.nby = (g.ny + by - 1) / by;
  g.nbz = (g.nz + bz - 1) / bz;
  g.nblocks = (int64_t)g.nbx * g.nby * g.nbz * c.nrep;
  g.nh_max = (bx + 2) * (by + 2) * (bz + 2);

  XS_HIP(hipMemsetAsync(c.dscalar.p + 2, 0, 2 * sizeof(unsigned long long), c.stream));
  k_block_halo_max<<<grid_for(g.nblocks), kBlock, 0, c.stream>>>(g, c.cell_off.p, c.dscalar.p + 2);
  XS_HIP(hipGetLastError());
  XS_HIP(hipMemcpyAsync(c.hscalar + 2, c.dscalar.p + 2, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        c.stream));
  XS_HIP(hipStreamSynchronize(c.stream));
  c.bcap = (int)((c.hscalar[2] + 63) / 64 * 64);
  c.bcentral = (int)c.hscalar[3];
}

size_t block_lds_bytes(const Ctx& c, bool with_type)
{
  return (size_t)c.bcap * sizeof(Pos4) + (size_t)c.bgeom.nh_max * sizeof(HaloCell) + 4 * sizeof(int) +
         (with_type ? (size_t)c.bcap : 0);
}

bool block_feasible(const Ctx& c)
{
  return c.bcap > 0 && c.bcap < 65536 && block_lds_bytes(c, true) <= kBlockLdsLimit;
}

unsigned block_grid(const Ctx& c)
{
  return (unsigned)((c.bgeom.nblocks + 7) / 8 * 8);
}

void block_build(Ctx& c, double rl2)
{
  const size_t lds = block_lds_bytes(c, false);
  for (int attempt = 0; attempt < 2; attempt++) {
    c.nbr16.reserve((size_t)c.nmax * c.npad);
    c.nnb.reserve(c.npad);
    c.brick_rows.reserve(c.bgeom.nblocks);
    XS_HIP(hipMemsetAsync(c.dscalar.p, 0, 2 * sizeof(unsigned long long), c.stream));
    k_block_build<<<block_grid(c), kPassThreads, lds, c.stream>>>(c.bgeom, c.cell_off.p, c.pos.p, c.box_d.p, rl2,
                                                            c.nbr16.p, c.nnb.p, c.nmax, c.npad, c.bcap,
                                                            c.bgeom.nh_max, c.dscalar.p + 4, c.brick_rows.p,
                                                            c.dscalar.p);
    XS_HIP(hipGetLastError());
    XS_HIP(hipMemcpyAsync(c.hscalar, c.dscalar.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          c.stream));
    XS_HIP(hipStreamSynchronize(c.stream));
    if (c.hscalar[1] != 0) throw Error(XS_E_STATE, "block halo exceeds its measured capacity");
    const unsigned long long need = c.hscalar[0];
    if (need == 0) return;
    if (attempt == 1) throw Error(XS_E_STATE, "neighbour list overflow after resize");
    c.nmax = (int)((need + 7) / 8 * 8 + 8);
  }
}

void block_to_global(Ctx& c)
{
  c.nbr.reserve((size_t)c.nmax * c.npad);
  const size_t lds = block_lds_bytes(c, false) - (size_t)c.bcap * sizeof(Pos4);
  k_block_to_global<<<block_grid(c), kBlock, lds, c.stream>>>(c.bgeom, c.cell_off.p, c.nbr16.p, c.nnb.p, c.nbr.p,
                                                              c.npad, c.bgeom.nh_max);
  XS_HIP(hipGetLastError());
}

}  // namespace xsmd

namespace xsmd {

// EAM on the block list.  The embedding step alone (eam_rho2emb after a separate
// eam_rho, as in the split-pass deck) reads no list and runs the per-atom kernel of
// the global path; its arithmetic is the same as the fused block pass.
void launch_eam_block(Ctx& c, double rcut, bool do_rho, bool do_emb, bool do_force, unsigned flags,
                      const ForceEpilogue& epi)
{
  const bool mixed = (flags & XS_MIXED) != 0;
  const bool multi = c.eam.nel > 1;
  if (do_emb && !do_rho) {
    const ListKind keep = c.list_kind;
    c.list_kind = ListKind::Global;
    try {
      launch_eam(c, rcut, false, true, false, flags, epi);
    } catch (...) {
      c.list_kind = keep;
      throw;
    }
    c.list_kind = keep;
    if (!do_force) return;
    do_emb = false;
  }
  if (mixed) {
    if (multi) block_eam_dispatch<float, true>(c, rcut, do_rho, do_emb, do_force, flags, epi);
    else block_eam_dispatch<float, false>(c, rcut, do_rho, do_emb, do_force, flags, epi);
  } else {
    if (multi) block_eam_dispatch<double, true>(c, rcut, do_rho, do_emb, do_force, flags, epi);
    else block_eam_dispatch<double, false>(c, rcut, do_rho, do_emb, do_force, flags, epi);
  }
}

void ensure_global_list(Ctx& c)
{
  if (c.list_kind == ListKind::Global || c.global_from_block) return;
  block_to_global(c);
  c.global_from_block = true;
}

}  // namespace xsmd
