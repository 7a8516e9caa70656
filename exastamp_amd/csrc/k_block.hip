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

// ---------------------------------------------------------------------------
// list build
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_block_build(BlockGeom g, const int* __restrict__ cell_off,
                                                     const Pos4* __restrict__ pos, const BoxDev* __restrict__ box,
                                                     double rl2, uint16_t* __restrict__ nbr16, int* __restrict__ nnb,
                                                     int nmax, int64_t npad, int cap, int nh_max,
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
  stage(s, l.hc, total, pos, box[s.r].L, nullptr, l.rec, nullptr);
  for (int q = threadIdx.x;; q += blockDim.x) {
    Central a;
    if (!central_atom(s, l.hc, q, a)) break;
    const Pos4 pi = l.rec.get(a.local);
    const int ix = a.h % s.hx, iy = (a.h / s.hx) % s.hy, iz = a.h / (s.hx * s.hy);
    int cnt = 0;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const HaloCell e = l.hc[((iz + dz) * s.hy + (iy + dy)) * s.hx + (ix + dx)];
          for (int c = e.lstart; c < e.lstart + e.count; c++) {
            if (c == a.local) continue;
            const Pos4 pj = l.rec.get(c);
            if (xsmd_d2(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z) < rl2) {
              if (cnt < nmax) nbr16[(int64_t)cnt * npad + a.global] = (uint16_t)c;
              cnt++;
            }
          }
        }
    nnb[a.global] = cnt;
    if (cnt > nmax) atomicMax(flags, (unsigned long long)cnt);
  }
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
constexpr int kPassThreads = 512;
constexpr int kPassAtoms = kPassThreads / 2;
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

// sum of v over the two halves of the wave (lanes l and l ^ 32)
__device__ __forceinline__ double half_sum(double v) { return v + __shfl_xor(v, 32); }

// this lane's atom slot and list parity
__device__ __forceinline__ int pass_slot() { return (threadIdx.x >> 6) * 32 + (threadIdx.x & 31); }
__device__ __forceinline__ int pass_half() { return (threadIdx.x >> 5) & 1; }

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

template <class Real, bool MULTI, bool ENERGY, bool VIRIAL, bool FUSED>
__global__ void __launch_bounds__(kPassThreads, kPassWaves<VIRIAL>) k_block_eam_force(BlockArgs a, EamTabDev<Real> t,
                                                                  const Real* __restrict__ pair, pm::Outputs o)
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
    if (half == 0) pm::write_outputs<ENERGY, VIRIAL, FUSED>(o, ca.global, fx, fy, fz, e, v);
  }
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
    ProfScope prof(c.prof, c.stream, XS_PROF_EAM_FORCE);
#define XS_BF(EN, VI, FU) \
  k_block_eam_force<Real, MULTI, EN, VI, FU><<<g, kPassThreads, lds, c.stream>>>(a, t, pair, o)
    if (epi.fused) {
      if (energy && vir) XS_BF(true, true, true);
      else if (energy) XS_BF(true, false, true);
      else if (vir) XS_BF(false, true, true);
      else XS_BF(false, false, true);
    } else {
      if (energy && vir) XS_BF(true, true, false);
      else if (energy) XS_BF(true, false, false);
      else if (vir) XS_BF(false, true, false);
      else XS_BF(false, false, false);
    }
#undef XS_BF
    XS_HIP(hipGetLastError());
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
    XS_HIP(hipMemsetAsync(c.dscalar.p, 0, 2 * sizeof(unsigned long long), c.stream));
    k_block_build<<<block_grid(c), kBlock, lds, c.stream>>>(c.bgeom, c.cell_off.p, c.pos.p, c.box_d.p, rl2,
                                                            c.nbr16.p, c.nnb.p, c.nmax, c.npad, c.bcap,
                                                            c.bgeom.nh_max, c.dscalar.p);
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
