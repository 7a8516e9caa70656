// k_block.hip -- cell-block (LDS-staged) neighbour list and EAM passes.
//
// The cell grid is cut into bricks of bx*by*bz cells.  One workgroup owns one
// brick: it stages the brick's halo -- the brick plus one cell on every side,
// periodic images applied once at staging -- into LDS as arrays of doubles
// (x, y, z, and w = F'(rho) in the force pass), then two lanes walk the list of
// each central atom.  List entries are 16-bit halo-local slots, so a pair costs
// three 8-byte LDS reads instead of a global gather, and the list is half the
// size of the global one.
//
// The halo of a brick is an (ex+2)x(ey+2)x(ez+2) box of cells enumerated z-major,
// x fastest.  Atoms of one cell are contiguous in HBM (cell-sorted order) and are
// staged contiguously, cells in enumeration order, so ascending slot order equals
// the visit order of the global build.  The build finds the same pairs as the
// global build; k_list_sort then orders each list by distance, and
// k_block_to_global restores slot order, so a converted block list equals the
// global list entry for entry.
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

// Staged halo records in LDS: {x, y, z} of slot c as three doubles at 24 c (one
// address and immediate offsets per neighbour) and, for the passes that read it, w[c]
// in an array of its own.  Random 8-byte reads at a 24-byte stride spread over the
// same 32 bank pairs as reads of 8-byte arrays (6 c mod 64 covers the even banks),
// while 16-byte reads conflict more: the passes measured faster with 8-byte reads
// per component than with the two 16-byte reads of {x, y}, {z, w} records.  Passes
// without w alias w to the x of the records (never read).
struct LRec {
  double* xyz;
  double* w;
  bool has_w;
  __device__ __forceinline__ Pos4 get(int c) const
  {
    const double* r = xyz + 3 * c;
    return {r[0], r[1], r[2], w[has_w ? c : 3 * c]};
  }
  __device__ __forceinline__ void put(int c, const Pos4& p) const
  {
    double* r = xyz + 3 * c;
    r[0] = p.x;
    r[1] = p.y;
    r[2] = p.z;
    if (has_w) w[c] = p.w;
  }
};

// records at p for cap slots plus the far record at slot cap (see stage); returns
// the bytes they take
__device__ __forceinline__ size_t lrec_at(unsigned char* p, int cap, bool with_w, LRec& r)
{
  r.xyz = reinterpret_cast<double*>(p);
  r.w = with_w ? r.xyz + 3 * (size_t)(cap + 1) : r.xyz;
  r.has_w = with_w;
  return (size_t)(cap + 1) * sizeof(double) * (with_w ? 4 : 3);
}

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

// Stage halo atoms: f(c, p, j) for slot c, position p (+ image shift; under a
// non-identity xform mapped to physical space, so that the passes' displacements
// are physical) and, with WITH_W, w (else 0: the 24 bytes of x, y, z are read, not
// the whole record), global index j; ends with __syncthreads().  slot_cell (total <= its capacity)
// first receives the halo cell of every slot, so that a thread finds the cell of
// a slot with one LDS read instead of a binary search over the halo table.
// (Issuing several slots' position loads per thread before using them measured
// slower: the force pass lost 4 %.)
template <bool WITH_W, class F>
__device__ void stage_each(const BlockShape& s, const HaloCell* hc, int total, const Pos4* __restrict__ pos,
                           const double* L, const XformDev* xf, uint16_t* slot_cell, F f)
{
  const int nh = s.hx * s.hy * s.hz;
  for (int h = threadIdx.x; h < nh; h += blockDim.x) {
    const HaloCell e = hc[h];
    for (int k = 0; k < e.count; k++) slot_cell[e.lstart + k] = (uint16_t)h;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < total; c += blockDim.x) {
    const HaloCell e = hc[slot_cell[c]];
    const int j = e.gstart + (c - e.lstart);
    Pos4 p;
    if constexpr (WITH_W) {
      p = load_pos(pos, j);
    } else {
      const double* q = reinterpret_cast<const double*>(pos + j);
      const double2 a = *reinterpret_cast<const double2*>(q);
      p = {a.x, a.y, q[2], 0.0};
    }
    if (e.img) {
      const uint32_t code = (uint32_t)e.img << XSMD_NBR_INDEX_BITS;
      p.x = xsmd_image(p.x, xsmd_nbr_img_x(code), L[0]);
      p.y = xsmd_image(p.y, xsmd_nbr_img_y(code), L[1]);
      p.z = xsmd_image(p.z, xsmd_nbr_img_z(code), L[2]);
    }
    if (xf) xsmd_xform_apply(xf->m, &p.x, &p.y, &p.z);
    f(c, p, j);
  }
  __syncthreads();
}

// Stage halo atoms into the records (and types when requested), and the far record
// at slot cap: list entries past an atom's count point there (lslots), so their
// displacement is finite and longer than any cutoff, and the passes mask them by
// d2 <= rc2 alone.  It lies at least 10^4 A from every staged position (which are
// within the box and its images) and close enough that its knot index fits an int.
__device__ void stage(const BlockShape& s, const HaloCell* hc, int total, const Pos4* __restrict__ pos,
                      const BoxDev& box, const XformDev* xf, const uint8_t* __restrict__ type, const LRec& rec,
                      uint8_t* ltype, uint16_t* slot_cell, int cap)
{
  const double* L = box.L;
  if (threadIdx.x == 0) {  // visible after stage_each's barriers
    const double far = 4.0 * (fabs(box.lo[0]) + fabs(box.lo[1]) + fabs(box.lo[2]) + L[0] + L[1] + L[2]) + 1e4;
    Pos4 f = {far, far, far, 0.0};
    if (xf) {
      xsmd_xform_apply(xf->m, &f.x, &f.y, &f.z);
      const double g = 1e4 + 4.0 * fabs(f.x - far) + 4.0 * fabs(f.y - far) + 4.0 * fabs(f.z - far);
      f = {f.x + g, f.y + g, f.z + g, 0.0};  // still far in physical space
    }
    double* r = rec.xyz + 3 * cap;
    r[0] = f.x;
    r[1] = f.y;
    r[2] = f.z;
    if (rec.has_w) rec.w[cap] = 0.0;
    if (ltype) ltype[cap] = 0;
  }
  const auto put = [&](int c, const Pos4& p, int j) {
    rec.put(c, p);
    if (ltype) ltype[c] = type[j];
  };
  if (rec.has_w) stage_each<true>(s, hc, total, pos, L, xf, slot_cell, put);
  else stage_each<false>(s, hc, total, pos, L, xf, slot_cell, put);
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

// Local slots of the central atoms in q order (the order of central_atom):
// central_local[q] for q < count, count in *count_out.  Wave 0 takes one central
// row (iz, iy) per lane, scans the row lengths with shuffles and writes the rows'
// slots; entries beyond ccap are not written (the caller checks count <= ccap).
// The caller's next __syncthreads publishes the table.
__device__ void central_table(const BlockShape& s, const HaloCell* hc, uint16_t* central_local, int ccap,
                              int* count_out)
{
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  const int nrows = s.ez * s.ey;
  int carry = 0;
  for (int base = 0; base < nrows; base += 64) {
    const int r = base + lane;
    int len = 0, l0 = 0;
    if (r < nrows) {
      const int iz = 1 + r / s.ey, iy = 1 + r % s.ey;
      const int h0 = (iz * s.hy + iy) * s.hx + 1;
      const int h1 = h0 + s.ex - 1;
      l0 = hc[h0].lstart;
      len = hc[h1].lstart + hc[h1].count - l0;
    }
    int v = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(v, off);
      if (lane >= off) v += t;
    }
    const int q0 = carry + v - len;
    for (int k = 0; k < len && q0 + k < ccap; k++) central_local[q0 + k] = (uint16_t)(l0 + k);
    carry += __shfl(v, 63);
  }
  if (lane == 0) *count_out = carry;
}

// central atom q from the tables (slot_cell and central_local filled)
__device__ __forceinline__ bool central_at(const HaloCell* hc, const uint16_t* slot_cell,
                                           const uint16_t* central_local, int count, int q, Central& out)
{
  if (q >= count) return false;
  const int local = central_local[q];
  const int h = slot_cell[local];
  out.h = h;
  out.local = local;
  out.global = (int64_t)hc[h].gstart + (local - hc[h].lstart);
  return true;
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
  uint16_t* slot_cell;      // halo cell of each slot (staging)
  uint16_t* central_local;  // local slots of the central atoms
};

// LDS of the block kernels: records, halo table, counters (total[0] halo atoms,
// total[2] central atoms), slot cells, central slots, per-record types
// (block_lds_bytes on the host adds up the same sizes)
__device__ __forceinline__ LdsLayout lds_layout(int cap, int ccap, int nh_max, bool with_type, bool with_w)
{
  extern __shared__ __align__(32) unsigned char smem[];
  LdsLayout l;
  const size_t rb = lrec_at(smem, cap, with_w, l.rec);
  l.hc = reinterpret_cast<HaloCell*>(smem + rb);
  l.total = reinterpret_cast<int*>(l.hc + nh_max);
  l.slot_cell = reinterpret_cast<uint16_t*>(l.total + 4);
  l.central_local = l.slot_cell + cap;
  l.ltype = with_type ? reinterpret_cast<uint8_t*>(l.central_local + ccap) : nullptr;
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
// number of central atoms of a brick (the row walk of central_atom)
__device__ __forceinline__ int central_count(const BlockShape& s, const HaloCell* hc)
{
  int n = 0;
  for (int iz = 1; iz <= s.ez; iz++)
    for (int iy = 1; iy <= s.ey; iy++) {
      const int h0 = (iz * s.hy + iy) * s.hx + 1;
      const int h1 = h0 + s.ex - 1;
      n += hc[h1].lstart + hc[h1].count - hc[h0].lstart;
    }
  return n;
}

// Destination lane of this lane's key in one stable split pass: keys with a 0 bit
// first (in lane order), then those with a 1.  zeros = ballot of the 0 keys; mbcnt
// counts the 0 keys below this lane, so the 1 keys below it are lane minus that.
// Both candidates are formed and one is selected, so the pass has no branch.
__device__ __forceinline__ int split_dest(bool one, unsigned long long zeros, int lane)
{
  const int zb = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(zeros >> 32),
                                                __builtin_amdgcn_mbcnt_lo((uint32_t)zeros, 0u));
  const int nz = __popcll(zeros);
  const int d1 = nz + lane - zb;
  return one ? d1 : zb;
}

// Stable sort of one 32-bit key per lane over the 64 lanes of a wave by key bits
// [lo, lo + nbits): least-significant-digit radix sort, one bit per pass (ballot,
// prefix counts, forward permute).  Keys whose sort bits are all ones go last.
__device__ __forceinline__ uint32_t wave_radix_sort(uint32_t key, int lo, int nbits)
{
  const int lane = threadIdx.x & 63;
  for (int bit = lo; bit < lo + nbits; bit++) {
    const bool one = (key >> bit) & 1u;
    const int dest = split_dest(one, __ballot(!one), lane);
    key = (uint32_t)__builtin_amdgcn_ds_permute(dest << 2, (int)key);
  }
  return key;
}

// two independent sorts interleaved (the permutes of one wait on the LDS crossbar
// while the other's ballots issue)
__device__ __forceinline__ void wave_radix_sort2(uint32_t& k0, uint32_t& k1, int lo, int nbits)
{
  const int lane = threadIdx.x & 63;
  for (int bit = lo; bit < lo + nbits; bit++) {
    const bool one0 = (k0 >> bit) & 1u, one1 = (k1 >> bit) & 1u;
    const int d0 = split_dest(one0, __ballot(!one0), lane);
    const int d1 = split_dest(one1, __ballot(!one1), lane);
    k0 = (uint32_t)__builtin_amdgcn_ds_permute(d0 << 2, (int)k0);
    k1 = (uint32_t)__builtin_amdgcn_ds_permute(d1 << 2, (int)k1);
  }
}

// Lists of at most 64 entries are kept sorted by distance (see k_list_sort).
constexpr int kSortRows = 64;

// Block-list storage.  The entries of atom i come in groups of 8: group q holds
// entries 8q .. 8q + 7 in the 16 bytes at nbr16 + (q * npad + i) * 8, the even
// entries first (8q, 8q + 2, 8q + 4, 8q + 6, then 8q + 1, ..., 8q + 7).  The two
// lanes of an atom walk its even and its odd entries, so each lane reads four of
// its entries with one 8-byte load and the 32 atoms of a wave read 512 contiguous
// bytes.  The list capacity (nmax) is a multiple of 8.
__device__ __forceinline__ int64_t bl_index(int64_t npad, int64_t i, int k)
{
  return (((int64_t)(k >> 3) * npad + i) << 3) + ((k & 1) << 2) + ((k & 7) >> 1);
}

// Block list: stencil order dz, dy, dx, in-cell order ascending (the order of the
// global build, ascending halo slot).  One lane per central atom walks the
// candidates itself, two at a time (their LDS reads overlap), and packs its hits
// into the current 8-entry group in registers; a full group goes out as one
// 16-byte store.  A neighbour cell whose box (widened by a margin that covers the
// rounding of cell assignment; unbounded on the outer side of a non-periodic edge
// cell, which also holds the atoms clamped into it) lies beyond rl is trimmed from
// the walk (its atoms would all fail the exact test), unless some atom sits outside
// its cell's box (nocull, set by the cell-key kernel).  Entries past nmax are counted, not
// stored (the host grows nmax and builds again).
constexpr int kBuildThreads = 256;

// The current 8-entry group of a lane is a 128-bit FIFO of 16-bit slots in w[0..3]:
// a new entry enters at the top and the others move down one place (a funnel
// shift per word), so after eight entries w holds them in list order, two per
// word, the first in the low half of w[0].
__device__ __forceinline__ void fifo_push(uint32_t (&w)[4], uint32_t c)
{
  w[0] = __builtin_amdgcn_alignbit(w[1], w[0], 16);
  w[1] = __builtin_amdgcn_alignbit(w[2], w[1], 16);
  w[2] = __builtin_amdgcn_alignbit(w[3], w[2], 16);
  w[3] = __builtin_amdgcn_alignbit(c, w[3], 16);
}

// list order (0, 1), (2, 3), (4, 5), (6, 7) -> the bl_index order (0, 2), (4, 6),
// (1, 3), (5, 7)
__device__ __forceinline__ uint4 group_interleave(const uint32_t (&o)[4])
{
  return make_uint4((o[0] & 0xffffu) | (o[1] << 16), (o[2] & 0xffffu) | (o[3] << 16),
                    (o[0] >> 16) | (o[1] & 0xffff0000u), (o[2] >> 16) | (o[3] & 0xffff0000u));
}

// LDS of the list build: cap + 1 position records of 24 bytes, 16-byte aligned
__host__ __device__ inline size_t build_rec_bytes(int cap) { return ((size_t)(cap + 1) * 24 + 15) / 16 * 16; }

// 8-bit distance keys of a group, list order (byte j of b[0..1] = entry j) -> the
// bl_index order (0, 2, 4, 6 | 1, 3, 5, 7)
__device__ __forceinline__ uint2 bins_interleave(const uint32_t (&b)[2])
{
  const uint32_t e0 = b[0] & 0xffu, e1 = (b[0] >> 8) & 0xffu, e2 = (b[0] >> 16) & 0xffu, e3 = b[0] >> 24;
  const uint32_t e4 = b[1] & 0xffu, e5 = (b[1] >> 8) & 0xffu, e6 = (b[1] >> 16) & 0xffu, e7 = b[1] >> 24;
  return make_uint2(e0 | (e2 << 8) | (e4 << 16) | (e6 << 24), e1 | (e3 << 8) | (e5 << 16) | (e7 << 24));
}

// BINS: each entry also gets its distance key for the sort, min(2^nbits - 1,
// floor(d2 * bin_scale)) with bin_scale = 2^nbits / rl^2 (bins uniform in d^2), in
// bins8, grouped like the list.
// XF: a non-identity domain xform.  The halo stays in grid space and each candidate's
// displacement is mapped by xf->m before the test, in the operation order of the
// global build (so the lists stay bit-identical); grid-space cell boxes do not
// bound physical distances, so no cell is culled.
template <bool BINS, bool XF>
__global__ void __launch_bounds__(kBuildThreads) k_block_build(BlockGeom g, const int* __restrict__ cell_off,
                                                               const Pos4* __restrict__ pos,
                                                               const BoxDev* __restrict__ box, double rl2,
                                                               uint16_t* __restrict__ nbr16, int* __restrict__ nnb,
                                                               int nmax, int64_t npad, int cap, int ccap, int nh_max,
                                                               const unsigned long long* __restrict__ nocull,
                                                               int* __restrict__ brick_rows,
                                                               unsigned long long* __restrict__ flags,
                                                               uint8_t* __restrict__ bins8, float bin_scale,
                                                               int bin_max, const XformDev* __restrict__ xf)
{
  const int64_t b = lblock();
  if (b >= g.nblocks) return;
  // LDS: cap + 1 position records {x, y, z} (24 bytes; the extra one pads the
  // second read of a candidate pair), then the halo table
  extern __shared__ __align__(32) unsigned char smem[];
  double3* P = reinterpret_cast<double3*>(smem);
  struct {
    HaloCell* hc;
    int* total;
    uint16_t* slot_cell;
    uint16_t* central_local;
  } l;
  l.hc = reinterpret_cast<HaloCell*>(smem + build_rec_bytes(cap));
  l.total = reinterpret_cast<int*>(l.hc + nh_max);
  l.slot_cell = reinterpret_cast<uint16_t*>(l.total + 4);
  l.central_local = l.slot_cell + cap;
  const BlockShape s = block_shape(g, b);
  const int total = build_halo(g, s, cell_off, l.hc, l.total);
  if (total > cap) {  // cannot happen when cap = measured maximum; never write past LDS
    if (threadIdx.x == 0) atomicMax(flags + 1, (unsigned long long)total);
    return;
  }
  const BoxDev bx = box[s.r];
  const bool cull = !XF && *nocull == 0ull;
  if (threadIdx.x == 0) l.total[1] = 0;  // longest list of the brick (made visible by stage's barrier)
  central_table(s, l.hc, l.central_local, ccap, l.total + 2);
  stage_each<false>(s, l.hc, total, pos, bx.L, nullptr, l.slot_cell,
                    [&](int c, const Pos4& p, int) { P[c] = double3{p.x, p.y, p.z}; });
  const int ncentral = l.total[2];
  if (ncentral > ccap) {  // cannot happen when ccap = measured maximum
    if (threadIdx.x == 0) atomicMax(flags + 1, (unsigned long long)ncentral);
    return;
  }
  for (int q = threadIdx.x;; q += blockDim.x) {
    Central a;
    if (!central_at(l.hc, l.slot_cell, l.central_local, ncentral, q, a)) break;
    const double xi = P[a.local].x, yi = P[a.local].y, zi = P[a.local].z;
    const Pos4 pi = {xi, yi, zi, 0.0};
    const int ix = a.h % s.hx, iy = (a.h / s.hx) % s.hy, iz = a.h / (s.hx * s.hy);
    int cnt = 0;
    uint32_t w[4] = {0u, 0u, 0u, 0u};
    // A hit at slot c goes into the group FIFO w; a full group moves to r and is
    // stored (when it fits in nmax) after the current stencil row, so that the
    // wave issues one store per row rather than one per candidate pair.
    // gp is the atom's next group (bl_index(npad, i, 8 q)); gq its index q.
    uint32_t r[4] = {0u, 0u, 0u, 0u};
    uint32_t bw[2] = {0u, 0u}, br[2] = {0u, 0u};  // distance keys of w and r (BINS)
    bool ready = false;
    uint16_t* gp = nbr16 + bl_index(npad, a.global, 0);
    uint8_t* bp = BINS ? bins8 + bl_index(npad, a.global, 0) : nullptr;
    const int64_t gstep = npad * 8;
    const int gmax = nmax >> 3;  // groups that fit
    int gq = 0;
    auto flush = [&]() {
      if (gq < gmax) {
        *reinterpret_cast<uint4*>(gp) = group_interleave(r);
        if constexpr (BINS) *reinterpret_cast<uint2*>(bp) = bins_interleave(br);
      }
      gp += gstep;
      if constexpr (BINS) bp += gstep;
      gq++;
      ready = false;
    };
    auto hit = [&](int c, double d2) {
      fifo_push(w, (uint32_t)c);
      if constexpr (BINS) {
        const uint32_t bin = (uint32_t)min(bin_max, (int)((float)d2 * bin_scale));
        bw[0] = __builtin_amdgcn_alignbit(bw[1], bw[0], 8);
        bw[1] = __builtin_amdgcn_alignbit(bin, bw[1], 8);
      }
      cnt++;
      if ((cnt & 7) == 0) {
        if (ready) flush();  // a second full group within one row
        r[0] = w[0];
        r[1] = w[1];
        r[2] = w[2];
        r[3] = w[3];
        br[0] = bw[0];
        br[1] = bw[1];
        ready = true;
      }
    };
    // The three cells dx = -1, 0, 1 of one (dy, dz) row of the stencil hold
    // consecutive halo slots, so a row is walked as one slot range: from the first
    // to the last of its cells that the box test keeps (a culled cell between them
    // is walked, and its atoms fail the exact test).  Nine walks of about three
    // cells each, rather than 27 of one, even out the trip counts of the wave's
    // lanes and halve the odd-count tails of the pairwise walk.
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++) {
        const int h0 = ((iz + dz) * s.hy + (iy + dy)) * s.hx + ix - 1;
        int beg = l.hc[h0].lstart;
        int end = l.hc[h0 + 2].lstart + l.hc[h0 + 2].count;
        if (cull) {
          const int gx = s.cx0 - 2 + ix, gy = s.cy0 - 1 + iy + dy, gz = s.cz0 - 1 + iz + dz;
          const double uy = xsmd_cell_gap(pi.y, bx.lo[1], bx.L[1], g.ny, gy, g.py);
          const double uz = xsmd_cell_gap(pi.z, bx.lo[2], bx.L[2], g.nz, gz, g.pz);
          bool keep[3];
#pragma unroll
          for (int k = 0; k < 3; k++)
            keep[k] = xsmd_d2(xsmd_cell_gap(pi.x, bx.lo[0], bx.L[0], g.nx, gx + k, g.px), uy, uz) < rl2;
          if (!keep[0]) beg = keep[1] ? l.hc[h0 + 1].lstart : l.hc[h0 + 2].lstart;
          if (!keep[2]) end = keep[1] ? l.hc[h0 + 2].lstart : l.hc[h0 + 1].lstart;
        }
        for (int c = beg; c < end; c += 2) {
          // record c + 1 <= cap exists (the next cell's first atom, or padding)
          const double3 p0 = P[c], p1 = P[c + 1];
          double u0 = p0.x - xi, v0 = p0.y - yi, w0 = p0.z - zi;
          double u1 = p1.x - xi, v1 = p1.y - yi, w1 = p1.z - zi;
          if constexpr (XF) {
            xsmd_xform_apply(xf->m, &u0, &v0, &w0);
            xsmd_xform_apply(xf->m, &u1, &v1, &w1);
          }
          const double d0 = xsmd_d2(u0, v0, w0);
          const double d1 = xsmd_d2(u1, v1, w1);
          if (d0 < rl2 && c != a.local) hit(c, d0);
          if (c + 1 < end && d1 < rl2 && c + 1 != a.local) hit(c + 1, d1);
        }
        if (ready) flush();
      }
    if ((cnt & 7) != 0) {  // the last, partial group: move it down (unused places hold slot 0)
      for (int k = cnt & 7; k < 8; k++) {
        fifo_push(w, 0u);
        bw[0] = __builtin_amdgcn_alignbit(bw[1], bw[0], 8);
        bw[1] = __builtin_amdgcn_alignbit(0u, bw[1], 8);
      }
      r[0] = w[0];
      r[1] = w[1];
      r[2] = w[2];
      r[3] = w[3];
      br[0] = bw[0];
      br[1] = bw[1];
      flush();
    }
    nnb[a.global] = cnt;
    atomicMax(l.total + 1, cnt);
    if (cnt > nmax) atomicMax(flags, (unsigned long long)cnt);
  }
  __syncthreads();
  if (threadIdx.x == 0) brick_rows[b] = l.total[1];
}

// Orders each atom's block-list entries by their distance keys (written by the
// build; stable, so equal keys keep ascending slot order), for lists of at most
// kSortRows entries.  At a given list position the atoms of a wave then look up
// nearby spline rows, which cuts the cache lines each table gather of the EAM
// passes touches by about 40 %.  One wave takes eight atoms: each lane loads one
// 8-entry group (slots and keys) of one atom, the groups go through LDS to one row
// per atom, and the wave sorts two rows at a time (ballot, popcount, ds_permute).
constexpr int kSortThreads = 256;
constexpr int kSortWaveAtoms = 8;
constexpr int kSortRowWords = 65;  // 64 keys + 1: the lanes of a group load start in distinct banks

__global__ void __launch_bounds__(kSortThreads) k_list_sort(int64_t n, uint16_t* __restrict__ nbr16,
                                                            const uint8_t* __restrict__ bins8,
                                                            const int* __restrict__ nnb, int64_t npad, int nbits)
{
  __shared__ uint32_t keys[kSortThreads / 64][kSortWaveAtoms][kSortRowWords];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t a0 = ((int64_t)blockIdx.x * (kSortThreads / 64) + w) * kSortWaveAtoms;
  const int la = lane >> 3, grp = lane & 7;
  const int64_t a = a0 + la;
  const int cnt = a < n ? nnb[a] : 0;
  const bool sorts = cnt >= 2 && cnt <= kSortRows;
  uint32_t* row = keys[w][la];
  if (sorts && 8 * grp < cnt) {
    const int64_t at = bl_index(npad, a, 8 * grp);
    const uint4 sl = *reinterpret_cast<const uint4*>(nbr16 + at);
    const uint2 bn = *reinterpret_cast<const uint2*>(bins8 + at);
    const uint32_t h[4] = {sl.x, sl.y, sl.z, sl.w};
#pragma unroll
    for (int j = 0; j < 8; j++) {  // storage order (0, 2, 4, 6, 1, 3, 5, 7)
      const int k = 8 * grp + (j < 4 ? 2 * j : 2 * (j - 4) + 1);
      const uint32_t slot = (h[j >> 1] >> (16 * (j & 1))) & 0xffffu;
      const uint32_t bin = ((j < 4 ? bn.x : bn.y) >> (8 * (j & 3))) & 0xffu;
      row[k] = (bin << 16) | slot;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortWaveAtoms; r += 2) {
    const int c0 = __shfl(cnt, 8 * r), c1 = __shfl(cnt, 8 * (r + 1));
    const bool s0 = c0 >= 2 && c0 <= kSortRows, s1 = c1 >= 2 && c1 <= kSortRows;
    if (!s0 && !s1) continue;  // uniform across the wave
    uint32_t k0 = (s0 && lane < c0) ? keys[w][r][lane] : 0xffffffffu;
    uint32_t k1 = (s1 && lane < c1) ? keys[w][r + 1][lane] : 0xffffffffu;
    wave_radix_sort2(k0, k1, 16, nbits);
    if (s0 && lane < c0) keys[w][r][lane] = k0;
    if (s1 && lane < c1) keys[w][r + 1][lane] = k1;
  }
  __syncthreads();
  if (sorts && 8 * grp < cnt) {
    uint32_t o[4];
#pragma unroll
    for (int j2 = 0; j2 < 4; j2++)  // list order (2 j2, 2 j2 + 1)
      o[j2] = (row[8 * grp + 2 * j2] & 0xffffu) | (row[8 * grp + 2 * j2 + 1] << 16);
    *reinterpret_cast<uint4*>(nbr16 + bl_index(npad, a, 8 * grp)) = group_interleave(o);
  }
}

// Block-local entries -> global (index, image) entries of the same list, in the
// stencil order of the global build: a distance-sorted list goes back to ascending
// halo slot order (the stencil order) first.  One wave per central atom.
__global__ void __launch_bounds__(256) k_block_to_global(BlockGeom g, const int* __restrict__ cell_off,
                                                         const uint16_t* __restrict__ nbr16,
                                                         const int* __restrict__ nnb, uint32_t* __restrict__ nbr,
                                                         int64_t npad, int nh_max)
{
  const int64_t b = lblock();
  if (b >= g.nblocks) return;
  LdsLayout l = lds_layout(0, 0, nh_max, false, false);
  const BlockShape s = block_shape(g, b);
  const int nh = s.hx * s.hy * s.hz;
  build_halo(g, s, cell_off, l.hc, l.total);
  const int lane = threadIdx.x & 63;
  for (int q = threadIdx.x >> 6;; q += blockDim.x >> 6) {
    Central a;
    if (!central_atom(s, l.hc, q, a)) break;
    const int cnt = nnb[a.global];
    if (cnt <= kSortRows) {
      uint32_t key = 0xffffu;
      if (lane < cnt) key = nbr16[bl_index(npad, a.global, lane)];
      key = wave_radix_sort(key, 0, 16);
      if (lane < cnt) {
        const int c = (int)key;
        const HaloCell e = l.hc[find_cell(l.hc, nh, c)];
        nbr[(int64_t)lane * npad + a.global] =
            (uint32_t)(e.gstart + (c - e.lstart)) | ((uint32_t)e.img << XSMD_NBR_INDEX_BITS);
      }
    } else {
      for (int k = lane; k < cnt; k += 64) {
        const int c = nbr16[bl_index(npad, a.global, k)];
        const HaloCell e = l.hc[find_cell(l.hc, nh, c)];
        nbr[(int64_t)k * npad + a.global] =
            (uint32_t)(e.gstart + (c - e.lstart)) | ((uint32_t)e.img << XSMD_NBR_INDEX_BITS);
      }
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
  int cap;   // halo slots (LDS records)
  int ccap;  // central atoms per brick at most
  double rc2;
  // set to 1 by a brick whose halo or central atoms exceed cap or ccap (the maxima
  // the host measured at list build, which hold while the cells do not change); such
  // a brick leaves its atoms unwritten and the host reports the call as failed
  unsigned long long* overflow;
  const XformDev* xf;  // the staging maps positions to physical space when not NULL
};

// a brick beyond the measured capacities: raise the flag, skip the brick
__device__ __forceinline__ bool over_capacity(const BlockArgs& a, int count, int cap)
{
  if (count <= cap) return false;
  if (threadIdx.x == 0) *a.overflow = 1ull;
  return true;
}

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

// List slots of entries k0, k0 + 2, k0 + 4, k0 + 6 of atom gi (this lane's parity of
// the list).  k0 = half + 8 q, so one 8-byte word brings the four entries of group q
// (in bounds for k0 < nmax); lslots turns entries past this atom's count into the far
// record (a pair longer than any cutoff).  The passes load the next group's word one
// iteration ahead (none at the end of the list).  In the bl_index layout the word of
// (gi, half + 8 q) is word 2 gi + half of group q, whose words start q * 2 npad words
// into the array.  Every lane of a wave is at the same q in a given iteration, so
// the group's start is one scalar address and each lane adds its own fixed 32-bit
// offset (the list holds at most 2^26 atoms x 64 entries = 2^30 words).
struct LWords {
  const uint2* base;
  uint32_t lane;  // this lane's word within a group
  int64_t step;   // words per group
  __device__ __forceinline__ uint2 get(int q) const { return (base + q * step)[lane]; }
};
__device__ __forceinline__ LWords lwords(const BlockArgs& a, int64_t gi, int half)
{
  return {reinterpret_cast<const uint2*>(a.nbr16), (uint32_t)(2 * gi + half), 2 * a.npad};
}
__device__ __forceinline__ void lslots(uint2 w, int far_slot, int k0, int cnt, int (&c)[4])
{
  const int raw[4] = {(int)(w.x & 0xffffu), (int)(w.x >> 16), (int)(w.y & 0xffffu), (int)(w.y >> 16)};
  // slots are 16-bit: saying so of the far slot too keeps the record addresses 24-bit
  // multiplies instead of 64-bit ones
  const int fs = far_slot & 0xffff;
#pragma unroll
  for (int u = 0; u < 4; u++) c[u] = (k0 + 2 * u < cnt) ? raw[u] : fs;
}

// neighbour records of slots c[0..U-1] (entries k0, k0 + 2, ...)
template <int U>
__device__ __forceinline__ void lgather(const BlockArgs& a, const LRec& rec, const int* c, int k0, int cnt,
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
    nb[u].in = nb[u].d2 <= a.rc2;  // entries past the count read the far record
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
  LdsLayout l = lds_layout(a.cap, a.ccap, a.g.nh_max, MULTI, false);
  const BlockShape s = block_shape(a.g, b);
  const int total = build_halo(a.g, s, a.cell_off, l.hc, l.total);
  if (over_capacity(a, total, a.cap)) return;
  central_table(s, l.hc, l.central_local, a.ccap, l.total + 2);
  stage(s, l.hc, total, a.pos, a.box[s.r], a.xf, a.type, l.rec, l.ltype, l.slot_cell, a.cap);
  const int ncentral = l.total[2];
  if (over_capacity(a, ncentral, a.ccap)) return;
  const int half = pass_half();
  for (int q = pass_slot();; q += kPassAtoms) {
    Central ca;
    if (!central_at(l.hc, l.slot_cell, l.central_local, ncentral, q, ca)) break;
    const Pos4 pi = l.rec.get(ca.local);
    const int ti = MULTI ? l.ltype[ca.local] : 0;
    const bool oni = !MULTI || pm::eam_on(t, ti);
    const int cnt = oni ? a.nnb[ca.global] : 0;
    double rho = 0.0;
    LWords lw = lwords(a, ca.global, half);
    uint2 wnext = lw.get(0);
    for (int k0 = half, q = 0; k0 < cnt; k0 += 2 * kBBatch, q++) {
      const uint2 w = wnext;
      if (k0 + 8 < cnt) wnext = lw.get(q + 1);
      int c[kBBatch];
      lslots(w, a.cap, k0, cnt, c);
      LNb nb[kBBatch];
      lgather<kBBatch>(a, l.rec, c, k0, cnt, pi, nb);
      if constexpr (MULTI) {
#pragma unroll
        for (int u = 0; u < kBBatch; u++) nb[u].in = nb[u].in && pm::eam_on(t, l.ltype[nb[u].c]);
      }
      // a batch with no pair inside rc on any lane of the wave (the tails of the
      // distance-sorted lists) skips the table reads; skipped terms would add 0
      bool any = false;
#pragma unroll
      for (int u = 0; u < kBBatch; u++) any |= nb[u].in;
      if (!__any(any)) continue;
#pragma unroll
      for (int u = 0; u < kBBatch; u++) {
        const int tj = MULTI ? l.ltype[nb[u].c] : 0;
        // every d2 here is finite and positive (entries past the count read the far
        // record), so the knot index needs no guard value; the sum masks by `in`
        const Real val = pm::eam_rho_knot<Real, MULTI>(t, pair, ti, tj, nb[u].d2, true);
        rho += (double)(nb[u].in ? val : Real(0));
      }
    }
    rho = half_sum(rho);
    if (half == 0) {
      rho_io[ca.global] = rho;
      if constexpr (EMB) {
        Real F, Fp;
        pm::embed(t, ti, rho, F, Fp);
        if (!oni) F = Fp = Real(0);  // a disabled species keeps rho_dEmb = 0 and adds no F
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
  LdsLayout l = lds_layout(a.cap, a.ccap, a.g.nh_max, MULTI, true);
  const BlockShape s = block_shape(a.g, b);
  const int total = build_halo(a.g, s, a.cell_off, l.hc, l.total);
  if (over_capacity(a, total, a.cap)) return;
  stage(s, l.hc, total, a.pos, a.box[s.r], a.xf, a.type, l.rec, l.ltype, l.slot_cell, a.cap);
  const int half = pass_half();
  for (int q = pass_slot();; q += kPassAtoms) {
    Central ca;
    if (!central_atom(s, l.hc, q, ca)) break;
    const Pos4 pi = l.rec.get(ca.local);
    const int ti = MULTI ? l.ltype[ca.local] : 0;
    const Real fpi = (Real)pi.w;
    const int cnt = (!MULTI || pm::eam_on(t, ti)) ? a.nnb[ca.global] : 0;
    double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    LWords lw = lwords(a, ca.global, half);
    uint2 wnext = lw.get(0);
    for (int k0 = half, q = 0; k0 < cnt; k0 += 2 * kBBatch, q++) {
      const uint2 w = wnext;
      if (k0 + 8 < cnt) wnext = lw.get(q + 1);
      int c[kBBatch];
      lslots(w, a.cap, k0, cnt, c);
      // two entries at a time: their four table rows stay in flight together within
      // the register budget of six waves per SIMD (the barrier keeps the compiler
      // from hoisting the second pair's gathers above the first pair's rows)
#pragma unroll
      for (int sub = 0; sub < kBBatch; sub += 2) {
        LNb nb[2];
        lgather<2>(a, l.rec, c + sub, k0 + 2 * sub, cnt, pi, nb);
        if constexpr (MULTI) {
          nb[0].in = nb[0].in && pm::eam_on(t, l.ltype[nb[0].c]);
          nb[1].in = nb[1].in && pm::eam_on(t, l.ltype[nb[1].c]);
        }
        if (!__any(nb[0].in || nb[1].in)) continue;  // as in the density pass
        auto add = [&](const LNb& q, bool in, Real fp, Real ph) {
          const double fpair = in ? (double)fp : 0.0;
          const double fe0 = q.x * fpair, fe1 = q.y * fpair, fe2 = q.z * fpair;
          fx += fe0;
          fy += fe1;
          fz += fe2;
          if constexpr (ENERGY) e += in ? 0.5 * (double)ph : 0.0;
          if constexpr (VIRIAL) pm::add_virial(v, fe0, fe1, fe2, q.x, q.y, q.z);
        };
        if constexpr (pm::kSTable<Real, MULTI>) {
          // FP32 single element: the s-table; the knot rows for pairs outside its range
          const int tj[2] = {0, 0};
          Real fp_r[2], phi[2] = {Real(0), Real(0)};
          bool slow[2];
          pm::eam_pair_batch<Real, MULTI, ENERGY, 2>(t, pair, ti, tj, fpi, nb, fp_r, phi, slow);
#pragma unroll
          for (int u = 0; u < 2; u++) add(nb[u], nb[u].in && !slow[u], fp_r[u], phi[u]);
          if (__builtin_expect(__any(slow[0] || slow[1]), 0)) {
#pragma unroll
            for (int u = 0; u < 2; u++)
              if (slow[u]) {
                Real fp, ph = Real(0);
                pm::eam_pair_knot<Real, false>(t, pair, 0, 0, fpi, (Real)nb[u].w, nb[u].d2, true, fp, ph);
                add(nb[u], true, fp, ph);
              }
          }
        } else {
#pragma unroll
          for (int u = 0; u < 2; u++) {
            const int tj = MULTI ? l.ltype[nb[u].c] : 0;
            Real fp_r, phi;
            pm::eam_pair_knot<Real, MULTI>(t, pair, ti, tj, fpi, (Real)nb[u].w, nb[u].d2, true, fp_r, phi);
            add(nb[u], nb[u].in, fp_r, phi);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
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

// Gather ceiling of the mixed single-element EAM passes (a measurement, not an
// operator): the staging, list walk, LDS neighbour reads, FP64 displacement, cutoff
// test and table-row index of the density pass (knot rows, FORCE = false) or of the
// force pass (s-table rows, FORCE = true), with the same table gathers, but none of
// the spline or pair arithmetic: every gathered row is summed into a per-atom sink,
// so no load is dead, and the sink is the only output.  Its time against the pass's
// own is the fraction of the pass that its data movement accounts for.  LEVEL takes
// parts away to split that time: 0 everything above; 1 without the table-row
// gathers (the row index still computed and summed); 2 the list words only (no
// neighbour records, no displacement); 3 the staging only.
template <bool FORCE, int LEVEL>
__global__ void __launch_bounds__(kPassThreads, kPassWaves<false>) k_block_gather_ceiling(BlockArgs a,
                                                                                            EamTabDev<float> t,
                                                                                            const float* __restrict__ pair,
                                                                                            double* __restrict__ sink)
{
  const int64_t b = lblock();
  if (b >= a.g.nblocks) return;
  LdsLayout l = lds_layout(a.cap, a.ccap, a.g.nh_max, false, FORCE);
  const BlockShape s = block_shape(a.g, b);
  const int total = build_halo(a.g, s, a.cell_off, l.hc, l.total);
  if (over_capacity(a, total, a.cap)) return;
  if constexpr (!FORCE) central_table(s, l.hc, l.central_local, a.ccap, l.total + 2);
  stage(s, l.hc, total, a.pos, a.box[s.r], a.xf, nullptr, l.rec, l.ltype, l.slot_cell, a.cap);
  int ncentral = 0;
  if constexpr (!FORCE) {
    ncentral = l.total[2];
    if (over_capacity(a, ncentral, a.ccap)) return;
  }
  const int half = pass_half();
  for (int q = pass_slot();; q += kPassAtoms) {
    Central ca;
    if constexpr (FORCE) {
      if (!central_atom(s, l.hc, q, ca)) break;
    } else {
      if (!central_at(l.hc, l.slot_cell, l.central_local, ncentral, q, ca)) break;
    }
    const Pos4 pi = l.rec.get(ca.local);
    const int cnt = a.nnb[ca.global];
    float acc = 0.0f;
    if constexpr (LEVEL == 3) {
      if (half == 0) sink[ca.global] = pi.x + cnt;
      continue;
    }
    LWords lw = lwords(a, ca.global, half);
    uint2 wnext = lw.get(0);
    for (int k0 = half, q = 0; k0 < cnt; k0 += 2 * kBBatch, q++) {
      const uint2 w = wnext;
      if (k0 + 8 < cnt) wnext = lw.get(q + 1);
      int c[kBBatch];
      lslots(w, a.cap, k0, cnt, c);
      if constexpr (LEVEL == 2) {
        acc += (float)((c[0] + c[1]) + (c[2] + c[3]));
        continue;
      }
      LNb nb[kBBatch];
      lgather<kBBatch>(a, l.rec, c, k0, cnt, pi, nb);
      bool any = false;
#pragma unroll
      for (int u = 0; u < kBBatch; u++) any |= nb[u].in;
      if (!__any(any)) continue;
#pragma unroll
      for (int u = 0; u < kBBatch; u++) {
        int m;
        float pp;
        if constexpr (FORCE) {
          bool out;
          pm::s_knot(t, nb[u].d2, m, pp, out);
          float v;
          if constexpr (LEVEL == 0) {
            const float* row = pm::at32(pair + t.s_force, (uint32_t)m * 8u);
            const pm::Q3 gr = pm::load3(row);
            const pm::Q3 gz = pm::load3(row + 4);
            v = ((gr.a + gr.b) + (gr.c + gz.a)) + ((gz.b + gz.c) + pp) + (float)nb[u].w;
          } else {
            v = (pp + (float)m) + (float)nb[u].w;
          }
          acc += nb[u].in ? v : 0.0f;
        } else {
          float rinv;
          pm::knot_dist(t, nb[u].d2, m, pp, rinv);
          float v;
          if constexpr (LEVEL == 0) {
            const pm::Q4<float> r = pm::load4(pm::pair_rho_row(pair, m));
            v = ((r.a + r.b) + (r.c + r.d)) + (pp + rinv);
          } else {
            v = (pp + rinv) + (float)m;
          }
          acc += nb[u].in ? v : 0.0f;
        }
      }
    }
    const double sum = half_sum((double)acc);
    if (half == 0) sink[ca.global] = sum;
  }
}

// LJ forces (lj_compute_force, pair_potential_force_op_singlemat.h:45-218) on the
// block list: the staging, list walk and epilogues of k_block_eam_force with the
// pair math of the global-list kernel (pm::lj_pair).
// MULTI (lj_multi_force): types staged with the records; each pair takes the
// parameters and cutoff of its two species.
template <class Real, bool ENERGY, bool VIRIAL, bool FUSED, bool DRIFT, bool MULTI = false>
__global__ void __launch_bounds__(kPassThreads, kPassWaves<VIRIAL>) k_block_lj(BlockArgs a, pm::LJ<Real> lp,
                                                                           pm::Outputs o, pm::Drift dr,
                                                                           pm::LJMultiDev lm = {})
{
  static_assert(!DRIFT || (FUSED && !ENERGY && !VIRIAL), "the drift epilogue is the plain MD-loop step");
  const int64_t b = lblock();
  if (b >= a.g.nblocks) return;
  LdsLayout l = lds_layout(a.cap, a.ccap, a.g.nh_max, MULTI, false);
  const BlockShape s = block_shape(a.g, b);
  const int total = build_halo(a.g, s, a.cell_off, l.hc, l.total);
  if (over_capacity(a, total, a.cap)) return;
  central_table(s, l.hc, l.central_local, a.ccap, l.total + 2);
  stage(s, l.hc, total, a.pos, a.box[s.r], a.xf, MULTI ? a.type : nullptr, l.rec, l.ltype, l.slot_cell,
        a.cap);
  const int ncentral = l.total[2];
  if (over_capacity(a, ncentral, a.ccap)) return;
  const int half = pass_half();
  for (int q = pass_slot();; q += kPassAtoms) {
    Central ca;
    if (!central_at(l.hc, l.slot_cell, l.central_local, ncentral, q, ca)) break;
    const Pos4 pi = l.rec.get(ca.local);
    const int ti = MULTI ? l.ltype[ca.local] : 0;
    const int cnt = a.nnb[ca.global];
    double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    LWords lw = lwords(a, ca.global, half);
    uint2 wnext = lw.get(0);
    for (int k0 = half, q = 0; k0 < cnt; k0 += 2 * kBBatch, q++) {
      const uint2 w = wnext;
      if (k0 + 8 < cnt) wnext = lw.get(q + 1);
      int c[kBBatch];
      lslots(w, a.cap, k0, cnt, c);
      LNb nb[kBBatch];
      lgather<kBBatch>(a, l.rec, c, k0, cnt, pi, nb);
#pragma unroll
      for (int u = 0; u < kBBatch; u++) {
        double de_r, eh;
        if constexpr (MULTI) {
          const int pid = pm::type_pair(ti, l.ltype[nb[u].c]);
          pm::lj_pair(pm::lj_multi_row<Real>(lm, pid), nb[u].d2, nb[u].in && pm::lj_multi_in(lm, pid, nb[u].d2),
                      de_r, eh);
        } else {
          pm::lj_pair(lp, nb[u].d2, nb[u].in, de_r, eh);
        }
        const double fe0 = nb[u].x * de_r, fe1 = nb[u].y * de_r, fe2 = nb[u].z * de_r;
        fx += fe0;
        fy += fe1;
        fz += fe2;
        if constexpr (ENERGY) e += eh;
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

// Fused next-step drift of the MD loop (ForceEpilogue::drift) for a force launch
// with the plain MD-loop epilogue; false when it does not apply.
bool drift_setup(Ctx& c, const ForceEpilogue& epi, unsigned flags, pm::Drift& dr)
{
  // under a non-identity xform the loop's own drift kernel (which maps the step by the
  // inverse xform) runs instead: the epilogue drifts from the staged record, which is
  // then a physical position, and it keeps the identity form's code
  if (!(epi.fused && epi.drift) || (flags & (XS_ENERGY | XS_VIRIAL)) || c.xf_dev()) return false;
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
  a.xf = c.xf_dev();
  a.nbr16 = c.nbr16.p;
  a.nnb = c.nnb.p;
  a.npad = c.npad;
  a.cap = c.bcap;
  a.ccap = c.bcentral;
  a.overflow = c.dscalar.p + 6;
  a.rc2 = rcut * rcut;
  const size_t lds_rho = block_lds_bytes(c, MULTI, false);
  const size_t lds = block_lds_bytes(c, MULTI, true);
  const unsigned g = block_grid(c);
  const bool energy = flags & XS_ENERGY;
  double* pos_w = reinterpret_cast<double*>(c.pos.p);
  if (do_rho) {
    ProfScope prof(c.prof, c.stream, XS_PROF_EAM_RHO);
#define XS_BRHO(EMB, EN, SET) \
  k_block_eam_rho<Real, MULTI, EMB, EN, SET><<<g, kPassThreads, lds_rho, c.stream>>>(a, t, pair, c.rho.p, pos_w, c.ep.p)
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

// Average time (ms) of `reps` launches of the gather-ceiling kernel of one pass at
// one level (see k_block_gather_ceiling) on the current block list and FP32
// single-element tables.
double gather_ceiling(Ctx& c, double rcut, bool force, int level, int reps)
{
  if (level < 0 || level > 3) throw Error(XS_E_ARG, "gather ceiling level must be 0..3");
  if (c.list_kind != ListKind::Block) throw Error(XS_E_STATE, "gather ceiling needs the cell-block list");
  if (!c.has_eam || eam_multi(c)) throw Error(XS_E_STATE, "gather ceiling needs single-element EAM tables");
  if (reps < 1) throw Error(XS_E_ARG, "reps must be >= 1");
  const EamTabDev<float> t = make_eam_tab<float>(c);
  if (force && c.eam.s_pieces == 0) throw Error(XS_E_STATE, "gather ceiling of the force pass needs the s-table");
  BlockArgs a;
  a.g = c.bgeom;
  a.cell_off = c.cell_off.p;
  a.pos = c.pos.p;
  a.type = c.type.p;
  a.box = c.box_d.p;
  a.xf = c.xf_dev();
  a.nbr16 = c.nbr16.p;
  a.nnb = c.nnb.p;
  a.npad = c.npad;
  a.cap = c.bcap;
  a.ccap = c.bcentral;
  a.overflow = c.dscalar.p + 6;
  a.rc2 = rcut * rcut;
  DBuf<double> sink;
  sink.reserve(c.npad);
  const unsigned g = block_grid(c);
  const size_t lds = block_lds_bytes(c, false, force);
  hipEvent_t e0, e1;
  XS_HIP(hipEventCreate(&e0));
  XS_HIP(hipEventCreate(&e1));
  XS_HIP(hipEventRecord(e0, c.stream));
  using K = void (*)(BlockArgs, EamTabDev<float>, const float*, double*);
  const K ks[2][4] = {{k_block_gather_ceiling<false, 0>, k_block_gather_ceiling<false, 1>,
                       k_block_gather_ceiling<false, 2>, k_block_gather_ceiling<false, 3>},
                      {k_block_gather_ceiling<true, 0>, k_block_gather_ceiling<true, 1>,
                       k_block_gather_ceiling<true, 2>, k_block_gather_ceiling<true, 3>}};
  const K k = ks[force ? 1 : 0][level];
  for (int r = 0; r < reps; r++) k<<<g, kPassThreads, lds, c.stream>>>(a, t, c.eam.pair_f.p, sink.p);
  XS_HIP(hipGetLastError());
  XS_HIP(hipEventRecord(e1, c.stream));
  XS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  XS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  check_pass_overflow(c);
  return ms / reps;
}

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

// Brick shape: the one that stages the fewest halo cells per grid cell (partial
// bricks at the grid edge count as whole ones) among the shapes whose expected
// central atoms stay within `bound` (one round of the 256 atom slots of a pass
// workgroup, with headroom for density fluctuations) and, unless XSMD_BLOCK_ATOMS
// set the bound, whose expected halo leaves LDS for three pass workgroups per CU
// (32-byte records, 15 % headroom).
void block_shape_choose(const Ctx& c, BlockGeom& g, double bound, double per_cell)
{
  const double lds_halo_cells = (double)(kLdsPerCu / 3 - 2048) / (32.0 * 1.15 * per_cell);
  int bx = 1, by = 1, bz = 1;
  double best = 1e300;
  for (int z = 1; z <= std::min(g.nz, 8); z++)
    for (int y = 1; y <= std::min(g.ny, 8); y++)
      for (int x = 1; x <= std::min(g.nx, 8); x++) {
        const int halo = (x + 2) * (y + 2) * (z + 2);
        if (x * y * z > 1 && (x * y * z * per_cell > bound || halo > kMaxHaloCells ||
                              (!c.block_atoms_set && halo > lds_halo_cells)))
          continue;
        const double nb = (double)((g.nx + x - 1) / x) * ((g.ny + y - 1) / y) * ((g.nz + z - 1) / z);
        const double cost = nb * halo / ((double)g.nx * g.ny * g.nz);
        if (cost < best * (1.0 - 1e-9)) {
          best = cost;
          bx = x;
          by = y;
          bz = z;
        }
      }
  g.bx = bx;
  g.by = by;
  g.bz = bz;
  g.nbx = (g.nx + bx - 1) / bx;
  g.nby = (g.ny + by - 1) / by;
  g.nbz = (g.nz + bz - 1) / bz;
  g.nblocks = (int64_t)g.nbx * g.nby * g.nbz * c.nrep;
  g.nh_max = (bx + 2) * (by + 2) * (bz + 2);
}

void block_setup(Ctx& c)
{
  BlockGeom& g = c.bgeom;
  g.nx = c.nc[0];
  g.ny = c.nc[1];
  g.nz = c.nc[2];
  g.px = c.periodic[0];
  g.py = c.periodic[1];
  g.pz = c.periodic[2];
  g.ncell_rep = c.ncell_rep;
  // The shape is chosen from the mean atoms per cell; where atoms crowd into part
  // of the grid (slabs, clusters) the measured halo can exceed the LDS budget, and
  // the bound shrinks until it fits (or the bricks are single cells).
  double per_cell = std::max((double)c.n / (double)c.ncell_tot, 1e-3);
  double bound = c.block_atoms;
  for (int attempt = 0;; attempt++) {
    block_shape_choose(c, g, bound, per_cell);
    XS_HIP(hipMemsetAsync(c.dscalar.p + 2, 0, 2 * sizeof(unsigned long long), c.stream));
    k_block_halo_max<<<grid_for(g.nblocks), kBlock, 0, c.stream>>>(g, c.cell_off.p, c.dscalar.p + 2);
    XS_HIP(hipGetLastError());
    XS_HIP(hipMemcpyAsync(c.hscalar + 2, c.dscalar.p + 2, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          c.stream));
    XS_HIP(hipStreamSynchronize(c.stream));
    c.bcap = (int)((c.hscalar[2] + 63) / 64 * 64);
    c.bcentral = (int)c.hscalar[3];
    const size_t budget = c.block_atoms_set ? kBlockLdsLimit : std::min(kBlockLdsLimit, kLdsPerCu / 3);
    if (block_lds_bytes(c, true, true) <= budget || g.nh_max == 27 || attempt == 6) break;
    // the densest halo set the capacity: size the next shape for that density
    per_cell = std::max(per_cell, (double)c.hscalar[2] / g.nh_max);
    bound *= 0.75;
  }
}

size_t block_lds_bytes(const Ctx& c, bool with_type, bool with_w)
{
  // bcap records and the far record, halo table, counters, slot cells, central slots,
  // per-record types (lds_layout)
  return (size_t)(c.bcap + 1) * sizeof(double) * (with_w ? 4 : 3) + (size_t)c.bgeom.nh_max * sizeof(HaloCell) +
         4 * sizeof(int) + (size_t)(c.bcap + c.bcentral) * sizeof(uint16_t) + (with_type ? (size_t)c.bcap + 1 : 0);
}

bool block_feasible(const Ctx& c)
{
  return c.bcap > 0 && c.bcap < 65536 && block_lds_bytes(c, true, true) <= kBlockLdsLimit;
}

unsigned block_grid(const Ctx& c)
{
  return (unsigned)((c.bgeom.nblocks + 7) / 8 * 8);
}

// Block list build, then (unless XSMD_LIST_SORT=0) the distance order of each list
// of at most 64 entries: the build writes an 8-bit distance key per entry
// (2^bits bins uniform in d^2 over [0, rl^2)), k_list_sort orders by it.
void block_build(Ctx& c, double rl2)
{
  const size_t lds = build_rec_bytes(c.bcap) + (size_t)c.bgeom.nh_max * sizeof(HaloCell) + 4 * sizeof(int) +
                     (size_t)(c.bcap + c.bcentral) * sizeof(uint16_t);
  const bool sorted = !c.list_sort_disabled;
  const int bits = std::max(1, std::min(c.sort_bits, 8));
  const float bin_scale = (float)((double)(1 << bits) / rl2);
  for (int attempt = 0; attempt < 2; attempt++) {
    c.nbr16.reserve((size_t)c.nmax * c.npad);
    if (sorted) c.bins8.reserve((size_t)c.nmax * c.npad);
    c.nnb.reserve(c.npad);
    c.brick_rows.reserve(c.bgeom.nblocks);
    XS_HIP(hipMemsetAsync(c.dscalar.p, 0, 2 * sizeof(unsigned long long), c.stream));
    auto k = c.xf_dev() ? (sorted ? k_block_build<true, true> : k_block_build<false, true>)
                     : (sorted ? k_block_build<true, false> : k_block_build<false, false>);
    k<<<block_grid(c), kBuildThreads, lds, c.stream>>>(c.bgeom, c.cell_off.p, c.pos.p, c.box_d.p, rl2, c.nbr16.p,
                                                       c.nnb.p, c.nmax, c.npad, c.bcap, c.bcentral, c.bgeom.nh_max,
                                                       c.dscalar.p + 4, c.brick_rows.p, c.dscalar.p,
                                                       sorted ? c.bins8.p : nullptr, bin_scale, (1 << bits) - 1,
                                                       c.xf_dev());
    XS_HIP(hipGetLastError());
    XS_HIP(hipMemcpyAsync(c.hscalar, c.dscalar.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          c.stream));
    XS_HIP(hipStreamSynchronize(c.stream));
    if (c.hscalar[1] != 0) throw Error(XS_E_STATE, "block halo or central atoms exceed their measured capacity");
    const unsigned long long need = c.hscalar[0];
    if (need == 0) {
      if (sorted && c.n > 0) {
        const int64_t per_block = (int64_t)(kSortThreads / 64) * kSortWaveAtoms;
        const unsigned grid = (unsigned)((c.n + per_block - 1) / per_block);
        k_list_sort<<<grid, kSortThreads, 0, c.stream>>>(c.n, c.nbr16.p, c.bins8.p, c.nnb.p, c.npad, bits);
        XS_HIP(hipGetLastError());
      }
      return;
    }
    if (attempt == 1) throw Error(XS_E_STATE, "neighbour list overflow after resize");
    c.nmax = (int)((need + 7) / 8 * 8 + 8);
  }
}

void block_to_global(Ctx& c)
{
  c.nbr.reserve((size_t)c.nmax * c.npad);
  const size_t lds = block_lds_bytes(c, false, false) - (size_t)c.bcap * 3 * sizeof(double);
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
  const bool multi = eam_multi(c);
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

void launch_lj_block(Ctx& c, double eps, double sigma, double rcut, unsigned flags, const ForceEpilogue& epi)
{
  const double ecut = lj_energy_cutoff(eps, sigma, rcut);
  BlockArgs a;
  a.g = c.bgeom;
  a.cell_off = c.cell_off.p;
  a.pos = c.pos.p;
  a.type = c.type.p;
  a.box = c.box_d.p;
  a.xf = c.xf_dev();
  a.nbr16 = c.nbr16.p;
  a.nnb = c.nnb.p;
  a.npad = c.npad;
  a.cap = c.bcap;
  a.ccap = c.bcentral;
  a.overflow = c.dscalar.p + 6;
  a.rc2 = rcut * rcut;
  const size_t lds = block_lds_bytes(c, false, false);
  const unsigned g = block_grid(c);
  const pm::Outputs o = make_outputs(c, epi);
  const bool energy = flags & XS_ENERGY, vir = flags & XS_VIRIAL;
  pm::Drift dr{};
  const bool drift = drift_setup(c, epi, flags, dr);
  ProfScope prof(c.prof, c.stream, XS_PROF_LJ);
  auto run = [&](auto zero) {
    using Real = decltype(zero);
    const pm::LJ<Real> lp{(Real)(sigma * sigma), (Real)(4.0 * eps), (Real)(24.0 * eps), (Real)ecut};
#define XS_BL(EN, VI, FU, DR) k_block_lj<Real, EN, VI, FU, DR><<<g, kPassThreads, lds, c.stream>>>(a, lp, o, dr)
    if (drift) {
      XS_BL(false, false, true, true);
    } else if (epi.fused) {
      if (energy && vir) XS_BL(true, true, true, false);
      else if (energy) XS_BL(true, false, true, false);
      else if (vir) XS_BL(false, true, true, false);
      else XS_BL(false, false, true, false);
    } else {
      if (energy && vir) XS_BL(true, true, false, false);
      else if (energy) XS_BL(true, false, false, false);
      else if (vir) XS_BL(false, true, false, false);
      else XS_BL(false, false, false, false);
    }
#undef XS_BL
  };
  if (flags & XS_MIXED) run(0.0f);
  else run(0.0);
  XS_HIP(hipGetLastError());
  if (drift) drift_commit(c);
}

void launch_lj_multi_block(Ctx& c, unsigned flags)
{
  BlockArgs a;
  a.g = c.bgeom;
  a.cell_off = c.cell_off.p;
  a.pos = c.pos.p;
  a.type = c.type.p;
  a.box = c.box_d.p;
  a.xf = c.xf_dev();
  a.nbr16 = c.nbr16.p;
  a.nnb = c.nnb.p;
  a.npad = c.npad;
  a.cap = c.bcap;
  a.ccap = c.bcentral;
  a.overflow = c.dscalar.p + 6;
  a.rc2 = c.ljm_rcmax * c.ljm_rcmax;
  const size_t lds = block_lds_bytes(c, true, false);
  const unsigned g = block_grid(c);
  const pm::Outputs o = make_outputs(c, ForceEpilogue{});
  const bool energy = flags & XS_ENERGY, vir = flags & XS_VIRIAL;
  const pm::Drift dr{};
  ProfScope prof(c.prof, c.stream, XS_PROF_LJ);
  auto run = [&](auto zero) {
    using Real = decltype(zero);
    const pm::LJ<Real> none{};
    const pm::LJMultiDev lm = ljm_dev<Real>(c);
#define XS_BLM(EN, VI) \
  k_block_lj<Real, EN, VI, false, false, true><<<g, kPassThreads, lds, c.stream>>>(a, none, o, dr, lm)
    if (energy && vir) XS_BLM(true, true);
    else if (energy) XS_BLM(true, false);
    else if (vir) XS_BLM(false, true);
    else XS_BLM(false, false);
#undef XS_BLM
  };
  if (flags & XS_MIXED) run(0.0f);
  else run(0.0);
  XS_HIP(hipGetLastError());
}

void ensure_global_list(Ctx& c)
{
  if (c.list_kind == ListKind::Global || c.global_from_block) return;
  block_to_global(c);
  c.global_from_block = true;
}

void check_pass_overflow(Ctx& c)
{
  XS_HIP(hipMemcpyAsync(c.hscalar + 6, c.dscalar.p + 6, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        c.stream));
  XS_HIP(hipStreamSynchronize(c.stream));
  if (c.hscalar[6] == 0) return;
  c.hscalar[6] = 0;
  XS_HIP(hipMemsetAsync(c.dscalar.p + 6, 0, sizeof(unsigned long long), c.stream));
  XS_HIP(hipStreamSynchronize(c.stream));
  throw Error(XS_E_STATE, "block pass: a brick holds more halo or central atoms than measured at the list build "
                          "(cells changed without a rebuild); its atoms were not computed");
}

void check_or_defer_overflow(Ctx& c)
{
  if (c.stream == c.own_stream) {
    check_pass_overflow(c);
    return;
  }
  XS_HIP(hipMemcpyAsync(c.hscalar + 6, c.dscalar.p + 6, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        c.stream));
  XS_HIP(hipEventRecord(c.flag_event, c.stream));
  c.flag_pending = true;
}

void poll_pass_overflow(Ctx& c, bool wait)
{
  if (!c.flag_pending) return;
  if (!wait) {
    const hipError_t q = hipEventQuery(c.flag_event);
    if (q == hipErrorNotReady) return;
    XS_HIP(q);
  } else {
    XS_HIP(hipEventSynchronize(c.flag_event));
  }
  c.flag_pending = false;
  if (c.hscalar[6] == 0) return;
  c.hscalar[6] = 0;
  XS_HIP(hipMemsetAsync(c.dscalar.p + 6, 0, sizeof(unsigned long long), c.stream));
  throw Error(XS_E_STATE, "block pass (an earlier call): a brick holds more halo or central atoms than measured at "
                          "the list build (cells changed without a rebuild); its atoms were not computed");
}

}  // namespace xsmd
