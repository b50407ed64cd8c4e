// CDNA4 (gfx950) kernels of the coupled D3Q19 + D3Q7 lattice-Boltzmann step, two-lattice pull scheme, templated on
// the storage format (storage.h: f32, f16 / bf16 deviations, f64 twin).
//
// Thread mapping: one thread per cell. The x-y plane is flattened into columns c = j*nx + i so that the 64 lanes of a
// wavefront touch 64 consecutive cells of every population stream (one contiguous access per stream per wavefront in
// the [q][cell] layout, one or two in the blocked f32 layout; stores and unshifted loads are aligned whenever nx*ny is
// a multiple of 64). blockIdx.y is the level k, starting at
// the lowest fluid level of the grid, so all-solid planes are never launched. Solid cells above that plane exit.
//
// Three streaming paths feed one collision:
//   * fast path: every fluid cell below the lid (face and near-solid cells included, see FaceFix and BounceFix; the
//     description here is of the pulls). All 26 pulls are plain loads through buffer
//     descriptors: one per population with the pull offset folded into the scalar base, so the only per-lane address
//     is the cell's byte offset. At or above the column's kfast level (the first level where every source cell in the
//     3 x 3 x 3 neighbourhood is fluid, host-computed) that is all; below it (near-solid cells) the populations whose
//     source is solid are then replaced by the bounce-back value, from a per-column word with the neighbours' ktop.
//     A wavefront mixing cells next to terrain or roofs with free cells thus runs one pull sequence, not two.
//   * lid fast path (LidPull): lid cells off the faces whose neighbourhood is fluid take the same pulls, with the
//     populations from above the lid replaced by the free-slip mirror (D3Q19) or the own cell's value (D3Q7).
//   * boundary path: every branch of js/live/kernel.js:159-243 per population (lid first, then lateral faces, then the
//     interior pull with half-way bounce-back from solid sources), with the 3 x 3 neighbour ktop values loaded once;
//     for lid cells on a face or next to a solid, and near-solid cells of columns whose neighbour word cannot hold
//     the heights.
#include <algorithm>
#include <type_traits>

#include "kernels.h"
#include "storage.h"

namespace aht {

template <typename R>
struct Outflow {
  R rn, uxn, uyn, uzn, usqn, uxc, uyc, usqc;
  R own[19];   // the own cell's pre-stream populations (set on an outlet face only, where the outlet reads them)
};

// element x of a column or profile table of the Grid. The tables live in device memory, but their pointers are read
// from the Grid in device memory, so the compiler cannot prove them global and would emit flat loads. A pending flat
// load counts against the LDS/scalar counter as well, which makes every later scalar load (the stream table) wait for
// it and turns the waits of the pulled loads conservative (vmcnt(0) in the middle of the pulls); global loads do not.
template <class T>
__device__ __forceinline__ T ld_tab(const T *p, size_t x) {
  static_assert(sizeof(T) % 4 == 0 && sizeof(T) <= 16, "table elements are one to four dwords");
  typedef uint32_t W __attribute__((ext_vector_type(sizeof(T) / 4)));
  return __builtin_bit_cast(T, ((const __attribute__((address_space(1))) W *)p)[x]);
}

// base address of population q shifted back by `off` cells (the pull offset), as a plain integer ([q][cell] layout)
template <class S>
__device__ __forceinline__ const void *shifted(const void *a, size_t N, int q, int off) {
  static_assert(S::LB == 0, "a shifted base exists only in the [q][cell] layout");
  return (const void *)((uintptr_t)a + (uintptr_t)((int64_t)q * (int64_t)N - off) * (uintptr_t)S::bytes);
}
// first element of population q of a lattice with Q populations (cell offsets from there: cell_off<S, Q>)
template <class S, int Q>
__device__ __forceinline__ void *plane(void *a, size_t N, int q) {
  return (void *)((uintptr_t)a + plane_off<S>(Q, N, q));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ void split_column(const Grid &g, int c, int &i, int &j) {
  int jj = (int)((float)c * g.inv_nx);
  int ii = c - jj * g.nx;
  if (ii < 0) { jj -= 1; ii += g.nx; }
  else if (ii >= g.nx) { jj += 1; ii -= g.nx; }
  i = ii; j = jj;
}

// Tile order. Workgroup b runs on XCD b % 8 and is the (b / 8)-th one there; `win` (0 or a power of two) picks the
// tile it takes:
//   win == 1: tile b (linear; the eight XCDs share one front moving through memory);
//   win >= 2: windows of 8 * win tiles, XCD x taking the x-th run of `win` consecutive tiles of each window, so the
//             lines that neighbouring tiles share (the +-1 cell and +-1 row pulls) mostly meet in one L2 while the
//             front stays compact; the last partial window keeps the linear order;
//   win == 0: XCD x takes the x-th contiguous eighth of all tiles (eight fronts).
// Shifts and masks only: this runs once per wavefront.
__device__ __forceinline__ uint32_t tile_of_block(uint32_t b, uint32_t G, uint32_t win) {
  const uint32_t x = b & (NUM_XCD - 1), i = b >> 3;
  if (win == 0) {
    const uint32_t q = G >> 3, r = G & (NUM_XCD - 1);
    return x * q + min(x, r) + i;
  }
  if (win == 1) return b;
  const uint32_t lw = (uint32_t)__builtin_ctz(win), span_mask = (win << 3) - 1u;
  if (b >= (G & ~span_mask)) return b;
  return ((i >> lw) << (lw + 3)) + (x << lw) + (i & (win - 1u));
}

// q = t / d, r = t - q d for a wave-uniform t < 2^26 and 1 <= d < 2^20, from a host-computed float 1/d (the float
// estimate is off by at most one, corrected once): a few instructions instead of an emulated integer division
__device__ __forceinline__ uint32_t div_by(uint32_t t, uint32_t d, float inv_d, uint32_t &r) {
  uint32_t q = (uint32_t)((float)t * inv_d);
  int32_t rr = (int32_t)(t - q * d);
  if (rr < 0) { q -= 1; rr += (int32_t)d; }
  else if (rr >= (int32_t)d) { q += 1; rr -= (int32_t)d; }
  r = (uint32_t)rr;
  return q;
}

// packed column word: kt | ktop << 10 | kfast << 20 | near << 30 (nz < 1023); near: the column's neighbour word
// (Grid::colnb) is exact, so its near-solid interior cells may take the fast path with the bounce-back fix-up
__device__ __forceinline__ int col_kt(uint32_t x) { return (int)(x & 1023u); }
__device__ __forceinline__ int col_ktop(uint32_t x) { return (int)((x >> 10) & 1023u); }
__device__ __forceinline__ int col_kfast(uint32_t x) { return (int)((x >> 20) & 1023u); }
__device__ __forceinline__ bool col_near(uint32_t x) { return (x >> 30) & 1u; }

// element access of population qq at cell index idx, decoded with weight w19<q> (deviation formats)
template <class S, int q>
__device__ __forceinline__ typename S::real ldf(const typename S::elem *__restrict__ f, size_t N, int qq, uint32_t idx) {
  return S::dec_f(f[lat_index<S>(19, N, qq, idx)], w19<q>());
}
__device__ __forceinline__ float w7q(int q) { return q == 0 ? 0.25f : 0.125f; }

// theta population qq at cell index idx (theta-referenced formats add w_qq times the stored theta of that cell)
template <class S>
__device__ __forceinline__ typename S::real ldg(const Grid &g, const void *gA, int qq, uint32_t idx) {
  using E = typename S::elem;
  const E e = ((const E *)gA)[lat_index<S>(7, g.N, qq, idx)];
  if constexpr (S::theta_ref) return S::dec_g(e, w7q(qq), ld_theta(gA, g.th_off + idx * 4u));
  else return S::dec_g(e, 0.0f, typename S::real(0));
}

// Stream bases. TAB = true (the single-domain step, batched launches) reads them from a host-built Streams table;
// TAB = false (the two-step sweep, the surface read) forms them from the lattice pointers. Forming them on the device
// keeps ~60 64-bit bases live in scalar registers: in the batched kernel that spilled ~980 SGPRs into VGPR lanes.
template <class S, bool TAB, int Q>
__device__ __forceinline__ void *plane_base(const Streams *sp, int slot, void *lat, size_t N, int q) {
  if constexpr (TAB) return (void *)sp->wr[slot];
  else return plane<S, Q>(lat, N, q);
}

// Random bits of the stochastic-rounding formats: two 32-bit words per cell and step from an integer mix of the cell
// index and the step index (the cell's own bits in any launch shape or batch); value v (0..18 the D3Q19 populations,
// 19..25 the D3Q7 ones) takes a rotation of one word, the words alternating between odd and even v so that the two
// populations of an opposite pair get independent bits.
struct Dither {
  uint32_t h[2];
  Dither() = default;
  __device__ __forceinline__ Dither(uint32_t n, uint32_t step) {
    uint32_t x = n * 0x9e3779b1u ^ step * 0x85ebca77u;
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    h[0] = x;
    x ^= 0x68e31da4u; x *= 0xb5297a4du; x ^= x >> 15; x *= 0x1b56c4e9u; x ^= x >> 16;
    h[1] = x;
  }
  __device__ __forceinline__ uint32_t bits(int v) const {
    return __builtin_amdgcn_alignbit(h[v & 1], h[v & 1], (uint32_t)(5 * (v >> 1)) & 31u);
  }
};

// the seven post-collision theta populations p of cell n: raw, or theta = sum_q p_q in f32 plus the deviations
template <class S, bool TAB>
__device__ __forceinline__ void store_g(const Grid &g, const Streams *sp, void *gB, uint32_t n,
                                        const typename S::real (&p)[7], const Dither &dz) {
  using R = typename S::real;
  const uint32_t boff = cell_off<S, 7>(n);
  R th = R(0);
  if constexpr (S::theta_ref) {
    th = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + p[6]);
    if constexpr (TAB) st_theta((void *)sp->wr[26], n * 4u, th);
    else st_theta(gB, g.th_off + n * 4u, th);
  }
#pragma unroll
  for (int q = 0; q < 7; q++) {
    typename S::elem e;
    if constexpr (S::SR) e = S::enc_g_sr(p[q], w7q(q), th, dz.bits(19 + q));
    else e = S::enc_g(p[q], w7q(q), th);
    S::st_raw(plane_base<S, TAB, 7>(sp, 19 + q, gB, g.N, q), boff, e);
  }
}

template <int q>
struct OwnPops {
  template <class S>
  __device__ __forceinline__ static void run(typename S::real (&a)[19], const typename S::elem *__restrict__ fA,
                                             size_t N, uint32_t n) {
    a[q] = ldf<S, q>(fA, N, q, n);
    OwnPops<q + 1>::template run<S>(a, fA, N, n);
  }
};
template <>
struct OwnPops<19> {
  template <class S>
  __device__ __forceinline__ static void run(typename S::real (&)[19], const typename S::elem *__restrict__, size_t,
                                             uint32_t) {}
};

// outflow-face prelude from the own cell's pre-stream populations (js/live/kernel.js:122-148), once o.own holds them
// (onb: the cell lies on an outlet face; otherwise o.own is not read)
template <typename R>
__device__ __forceinline__ void outflow_finish(const Grid &g, Outflow<R> &o, bool onb, int i, int j) {
  const Params &P = g.P;
  o.rn = R(1); o.uxn = R(0); o.uyn = R(0); o.uzn = R(0);
  if (onb) {
    const R(&a)[19] = o.own;
    R r = a[0], sx = R(0), sy = R(0), sz = R(0);
#pragma unroll
    for (int q = 1; q < 19; q++) {
      r += a[q];
      sx += R(CX[q]) * a[q];
      sy += R(CY[q]) * a[q];
      sz += R(CZ[q]) * a[q];
    }
    const R ir = rcp(r);
    o.rn = r; o.uxn = sx * ir; o.uyn = sy * ir; o.uzn = sz * ir;
  }
  o.usqn = o.uxn * o.uxn + o.uyn * o.uyn + o.uzn * o.uzn;
  R uxc = o.uxn, uyc = o.uyn;
  if (P.oclip != 0) {                       // backflow-clipped outlet: no inward normal velocity
    if (i == 0 && P.inflow_w == 0 && uxc > R(0)) uxc = R(0);
    if (i == g.nx - 1 && P.inflow_e == 0 && uxc < R(0)) uxc = R(0);
    if (j == 0 && P.inflow_s == 0 && uyc > R(0)) uyc = R(0);
    if (j == g.ny - 1 && P.inflow_n == 0 && uyc < R(0)) uyc = R(0);
  }
  o.uxc = uxc; o.uyc = uyc;
  o.usqc = uxc * uxc + uyc * uyc + o.uzn * o.uzn;
}

// the prelude with the own cell's populations loaded from lattice fA
template <class S>
__device__ __forceinline__ Outflow<typename S::real> outflow_prelude(const Grid &g,
                                                                     const typename S::elem *__restrict__ fA, size_t N,
                                                                     uint32_t n, int i, int j) {
  using R = typename S::real;
  const Params &P = g.P;
  Outflow<R> o;
  const bool onb = (i == 0 && P.inflow_w == 0) || (i == g.nx - 1 && P.inflow_e == 0) ||
                   (j == 0 && P.inflow_s == 0) || (j == g.ny - 1 && P.inflow_n == 0);
  if (onb) OwnPops<0>::template run<S>(o.own, fA, N, n);
  outflow_finish(g, o, onb, i, j);
  return o;
}

// population q pulled from outside the lateral faces (js/live/kernel.js:186-205): inflow equilibrium with rho = 1, or
// the pressure outlet's non-equilibrium extrapolation towards rho = 1 (optionally backflow-clipped). Shared by the
// boundary path and the fast path's face fix-up, so both give the same bits.
template <class S, int q>
__device__ __forceinline__ typename S::real lateral19(const Grid &g, int i, int j, typename S::real uI,
                                                      typename S::real vI, const Outflow<typename S::real> &ob) {
  using R = typename S::real;
  constexpr int cx = CX[q], cy = CY[q], cz = CZ[q];
  const Params &P = g.P;
  const int NX = g.nx, NY = g.ny;
  const bool inflow = (cx == 1 && i == 0 && P.inflow_w != 0) || (cx == -1 && i == NX - 1 && P.inflow_e != 0) ||
                      (cy == 1 && j == 0 && P.inflow_s != 0) || (cy == -1 && j == NY - 1 && P.inflow_n != 0);
  const R w = R(w19<q>());
  if (inflow) return w * epoly(cdot<cx, cy, 0>(uI, vI, R(0)), uI * uI + vI * vI);   // equilibrium, rho = 1
  const R cu = cdot<cx, cy, cz>(ob.uxn, ob.uyn, ob.uzn);
  const R own = ob.own[q];   // not an inflow face: an outlet face, whose prelude loaded the own cell
  if (P.oclip != 0) {
    const R cuc = cdot<cx, cy, cz>(ob.uxc, ob.uyc, ob.uzn);
    return own + w * (epoly(cuc, ob.usqc) - ob.rn * epoly(cu, ob.usqn));
  }
  return own + w * (R(1) - ob.rn) * epoly(cu, ob.usqn);
}

// post-stream population q of cell (i, j, k), every boundary branch (js/live/kernel.js:159-217)
template <class S, int q>
__device__ __forceinline__ typename S::real pull19(const Grid &g, const typename S::elem *__restrict__ fA, size_t N,
                                                   uint32_t n, int i, int j, int k, const int (&kn)[3][3],
                                                   typename S::real uI, typename S::real vI,
                                                   const Outflow<typename S::real> &ob) {
  constexpr int cx = CX[q], cy = CY[q], cz = CZ[q];
  const int NX = g.nx, NY = g.ny;
  if (cz == -1 && k == g.nz - 1) {
    // free-slip lid: the mirrored population of the lateral neighbour at the same level
    constexpr int m = MIRZ[q];
    const int si = i - cx, sj = j - cy;
    if (si < 0 || si >= NX || sj < 0 || sj >= NY) return ldf<S, q>(fA, N, m, n);
    if (k < kn[1 - cx][1 - cy]) return ldf<S, q>(fA, N, m, n);
    return ldf<S, q>(fA, N, m, n - (uint32_t)(cx + cy * NX));
  }
  const bool outx = (cx == 1 && i == 0) || (cx == -1 && i == NX - 1);
  const bool outy = (cy == 1 && j == 0) || (cy == -1 && j == NY - 1);
  if (outx || outy) return lateral19<S, q>(g, i, j, uI, vI, ob);
  // interior: half-way bounce-back when the source cell is solid
  const int sk = k - cz;
  if (sk < kn[1 - cx][1 - cy]) return ldf<S, q>(fA, N, OPP[q], n);
  const int off = cx + cy * NX + cz * (int)g.NXY;
  return ldf<S, q>(fA, N, q, (uint32_t)((int)n - off));
}

// theta population q pulled from outside the lateral faces (js/live/kernel.js:231-236): 0 at an inflow face, the own
// cell's value at an outlet
template <class S, int q>
__device__ __forceinline__ typename S::real lateral7(const Grid &g, const void *gA, uint32_t n, int i, int j) {
  using R = typename S::real;
  constexpr int cx = CX7[q], cy = CY7[q];
  const Params &P = g.P;
  const bool inflow = (cx == 1 && i == 0 && P.inflow_w != 0) || (cx == -1 && i == g.nx - 1 && P.inflow_e != 0) ||
                      (cy == 1 && j == 0 && P.inflow_s != 0) || (cy == -1 && j == g.ny - 1 && P.inflow_n != 0);
  return inflow ? R(0) : ldg<S>(g, gA, q, n);
}

// post-stream theta population q (js/live/kernel.js:219-243)
template <class S, int q>
__device__ __forceinline__ typename S::real pull7(const Grid &g, const void *gA, uint32_t n, int i, int j, int k,
                                                  const int (&kn)[3][3]) {
  constexpr int cx = CX7[q], cy = CY7[q], cz = CZ7[q];
  if (cz == -1 && k == g.nz - 1) return ldg<S>(g, gA, q, n);   // zero-gradient top
  const bool outx = (cx == 1 && i == 0) || (cx == -1 && i == g.nx - 1);
  const bool outy = (cy == 1 && j == 0) || (cy == -1 && j == g.ny - 1);
  if (outx || outy) return lateral7<S, q>(g, gA, n, i, j);
  const int sk = k - cz;
  if (sk < kn[1 - cx][1 - cy]) return ldg<S>(g, gA, OPP7[q], n);
  const int off = cx + cy * g.nx + cz * (int)g.NXY;
  return ldg<S>(g, gA, q, (uint32_t)((int)n - off));
}

template <int q>
struct Pull19 {
  template <class S>
  __device__ __forceinline__ static void run(typename S::real (&f)[19], const Grid &g,
                                             const typename S::elem *__restrict__ fA, size_t N, uint32_t n, int i,
                                             int j, int k, const int (&kn)[3][3], typename S::real uI,
                                             typename S::real vI, const Outflow<typename S::real> &ob) {
    f[q] = pull19<S, q>(g, fA, N, n, i, j, k, kn, uI, vI, ob);
    Pull19<q + 1>::template run<S>(f, g, fA, N, n, i, j, k, kn, uI, vI, ob);
  }
};
template <>
struct Pull19<19> {
  template <class S>
  __device__ __forceinline__ static void run(typename S::real (&)[19], const Grid &,
                                             const typename S::elem *__restrict__, size_t, uint32_t, int, int, int,
                                             const int (&)[3][3], typename S::real, typename S::real,
                                             const Outflow<typename S::real> &) {}
};

template <int q>
struct Pull7 {
  template <class S>
  __device__ __forceinline__ static void run(typename S::real (&gg)[7], const Grid &g, const void *gA, uint32_t n,
                                             int i, int j, int k, const int (&kn)[3][3]) {
    gg[q] = pull7<S, q>(g, gA, n, i, j, k, kn);
    Pull7<q + 1>::template run<S>(gg, g, gA, n, i, j, k, kn);
  }
};
template <>
struct Pull7<7> {
  template <class S>
  __device__ __forceinline__ static void run(typename S::real (&)[7], const Grid &, const void *, uint32_t, int, int,
                                             int, const int (&)[3][3]) {}
};

// fast path: all 26 sources are fluid interior cells; plain shifted loads through per-population descriptors
template <int q>
struct FastPull {
  // population q of both lattices (D3Q7 velocity q = D3Q19 velocity q for q < 7)
  template <class S, bool TAB>
  __device__ __forceinline__ static void one(typename S::real (&f)[19], typename S::real (&gg)[7], const Grid &g,
                                             const Streams *sp, const void *fA, const void *gA, uint32_t n) {
    if constexpr (S::LB != 0) {
      // blocked layout: the pull's uniform part is in the stream base, the lane picks one of two offsets of its own
      // cell (pull_voff); out-of-domain sources of face cells at levels 1 and nz - 2 read the allocation's guard
      // bytes and are replaced afterwards
      const void *bf;
      uint32_t o;
      if constexpr (TAB) {
        bf = (const void *)sp->rd[q];
        o = sp->lo[q];
      } else {
        const int64_t off = CX[q] + CY[q] * (int64_t)g.nx + CZ[q] * (int64_t)g.NXY;
        bf = (const void *)((uintptr_t)plane<S, 19>((void *)fA, g.N, q) - (intptr_t)pull_shift(S::LB, S::bytes, off));
        o = pull_rem(S::LB, off);
      }
      const uint32_t voff = pull_voff<S, 19>(n, o);
      f[q] = S::dec_f(S::ld_raw(bf, voff), w19<q>());
      if constexpr (q < 7) {                                    // D3Q7 velocity q = D3Q19 velocity q: same offset
        const void *bg;
        if constexpr (TAB) bg = (const void *)sp->rd[19 + q];
        else bg = (const void *)((uintptr_t)bf + ((uintptr_t)gA - (uintptr_t)fA));
        if constexpr (S::theta_ref) {                          // theta plane [cell], pulled as in [q][cell]
          const int o7 = CX7[q] + CY7[q] * g.nx + CZ7[q] * (int)g.NXY;
          const void *thb;
          if constexpr (TAB) thb = (const void *)sp->rd[26 + q];
          else thb = (const void *)((uintptr_t)gA + g.th_off - (intptr_t)o7 * 4);
          gg[q] = S::dec_g(S::ld_raw(bg, voff), w7q(q), ld_theta(thb, n * 4u));
        } else {
          gg[q] = S::dec_g(S::ld_raw(bg, voff), 0.0f, typename S::real(0));
        }
      }
    } else {   // [q][cell]: the pull offset is folded into each population's base
      const uint32_t boff = n * (uint32_t)S::bytes;
      const void *bf;
      if constexpr (TAB) bf = (const void *)sp->rd[q];
      else bf = shifted<S>(fA, g.N, q, CX[q] + CY[q] * g.nx + CZ[q] * (int)g.NXY);
      f[q] = S::dec_f(S::ld_raw(bf, boff), w19<q>());
      if constexpr (q < 7) {
        const int o7 = CX7[q] + CY7[q] * g.nx + CZ7[q] * (int)g.NXY;
        const void *bg;
        if constexpr (TAB) bg = (const void *)sp->rd[19 + q];
        else bg = shifted<S>(gA, g.N, q, o7);
        const typename S::elem e = S::ld_raw(bg, boff);
        if constexpr (S::theta_ref) {
          const void *thb;                                                  // theta plane, pulled
          if constexpr (TAB) thb = (const void *)sp->rd[26 + q];
          else thb = (const void *)((uintptr_t)gA + g.th_off - (intptr_t)o7 * 4);
          gg[q] = S::dec_g(e, w7q(q), ld_theta(thb, n * 4u));
        } else {
          gg[q] = S::dec_g(e, 0.0f, typename S::real(0));
        }
      }
    }
  }
  template <class S, bool TAB>
  __device__ __forceinline__ static void run(typename S::real (&f)[19], typename S::real (&gg)[7], const Grid &g,
                                             const Streams *sp, const void *fA, const void *gA, uint32_t n) {
    one<S, TAB>(f, gg, g, sp, fA, gA, n);
    FastPull<q + 1>::template run<S, TAB>(f, gg, g, sp, fA, gA, n);
  }
};
template <>
struct FastPull<19> {
  template <class S, bool TAB>
  __device__ __forceinline__ static void run(typename S::real (&)[19], typename S::real (&)[7], const Grid &,
                                             const Streams *, const void *, const void *, uint32_t) {}
};

// Lid cells (k = nz - 1) off the lateral faces whose 3 x 3 neighbourhood is fluid at the lid and the level below
// (k >= kfast): the fast path's pulls, except for the populations whose source lies above the lid, which are never
// issued. A D3Q19 population from above takes the free-slip mirror MIRZ(q) of the lateral neighbour at the same
// level (js/live/kernel.js:168-181); the D3Q7 one keeps the own cell's value (zero gradient, :225-229). These are
// the values the boundary path loads for such a cell.
template <int q>
struct LidPull {
  template <class S, bool TAB>
  __device__ __forceinline__ static void run(typename S::real (&f)[19], typename S::real (&gg)[7], const Grid &g,
                                             const Streams *sp, const void *fA, const void *gA, uint32_t n) {
    using R = typename S::real;
    if constexpr (CZ[q] != -1) {
      FastPull<q>::template one<S, TAB>(f, gg, g, sp, fA, gA, n);
    } else {
      constexpr int m = MIRZ[q];              // c_m = (c_x, c_y, +1)
      const int off = CX[q] + CY[q] * g.nx;   // the lateral neighbour at the same level
      if constexpr (S::LB != 0) {   // (the table's blocked bases fold in m's own pull: formed here instead)
        const uint32_t src = (uint32_t)min(max((int)n - off, 0), (int)g.N - 1);
        f[q] = S::dec_f(S::ld_raw(plane<S, 19>((void *)fA, g.N, m), cell_off<S, 19>(src)), w19<q>());
      } else {
        const void *bf;   // the table's base of m folds in c_z(m) = +1, one plane below; undo that
        if constexpr (TAB) bf = (const void *)(sp->rd[m] + (uint64_t)g.NXY * S::bytes);
        else bf = shifted<S>(fA, g.N, m, off);
        f[q] = S::dec_f(S::ld_raw(bf, n * (uint32_t)S::bytes), w19<q>());
      }
      if constexpr (q < 7) {                  // q = 6: the own cell's D3Q7 population
        if constexpr (S::LB != 0) {
          const void *bg = plane<S, 7>((void *)gA, g.N, q);
          if constexpr (S::theta_ref) {
            const void *thb;   // the table's theta base folds in c_z = -1, one plane above; undo that
            if constexpr (TAB) thb = (const void *)(sp->rd[26 + q] - (uint64_t)g.NXY * 4u);
            else thb = (const void *)((uintptr_t)gA + g.th_off);
            gg[q] = S::dec_g(S::ld_raw(bg, cell_off<S, 7>(n)), w7q(q), ld_theta(thb, n * 4u));
          } else {
            gg[q] = S::dec_g(S::ld_raw(bg, cell_off<S, 7>(n)), 0.0f, R(0));
          }
        } else {
          const void *bg;   // the table's base folds in c_z = -1, one plane above; undo that
          if constexpr (TAB) bg = (const void *)(sp->rd[19 + q] - (uint64_t)g.NXY * S::bytes);
          else bg = shifted<S>(gA, g.N, q, 0);
          const typename S::elem e = S::ld_raw(bg, n * (uint32_t)S::bytes);
          if constexpr (S::theta_ref) {
            const void *thb;
            if constexpr (TAB) thb = (const void *)(sp->rd[26 + q] - (uint64_t)g.NXY * 4u);
            else thb = (const void *)((uintptr_t)gA + g.th_off);
            gg[q] = S::dec_g(e, w7q(q), ld_theta(thb, n * 4u));
          } else {
            gg[q] = S::dec_g(e, 0.0f, R(0));
          }
        }
      }
    }
    LidPull<q + 1>::template run<S, TAB>(f, gg, g, sp, fA, gA, n);
  }
};
template <>
struct LidPull<19> {
  template <class S, bool TAB>
  __device__ __forceinline__ static void run(typename S::real (&)[19], typename S::real (&)[7], const Grid &,
                                             const Streams *, const void *, const void *, uint32_t) {}
};

// Near-solid cells below the lid (ktop <= k < kfast) share the fast path's pulls, which for a solid source cell read
// a value that is then replaced: population q whose source (i - cx, j - cy, k - cz) is solid takes the own cell's
// opp(q) instead (half-way bounce-back, js/live/kernel.js:209-210 and :240-241), exactly the value the boundary path
// loads. r = k - ktop; nb = the column's neighbour word (ktop of the neighbour minus ktop of the column plus one, per
// nibble; clamped at the faces like the boundary path's neighbours). Only lanes that need a replacement load it. On
// a face, populations pulled from outside the domain are replaced again afterwards (FaceFix), as the boundary path
// tests the faces before solidity.
template <int cx, int cy, int cz>
__device__ __forceinline__ bool source_solid(int r, uint32_t nb) {
  if constexpr (cx == 0 && cy == 0) {
    return r - cz < 0;
  } else {
    constexpr int dx = -cx, dy = -cy, n9 = (dy + 1) * 3 + (dx + 1), slot = n9 < 4 ? n9 : n9 - 1;
    return r - cz + 1 < (int)((nb >> (4 * slot)) & 15u);
  }
}

// a loaded element in a register of its own (16-bit elements zero-extended to 32 bits, so that the compiler does not
// pack two of them into one register, which would need the load to have completed)
template <class E>
using Slot = std::conditional_t<sizeof(E) == 2, uint32_t, E>;
template <class E>
__device__ __forceinline__ Slot<E> to_slot(E e) {
  if constexpr (sizeof(E) == 2) return (uint32_t)__builtin_bit_cast(uint16_t, e);
  else return e;
}
template <class E>
__device__ __forceinline__ E from_slot(Slot<E> v) {
  if constexpr (sizeof(E) == 2) return __builtin_bit_cast(E, (uint16_t)v);
  else return v;
}

template <int q>
struct BounceFix {
  // issue the loads of the replaced populations (lanes that need them only) before the first use of any of them, so
  // that a wavefront with several solid sources waits for memory once instead of once per population
  template <class S, class T>
  __device__ __forceinline__ static void load(T (&ef)[19], T (&eg)[7], bool (&sf)[19], bool (&sg)[7], const Grid &g,
                                              const void *fA, const void *gA, uint32_t n, int r, uint32_t nb) {
    sf[q] = source_solid<CX[q], CY[q], CZ[q]>(r, nb);
    if (sf[q]) ef[q] = to_slot(S::ld_raw(plane<S, 19>((void *)fA, g.N, OPP[q]), cell_off<S, 19>(n)));
    if constexpr (q < 7) {
      sg[q] = source_solid<CX7[q], CY7[q], CZ7[q]>(r, nb);
      if (sg[q]) eg[q] = to_slot(S::ld_raw(plane<S, 7>((void *)gA, g.N, OPP7[q]), cell_off<S, 7>(n)));
    }
    BounceFix<q + 1>::template load<S>(ef, eg, sf, sg, g, fA, gA, n, r, nb);
  }
  template <class S, class T>
  __device__ __forceinline__ static void apply(typename S::real (&f)[19], typename S::real (&gg)[7], const T (&ef)[19],
                                               const T (&eg)[7], const bool (&sf)[19], const bool (&sg)[7],
                                               typename S::real th) {
    using E = typename S::elem;
    if (sf[q]) f[q] = S::dec_f(from_slot<E>(ef[q]), w19<q>());
    if constexpr (q < 7) {
      if (sg[q]) {
        if constexpr (S::theta_ref) gg[q] = S::dec_g(from_slot<E>(eg[q]), w7q(q), th);
        else gg[q] = S::dec_g(from_slot<E>(eg[q]), 0.0f, typename S::real(0));
      }
    }
    BounceFix<q + 1>::template apply<S>(f, gg, ef, eg, sf, sg, th);
  }
  // one population at a time: load, wait, decode (fewer values live at once)
  template <class S>
  __device__ __forceinline__ static void run(typename S::real (&f)[19], typename S::real (&gg)[7], const Grid &g,
                                             const void *fA, const void *gA, uint32_t n, int r, uint32_t nb) {
    if (source_solid<CX[q], CY[q], CZ[q]>(r, nb))
      f[q] = S::dec_f(S::ld_raw(plane<S, 19>((void *)fA, g.N, OPP[q]), cell_off<S, 19>(n)), w19<q>());
    if constexpr (q < 7) {
      if (source_solid<CX7[q], CY7[q], CZ7[q]>(r, nb)) {
        const typename S::elem e = S::ld_raw(plane<S, 7>((void *)gA, g.N, OPP7[q]), cell_off<S, 7>(n));
        if constexpr (S::theta_ref) gg[q] = S::dec_g(e, w7q(q), ld_theta(gA, g.th_off + n * 4u));
        else gg[q] = S::dec_g(e, 0.0f, typename S::real(0));
      }
    }
    BounceFix<q + 1>::template run<S>(f, gg, g, fA, gA, n, r, nb);
  }
};
template <>
struct BounceFix<19> {
  template <class S, class T>
  __device__ __forceinline__ static void load(T (&)[19], T (&)[7], bool (&)[19], bool (&)[7], const Grid &,
                                              const void *, const void *, uint32_t, int, uint32_t) {}
  template <class S, class T>
  __device__ __forceinline__ static void apply(typename S::real (&)[19], typename S::real (&)[7], const T (&)[19],
                                               const T (&)[7], const bool (&)[19], const bool (&)[7],
                                               typename S::real) {}  template <class S>
  __device__ __forceinline__ static void run(typename S::real (&)[19], typename S::real (&)[7], const Grid &,
                                             const void *, const void *, uint32_t, int, uint32_t) {}
};

// an empty asm that takes a loaded element in and out: its decoding cannot move above it (into the conditional
// block of its load), so the waits for all the conditional loads before it collapse into one
template <class E>
__device__ __forceinline__ void hold(E &e) {
  static_assert(sizeof(E) == 4 || sizeof(E) == 8, "16-bit elements are held in their 32-bit slot");
  if constexpr (sizeof(E) == 4) {
    uint32_t b = __builtin_bit_cast(uint32_t, e);
    asm volatile("" : "+v"(b));
    e = __builtin_bit_cast(E, b);
  } else {
    uint64_t b = __builtin_bit_cast(uint64_t, e);
    asm volatile("" : "+v"(b));
    e = __builtin_bit_cast(E, b);
  }
}

// TWO_PHASE (the single-domain and the batched step, TAB = true): all replacement loads first, one wait, then the
// decodes. At 600 x 600 x 64 it ran the 16-bit formats 2-3 % faster than replacing one population at a time (which
// takes the step from 80 to 65-69 VGPRs, 6 to 7 waves per SIMD), f32 equally fast; the surface read (TAB = false)
// replaces one at a time
template <class S, bool TWO_PHASE>
__device__ __forceinline__ void bounce_fix(typename S::real (&f)[19], typename S::real (&gg)[7], const Grid &g,
                                           const void *fA, const void *gA, uint32_t n, int r, uint32_t nb) {
  if constexpr (!TWO_PHASE) {
    BounceFix<1>::template run<S>(f, gg, g, fA, gA, n, r, nb);
  } else {
    Slot<typename S::elem> ef[19], eg[7];
    bool sf[19], sg[7];
    BounceFix<1>::template load<S>(ef, eg, sf, sg, g, fA, gA, n, r, nb);
    float th = 0.0f;
    if constexpr (S::theta_ref) {   // the own cell's theta, the reference of its replaced D3Q7 deviations
      if (sg[1] || sg[2] || sg[3] || sg[4] || sg[5] || sg[6]) th = ld_theta(gA, g.th_off + n * 4u);
      hold(th);
    }
#pragma unroll
    for (int q = 1; q < 19; q++) hold(ef[q]);
#pragma unroll
    for (int q = 1; q < 7; q++) hold(eg[q]);
    BounceFix<1>::template apply<S>(f, gg, ef, eg, sf, sg, typename S::real(th));
  }
}

// Face cells (below the lid) also share the fast path's pulls: populations whose source lies outside the lateral
// faces are then replaced by the inflow / outlet value of the boundary path (lateral19, lateral7). The pulled values
// they replace come from in-bounds cells of the wrapped row or the neighbouring plane (k >= 1) and are discarded.
template <int q>
struct FaceFix {
  template <class S>
  __device__ __forceinline__ static void run(typename S::real (&f)[19], typename S::real (&gg)[7], const Grid &g,
                                             const void *gA, uint32_t n, int i, int j, typename S::real uI,
                                             typename S::real vI, const Outflow<typename S::real> &ob) {
    constexpr int cx = CX[q], cy = CY[q];
    if constexpr (cx != 0 || cy != 0) {
      const bool out = (cx == 1 && i == 0) || (cx == -1 && i == g.nx - 1) || (cy == 1 && j == 0) ||
                       (cy == -1 && j == g.ny - 1);
      if (out) f[q] = lateral19<S, q>(g, i, j, uI, vI, ob);
    }
    if constexpr (q < 7 && (CX7[q] != 0 || CY7[q] != 0)) {
      const bool out = (CX7[q] == 1 && i == 0) || (CX7[q] == -1 && i == g.nx - 1) || (CY7[q] == 1 && j == 0) ||
                       (CY7[q] == -1 && j == g.ny - 1);
      if (out) gg[q] = lateral7<S, q>(g, gA, n, i, j);
    }
    FaceFix<q + 1>::template run<S>(f, gg, g, gA, n, i, j, uI, vI, ob);
  }
};
template <>
struct FaceFix<19> {
  template <class S>
  __device__ __forceinline__ static void run(typename S::real (&)[19], typename S::real (&)[7], const Grid &,
                                             const void *, uint32_t, int, int, typename S::real, typename S::real,
                                             const Outflow<typename S::real> &) {}
};

template <class S, int q>
__device__ __forceinline__ typename S::elem enc_f_step(typename S::real v, const Dither &dz) {
  if constexpr (S::SR) return S::enc_f_sr(v, w19<q>(), dz.bits(q));
  else return S::enc_f(v, w19<q>());
}

template <int q>
struct StorePairs {
  template <class S, bool TAB>
  __device__ __forceinline__ static void run(const Streams *sp, void *fB, size_t N, uint32_t boff,
                                             const Moments<typename S::real> &m, const Forces<typename S::real> &F,
                                             const Relax<typename S::real> &R, const Dither &dz) {
    typename S::real a, b;
    collide_pair<q>(m, F, R, a, b);
    S::st_raw(plane_base<S, TAB, 19>(sp, q, fB, N, q), boff, enc_f_step<S, q>(a, dz));
    S::st_raw(plane_base<S, TAB, 19>(sp, q + 1, fB, N, q + 1), boff, enc_f_step<S, q + 1>(b, dz));
    StorePairs<q + 2>::template run<S, TAB>(sp, fB, N, boff, m, F, R, dz);
  }
};
template <>
struct StorePairs<19> {
  template <class S, bool TAB>
  __device__ __forceinline__ static void run(const Streams *, void *, size_t, uint32_t,
                                             const Moments<typename S::real> &, const Forces<typename S::real> &,
                                             const Relax<typename S::real> &, const Dither &) {}
};

__device__ __forceinline__ void load_neighbours(const Grid &g, int i, int j, int kg, int (&kn)[3][3]) {
  const int im = max(i - 1, 0), ip = min(i + 1, g.nx - 1), jm = max(j - 1, 0), jp = min(j + 1, g.ny - 1);
  const int xs[3] = {im, i, ip}, ys[3] = {jm, j, jp};
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++)
      kn[a][b] = (a == 1 && b == 1) ? kg : (int)ld_tab(g.ktop, (size_t)(ys[b] * g.nx + xs[a]));
}

// streaming of one fluid cell: the fast path below the lid, the lid's fast path, or the boundary path
template <class S, bool TAB>
__device__ __forceinline__ void stream_cell(const Grid &g, const Streams *sp, const void *fA, const void *gA,
                                            uint32_t n, int c, int i, int j, int k, int kg, int kfast, bool fast,
                                            bool lid, typename S::real uI, typename S::real vI,
                                            typename S::real (&f)[19], typename S::real (&gg)[7]) {
  using E = typename S::elem;
  if (fast) {
    FastPull<0>::template run<S, TAB>(f, gg, g, sp, fA, gA, n);
    if (k < kfast) bounce_fix<S, TAB>(f, gg, g, fA, gA, n, k - kg, ld_tab(g.colnb, (size_t)c));
    if (i == 0 || i == g.nx - 1 || j == 0 || j == g.ny - 1) {
      const Outflow<typename S::real> ob = outflow_prelude<S>(g, (const E *)fA, g.N, n, i, j);
      FaceFix<1>::template run<S>(f, gg, g, gA, n, i, j, uI, vI, ob);
    }
  } else if (lid) {
    LidPull<0>::template run<S, TAB>(f, gg, g, sp, fA, gA, n);
  } else {
    int kn[3][3];
    load_neighbours(g, i, j, kg, kn);
    const Outflow<typename S::real> ob = outflow_prelude<S>(g, (const E *)fA, g.N, n, i, j);
    Pull19<0>::template run<S>(f, g, (const E *)fA, g.N, n, i, j, k, kn, uI, vI, ob);
    Pull7<0>::template run<S>(gg, g, gA, n, i, j, k, kn);
  }
}

// D3Q7 collision (js/live/kernel.js:369-386): g^post = a g + b (1 + 4 c.u) + w S (1 - sg); src: the cell takes the
// column's rooftop source (k == ktop), of strength cqy * srcscale
template <typename R>
__device__ __forceinline__ void theta_post(const Params &P, bool src, float cqy, const Moments<R> &m,
                                           const Forces<R> &F, const Relax<R> &Rx, const R (&gg)[7], R (&gp)[7]) {
  const R Sq = src ? R(cqy) * R(P.srcscale) : R(0);
  const Scalar<R> sc = scalar_coeffs(m.th, Rx.omg, Sq, F.sg);
  const R b1 = sc.b + sc.cs1, b4 = R(4) * sc.b;
  gp[0] = sc.a * gg[0] + sc.cs0;
  gp[1] = sc.a * gg[1] + (b1 + b4 * F.ux);
  gp[2] = sc.a * gg[2] + (b1 - b4 * F.ux);
  gp[3] = sc.a * gg[3] + (b1 + b4 * F.uy);
  gp[4] = sc.a * gg[4] + (b1 - b4 * F.uy);
  gp[5] = sc.a * gg[5] + (b1 + b4 * F.uz);
  gp[6] = sc.a * gg[6] + (b1 - b4 * F.uz);
}

// collision of one cell from its moments, with the post-collision populations of both lattices stored to pair B
template <class S, bool TAB>
__device__ __forceinline__ void store_post(const Grid &g, const Streams *sp, void *fB, void *gB, uint32_t n, bool src,
                                           float cqy, const Moments<typename S::real> &m,
                                           const Forces<typename S::real> &F, const Relax<typename S::real> &Rx,
                                           const typename S::real (&gg)[7], uint32_t step) {
  using R = typename S::real;
  const uint32_t boff = cell_off<S, 19>(n);
  Dither dz;                                  // read by the stochastic-rounding formats only
  if constexpr (S::SR) dz = Dither(n, step);
  S::st_raw(fB, boff, enc_f_step<S, 0>(collide_rest(Rx), dz));
  StorePairs<1>::template run<S, TAB>(sp, fB, g.N, boff, m, F, Rx, dz);
  R gp[7];
  theta_post(g.P, src, cqy, m, F, Rx, gg, gp);
  store_g<S, TAB>(g, sp, gB, n, gp, dz);
}

// one cell of the fused step (genStepWGSL, js/live/kernel.js:395-431) with FP64 accumulation (ACC, and at run time
// acc_on, which must be uniform over the block) and debug outputs (OUT). Every thread of the block must call it (the
// ACC heat-content sum is a wavefront reduction).
template <class S, bool ACC, bool OUT, bool TAB>
__device__ __forceinline__ double step_cell_body(const Grid &g, const Streams *sp, const void *fA, void *fB,
                                                 const void *gA, void *gB, const Acc &acc, float *__restrict__ out4,
                                                 int c, int k, uint32_t ck, uint32_t step, bool acc_on = true) {
  using R = typename S::real;
  const int ktc = col_kt(ck), kg = col_ktop(ck);
  const bool fluid = c < (int)g.NXY && k >= kg && k < g.nz;
  double th_cell = 0.0;
  if (fluid) {
    const size_t N = g.N;
    const uint32_t n = (uint32_t)k * g.NXY + (uint32_t)c;
    int i, j;
    split_column(g, c, i, j);
    const float4 pr = ld_tab(g.prof, (size_t)max(k - ktc, 0));
    const int kfast = col_kfast(ck);
    const bool fast = (k >= kfast || col_near(ck)) && k < g.nz - 1;
    const bool lid = k == g.nz - 1 && k >= kfast && i > 0 && i < g.nx - 1 && j > 0 && j < g.ny - 1;
    // drag and source act at k == ktop only. The 16-bit formats load them after the pulls are issued, so that the
    // load does not delay the pulls (0.8 % faster per step); the f32 step measured 0.25 % faster with the load first
    float2 cq = make_float2(0.0f, 0.0f);
    if constexpr (S::bytes != 2) {
      if (k == kg) cq = ld_tab(g.colcq, (size_t)c);
    }
    R f[19], gg[7];
    stream_cell<S, TAB>(g, sp, fA, gA, n, c, i, j, k, kg, kfast, fast, lid, R(pr.x), R(pr.y), f, gg);
    if constexpr (S::bytes == 2) {
      if (k == kg) cq = ld_tab(g.colcq, (size_t)c);
    }
    const Moments<R> m = moments(f, gg);
    const Forces<R> F = forces<true, true>(g.P, g.rp, g.nz, k, kg, ktc, R(cq.x), R(pr.x), R(pr.y), m);
    const Relax<R> Rx = relax(g.P, g.rp, g.nx, g.ny, i, j, R(pr.z), m, F);
    store_post<S, TAB>(g, sp, fB, gB, n, k == kg, cq.y, m, F, Rx, gg, step);
    if (OUT) {
      out4[n] = (float)m.th;
      out4[N + n] = (float)F.ux;
      out4[2 * N + n] = (float)F.uy;
      out4[3 * N + n] = (float)F.uz;
    }
    if (ACC && acc_on) {
      th_cell = (double)m.th;
      if (k == kg) {
        acc.surf[c] += (double)m.th;
        acc.surf[g.NXY + c] += (double)F.ux;
        acc.surf[2 * (size_t)g.NXY + c] += (double)F.uy;
      }
      const int nx = g.nx, ny = g.ny, nz = g.nz;
      const size_t lw = (size_t)nz * ny, ls = (size_t)nz * nx;
      if (i == 0) {
        double *fw = acc.face;
        fw[(size_t)k * ny + j] += (double)m.th;
        fw[lw + (size_t)k * ny + j] += (double)F.ux;
      }
      if (i == nx - 1) {
        double *fe = acc.face + 2 * lw;
        fe[(size_t)k * ny + j] += (double)m.th;
        fe[lw + (size_t)k * ny + j] += (double)F.ux;
      }
      if (j == 0) {
        double *fs = acc.face + 4 * lw;
        fs[(size_t)k * nx + i] += (double)m.th;
        fs[ls + (size_t)k * nx + i] += (double)F.uy;
      }
      if (j == ny - 1) {
        double *fn = acc.face + 4 * lw + 2 * ls;
        fn[(size_t)k * nx + i] += (double)m.th;
        fn[ls + (size_t)k * nx + i] += (double)F.uy;
      }
    }
  }
  return th_cell;
}

// Heat content: one FP64 slot per (level, group of 64 consecutive columns), i.e. per wavefront of any launch shape
// (blocks are whole multiples of 64 columns), each slot updated by exactly one wavefront per step: deterministic, no
// atomics, and the same slots whatever the block size or levels per thread, so the summed content does not depend on
// the launch shape (a batch member gives the bits of the same domain stepped alone). k and c0 (the wavefront's first
// column) are wave-uniform; a wavefront entirely beyond the last column or level writes nothing.
__device__ __forceinline__ uint32_t content_slot(const Grid &g, int k, int c0) {
  return (uint32_t)k * ((g.NXY + 63u) >> 6) + ((uint32_t)c0 >> 6);
}
__device__ __forceinline__ void content_add(const Grid &g, const Acc &acc, int k, int c0, double th) {
  if (k >= g.nz || c0 >= (int)g.NXY) return;
  const double sum = wave_sum(th);
  if ((threadIdx.x & 63) == 0) acc.content[content_slot(g, k, c0)] += sum;
}

// one block per (column block, KB-level chunk) tile, numbered column block fastest, then level chunk, taken in the
// order g.win selects; a thread walks its column through the KB levels (column word and index math once per column).
// Plane slabs (sw > 0, g.win == WIN_PLANE_SLABS): the column blocks of a level chunk are cut into eight slabs of sw
// blocks, XCD x stepping slab x of every level chunk in turn, so the eight XCDs move through the levels together and
// the tiles that pull from one another across levels (the same columns one level up or down) meet in one XCD's L2;
// blocks past the last column block (at most 7 per level chunk) exit.
template <class S, bool ACC, bool OUT, int KB>
__global__ __launch_bounds__(256) void step_kernel(const Grid *__restrict__ gp, const Streams *__restrict__ sp,
                                                   const void *fA, void *fB, const void *gA, void *gB, Acc acc,
                                                   float *__restrict__ out4, uint32_t bx, float inv_bx, uint32_t sw,
                                                   float inv_sw, uint32_t step) {
  const Grid &g = *gp;
  uint32_t cb, kk;
  if (sw) {
    uint32_t r;
    kk = div_by(blockIdx.x >> 3, sw, inv_sw, r);
    cb = (blockIdx.x & (NUM_XCD - 1)) * sw + r;
    if (cb >= bx) return;
  } else {
    kk = div_by(tile_of_block(blockIdx.x, gridDim.x, g.win), bx, inv_bx, cb);
  }
  const int c = (int)(cb * blockDim.x + threadIdx.x);
  const int c0 = c - (int)(threadIdx.x & 63);
  const int k0 = g.kmin + (int)kk * KB;
  const uint32_t ck = ld_tab(g.colk, (size_t)(c < (int)g.NXY ? c : 0));
  for (int u = 0; u < KB; u++) {
    const double th = step_cell_body<S, ACC, OUT, true>(g, sp, fA, fB, gA, gB, acc, out4, c, k0 + u, ck, step);
    if (ACC) content_add(g, acc, k0 + u, c0, th);
  }
}

// ---- two steps per launch: a z sweep with LDS-staged halo planes -----------------------------------------------
// A workgroup of 256 threads owns a region of TX x TY columns, one column per thread, and walks it up through the
// levels of a chunk. For each level it runs two phases:
//   A: the first step (t -> t+1) of plane k+1 for every fluid cell of the region, streamed from lattice pair A exactly
//      as the single step streams it. The post-collision populations stay on chip: the 20 that a lateral neighbour
//      pulls go to LDS, grouped by the role the plane plays for the second step (12 pulled within the plane, 4 pulled
//      upward from it, 4 downward), and the thread keeps the other 10 of its own cell in registers;
//   B: the second step (t+1 -> t+2) of plane k for the (TX - 2) x (TY - 2) interior columns. Every pull reads LDS
//      (lateral sources, and the own cell's values that live there) or the thread's registers, under the boundary
//      path's rules for every cell (lid mirror, lateral faces, half-way bounce-back); collision and stores are the
//      single step's, into lattice pair B.
// The pulls of the next plane's first step are issued before phase B, so that their latency overlaps its arithmetic.
// The intermediate state never reaches HBM (208 B per cell for two steps instead of 416, f32). The price is
// recomputation: the halo ring of a region is also stepped by the neighbouring workgroup, and a chunk also steps the
// plane below and the plane above it. The results are bit-identical to two single steps (same expressions, same
// contraction, same pulled values). f32 pair layout only; neither FP64 means nor debug outputs (such steps take
// the single-step kernel).
constexpr int SWEEP_THREADS = 256;
// slots of the lateral populations in the LDS planes (-1: not lateral)
__device__ constexpr int SW_FLAT[19] = {-1, 0, 1, 2, 3, -1, -1, 4, 5, 6, 7, -1, -1, -1, -1, -1, -1, -1, -1};
__device__ constexpr int SW_UP[19] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 0, -1, -1, 1, 2, -1, -1, 3};
__device__ constexpr int SW_DOWN[19] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 0, 1, -1, -1, 2, 3, -1};
// the own cell's D3Q19 populations that are kept in registers (the others live in the plane's LDS slots)
__device__ constexpr int SW_REG[19] = {0, -1, -1, -1, -1, 1, 2, -1, -1, -1, -1, -1, 3, 4, -1, -1, 5, 6, -1};

// plane rings: flat[p & 1] holds plane p's in-plane populations (8 D3Q19, then D3Q7 1..4), up[p % 3] the four D3Q19
// populations with c_z = +1 and a lateral component, down the four with c_z = -1 of the newest plane only
struct SweepLds {
  float flat[2][12][SWEEP_THREADS];
  float up[3][4][SWEEP_THREADS];
  float down[4][SWEEP_THREADS];
};

// the own cell's post-collision populations that stay in registers: D3Q19 0, 5, 6, 12, 13, 16, 17 and D3Q7 0, 5, 6
struct SweepOwn {
  float f[7], g[3];
};

// where the own cell's populations of plane p live: its LDS slots (sf = p & 1, su = p % 3) and its registers
struct OwnAt {
  const SweepLds &L;
  int t, sf, su;
  const SweepOwn &r;
  template <int q>
  __device__ __forceinline__ float f() const {
    if constexpr (SW_REG[q] >= 0) return r.f[SW_REG[q]];
    else if constexpr (SW_FLAT[q] >= 0) return L.flat[sf][SW_FLAT[q]][t];
    else return L.up[su][SW_UP[q]][t];
  }
  template <int q>
  __device__ __forceinline__ float g() const {
    if constexpr (q == 0) return r.g[0];
    else if constexpr (q >= 5) return r.g[q - 4];
    else return L.flat[sf][8 + q - 1][t];
  }
};

template <int q>
struct CollidePairs {
  __device__ __forceinline__ static void run(const Moments<float> &m, const Forces<float> &F, const Relax<float> &Rx,
                                             float (&fp)[19]) {
    collide_pair<q>(m, F, Rx, fp[q], fp[q + 1]);
    CollidePairs<q + 2>::run(m, F, Rx, fp);
  }
};
template <>
struct CollidePairs<19> {
  __device__ __forceinline__ static void run(const Moments<float> &, const Forces<float> &, const Relax<float> &,
                                             float (&)[19]) {}
};

template <int q>
struct OwnAll {   // the own cell's 19 D3Q19 populations (outlet faces)
  __device__ __forceinline__ static void run(const OwnAt &o, float (&a)[19]) {
    a[q] = o.template f<q>();
    OwnAll<q + 1>::run(o, a);
  }
};
template <>
struct OwnAll<19> {
  __device__ __forceinline__ static void run(const OwnAt &, float (&)[19]) {}
};

// post-stream population q of the second step at cell (i, j, k): the rules of pull19, the sources taken from the
// sweep's planes (own = plane k of this column, f6n = population 6 of plane k + 1, f5m = population 5 of plane k - 1)
template <int TX, int q>
__device__ __forceinline__ float sweep_pull19(const Grid &g, const OwnAt &own, int i, int j, int k,
                                              const int (&kn)[3][3], float uI, float vI, const Outflow<float> &ob,
                                              int su_km1, float f6n, float f5m) {
  constexpr int cx = CX[q], cy = CY[q], cz = CZ[q];
  const int NX = g.nx, NY = g.ny;
  const SweepLds &L = own.L;
  if constexpr (cz == -1) {
    if (k == g.nz - 1) {
      constexpr int m = MIRZ[q];
      const int si = i - cx, sj = j - cy;
      if (si < 0 || si >= NX || sj < 0 || sj >= NY) return own.f<m>();
      if (k < kn[1 - cx][1 - cy]) return own.f<m>();
      if constexpr (cx == 0 && cy == 0) return own.f<m>();
      else return L.up[own.su][SW_UP[m]][own.t - cx - cy * TX];
    }
  }
  const bool outx = (cx == 1 && i == 0) || (cx == -1 && i == NX - 1);
  const bool outy = (cy == 1 && j == 0) || (cy == -1 && j == NY - 1);
  if (outx || outy) return lateral19<StoreF32, q>(g, i, j, uI, vI, ob);
  const int sk = k - cz;
  if (sk < kn[1 - cx][1 - cy]) return own.f<OPP[q]>();
  if constexpr (cx == 0 && cy == 0) {
    if constexpr (cz == 0) return own.f<q>();
    else if constexpr (cz == 1) return f5m;        // q = 5 from the level below
    else return f6n;                               // q = 6 from the level above
  } else {
    const int tn = own.t - cx - cy * TX;
    if constexpr (cz == 0) return L.flat[own.sf][SW_FLAT[q]][tn];
    else if constexpr (cz == 1) return L.up[su_km1][SW_UP[q]][tn];
    else return L.down[SW_DOWN[q]][tn];
  }
}

// post-stream theta population q of the second step (the rules of pull7)
template <int TX, int q>
__device__ __forceinline__ float sweep_pull7(const Grid &g, const OwnAt &own, int i, int j, int k,
                                             const int (&kn)[3][3], float g6n, float g5m) {
  constexpr int cx = CX7[q], cy = CY7[q], cz = CZ7[q];
  const Params &P = g.P;
  if (cz == -1 && k == g.nz - 1) return own.g<q>();
  const bool outx = (cx == 1 && i == 0) || (cx == -1 && i == g.nx - 1);
  const bool outy = (cy == 1 && j == 0) || (cy == -1 && j == g.ny - 1);
  if (outx || outy) {
    const bool inflow = (cx == 1 && i == 0 && P.inflow_w != 0) || (cx == -1 && i == g.nx - 1 && P.inflow_e != 0) ||
                        (cy == 1 && j == 0 && P.inflow_s != 0) || (cy == -1 && j == g.ny - 1 && P.inflow_n != 0);
    return inflow ? 0.0f : own.g<q>();
  }
  const int sk = k - cz;
  if (sk < kn[1 - cx][1 - cy]) return own.g<OPP7[q]>();
  if constexpr (cx == 0 && cy == 0) {
    if constexpr (cz == 0) return own.g<q>();
    else if constexpr (cz == 1) return g5m;
    else return g6n;
  } else {
    return own.L.flat[own.sf][8 + q - 1][own.t - cx - cy * TX];
  }
}

template <int TX, int q>
struct SweepPull {
  __device__ __forceinline__ static void run(float (&f)[19], float (&gg)[7], const Grid &g, const OwnAt &own, int i,
                                             int j, int k, const int (&kn)[3][3], float uI, float vI,
                                             const Outflow<float> &ob, int su_km1, const SweepOwn &nxt, float f5m,
                                             float g5m) {
    f[q] = sweep_pull19<TX, q>(g, own, i, j, k, kn, uI, vI, ob, su_km1, nxt.f[2], f5m);
    if constexpr (q < 7) gg[q] = sweep_pull7<TX, q>(g, own, i, j, k, kn, nxt.g[2], g5m);
    SweepPull<TX, q + 1>::run(f, gg, g, own, i, j, k, kn, uI, vI, ob, su_km1, nxt, f5m, g5m);
  }
};
template <int TX>
struct SweepPull<TX, 19> {
  __device__ __forceinline__ static void run(float (&)[19], float (&)[7], const Grid &, const OwnAt &, int, int, int,
                                             const int (&)[3][3], float, float, const Outflow<float> &, int,
                                             const SweepOwn &, float, float) {}
};

// a wave-uniform pointer the compiler must treat as new at this point: the population bases formed from it are then
// formed again in every loop iteration (a few scalar adds) instead of being kept live, and spilled, across the loop
template <class P>
__device__ __forceinline__ P *fresh(P *p) {
  uint64_t v = (uint64_t)(uintptr_t)p;
  asm volatile("" : "+s"(v));
  return (P *)(uintptr_t)v;
}

// the first step's streaming split in two (the values of stream_cell): the plain pulls of the fast and lid paths,
// issued early, and what depends on them or on solid neighbours (bounce-back and face fix-ups, the boundary path)
template <class S>
__device__ __forceinline__ void sweep_issue(const Grid &g, const void *fA, const void *gA, uint32_t n, bool fast,
                                            bool lid, float (&f)[19], float (&gg)[7]) {
  if (fast) FastPull<0>::template run<S, false>(f, gg, g, nullptr, fA, gA, n);
  else if (lid) LidPull<0>::template run<S, false>(f, gg, g, nullptr, fA, gA, n);
}
template <class S>
__device__ __forceinline__ void sweep_finish(const Grid &g, const void *fA, const void *gA, uint32_t n, int c, int i,
                                             int j, int k, int kg, int kfast, bool fast, bool lid, float uI, float vI,
                                             float (&f)[19], float (&gg)[7]) {
  using E = typename S::elem;
  if (fast) {
    if (k < kfast) bounce_fix<S, false>(f, gg, g, fA, gA, n, k - kg, ld_tab(g.colnb, (size_t)c));
    if (i == 0 || i == g.nx - 1 || j == 0 || j == g.ny - 1) {
      const Outflow<float> ob = outflow_prelude<S>(g, (const E *)fA, g.N, n, i, j);
      FaceFix<1>::template run<S>(f, gg, g, gA, n, i, j, uI, vI, ob);
    }
  } else if (!lid) {
    int kn[3][3];
    load_neighbours(g, i, j, kg, kn);
    const Outflow<float> ob = outflow_prelude<S>(g, (const E *)fA, g.N, n, i, j);
    Pull19<0>::template run<S>(f, g, (const E *)fA, g.N, n, i, j, k, kn, uI, vI, ob);
    Pull7<0>::template run<S>(gg, g, gA, n, i, j, k, kn);
  }
}

// Blocks: (tile, level chunk), tile = tx + tiles_x * ty fastest, taken in the single step's XCD-aware order (win).
// Region of tile (tx, ty): columns i = tx (TX - 2) - 1 + rx, j = ty (TY - 2) - 1 + ry, thread t = ry TX + rx.
// Population bases are formed from the lattice pointers (blocked layout: one scalar add each), not read from the
// stream table, whose 60 bases would stay live across the level loop in scalar registers.
template <class S, int TX>
__global__ __launch_bounds__(SWEEP_THREADS) void sweep2_kernel(const Grid *__restrict__ gp, const void *fA, void *fB,
                                                               const void *gA, void *gB, uint32_t tiles_x,
                                                               uint32_t tiles_xy, int kc, uint32_t win,
                                                               uint32_t step) {
  static_assert(S::bytes == 4 && !S::theta_ref && !S::SR && S::LB != 0, "the sweep runs the f32 pair layout");
  constexpr int TY = SWEEP_THREADS / TX;
  __shared__ SweepLds L;
  const Grid &g = *gp;
  const int t = (int)threadIdx.x;
  const uint32_t b = tile_of_block(blockIdx.x, gridDim.x, win);
  const uint32_t chunk = b / tiles_xy, tile = b - chunk * tiles_xy, ty = tile / tiles_x, tx = tile - ty * tiles_x;
  const int rx = t % TX, ry = t / TX;
  const int i = (int)tx * (TX - 2) - 1 + rx, j = (int)ty * (TY - 2) - 1 + ry;
  const bool indom = i >= 0 && i < g.nx && j >= 0 && j < g.ny;
  const bool inner = indom && rx >= 1 && rx <= TX - 2 && ry >= 1 && ry <= TY - 2;
  const int c = indom ? j * g.nx + i : 0;
  const uint32_t ck = ld_tab(g.colk, (size_t)c);
  const int ktc = col_kt(ck), kg = indom ? col_ktop(ck) : g.nz, kfast = col_kfast(ck);
  const bool near = col_near(ck);
  float2 cq = make_float2(0.0f, 0.0f);
  if (kg < g.nz) cq = ld_tab(g.colcq, (size_t)c);
  int kn[3][3];
  if (inner) load_neighbours(g, i, j, kg, kn);
  const int kb0 = g.kmin + (int)chunk * kc, kb1 = min(kb0 + kc, g.nz);
  auto is_fast = [&](int ka) { return (ka >= kfast || near) && ka < g.nz - 1; };
  auto is_lid = [&](int ka) {
    return ka == g.nz - 1 && ka >= kfast && i > 0 && i < g.nx - 1 && j > 0 && j < g.ny - 1;
  };
  SweepOwn cur, nxt;
  float f5m = 0.0f, g5m = 0.0f;
  float fa[19], ga[7];   // pulls of the pending first-step plane
  {
    const int ka = kb0 - 1;
    if (ka >= kg && ka < g.nz)
      sweep_issue<S>(g, fresh(fA), fresh(gA), (uint32_t)ka * g.NXY + (uint32_t)c, is_fast(ka), is_lid(ka), fa, ga);
  }
  for (int ka = kb0 - 1; ka <= kb1; ka++) {
    // phase A: the first step of plane ka, from the pulls issued one iteration earlier
    if (ka >= kg && ka < g.nz) {
      const uint32_t n = (uint32_t)ka * g.NXY + (uint32_t)c;
      const float4 pr = ld_tab(g.prof, (size_t)max(ka - ktc, 0));
      sweep_finish<S>(g, fresh(fA), fresh(gA), n, c, i, j, ka, kg, kfast, is_fast(ka), is_lid(ka), pr.x, pr.y, fa, ga);
      const Moments<float> m = moments(fa, ga);
      const Forces<float> F = forces<true, true>(g.P, g.rp, g.nz, ka, kg, ktc, ka == kg ? cq.x : 0.0f, pr.x, pr.y, m);
      const Relax<float> Rx = relax(g.P, g.rp, g.nx, g.ny, i, j, pr.z, m, F);
      float fp[19], gp7[7];
      fp[0] = collide_rest(Rx);
      CollidePairs<1>::run(m, F, Rx, fp);
      theta_post(g.P, ka == kg, cq.y, m, F, Rx, ga, gp7);
      const int sf = ka & 1, su = ka % 3;
#pragma unroll
      for (int q = 0; q < 19; q++) {
        if (SW_FLAT[q] >= 0) L.flat[sf][SW_FLAT[q]][t] = fp[q];
        if (SW_UP[q] >= 0) L.up[su][SW_UP[q]][t] = fp[q];
        if (SW_DOWN[q] >= 0) L.down[SW_DOWN[q]][t] = fp[q];
        if (SW_REG[q] >= 0) nxt.f[SW_REG[q]] = fp[q];
      }
#pragma unroll
      for (int q = 1; q < 5; q++) L.flat[sf][8 + q - 1][t] = gp7[q];
      nxt.g[0] = gp7[0];
      nxt.g[1] = gp7[5];
      nxt.g[2] = gp7[6];
    }
    __syncthreads();
    {   // the pulls of the next first-step plane, in flight during phase B
      const int kn1 = ka + 1;
      if (kn1 <= kb1 && kn1 >= kg && kn1 < g.nz)
        sweep_issue<S>(g, fresh(fA), fresh(gA), (uint32_t)kn1 * g.NXY + (uint32_t)c, is_fast(kn1), is_lid(kn1), fa,
                       ga);
    }
    // phase B: the second step of plane k = ka - 1
    const int k = ka - 1;
    if (inner && k >= kb0 && k < kb1 && k >= kg) {
      const uint32_t n = (uint32_t)k * g.NXY + (uint32_t)c;
      const float4 pr = ld_tab(g.prof, (size_t)max(k - ktc, 0));
      const Params &P = g.P;
      const OwnAt own{L, t, k & 1, k % 3, cur};
      Outflow<float> ob;
      const bool onb = (i == 0 && P.inflow_w == 0) || (i == g.nx - 1 && P.inflow_e == 0) ||
                       (j == 0 && P.inflow_s == 0) || (j == g.ny - 1 && P.inflow_n == 0);
      if (onb) OwnAll<0>::run(own, ob.own);
      outflow_finish(g, ob, onb, i, j);
      float f[19], gg[7];
      SweepPull<TX, 0>::run(f, gg, g, own, i, j, k, kn, pr.x, pr.y, ob, (k + 2) % 3, nxt, f5m, g5m);
      const Moments<float> m = moments(f, gg);
      const Forces<float> F = forces<true, true>(g.P, g.rp, g.nz, k, kg, ktc, k == kg ? cq.x : 0.0f, pr.x, pr.y, m);
      const Relax<float> Rx = relax(g.P, g.rp, g.nx, g.ny, i, j, pr.z, m, F);
      store_post<S, false>(g, nullptr, fresh(fB), fresh(gB), n, k == kg, cq.y, m, F, Rx, gg, step + 1u);
    }
    __syncthreads();
    f5m = cur.f[1];
    g5m = cur.g[1];
    cur = nxt;
  }
}

// ---- many domains in one launch ----------------------------------------------------------------------------------
// One grid of G blocks of BS threads walks one step of the batch members that step in this launch. Member d
// contributes tiles (BS-column block, one level) for its fluid levels, concatenated in table order. The
// host orders the table by step count, longest first, so the members stepping in launch `it` of a chunk are a prefix
// of it and the launch covers only their tiles (nmem, total): a member that sits out a launch costs no workgroups.
// The grid strides through the list in windows of G tiles, taking the tiles of a window in the XCD-aware order of the
// single-domain step (runs of 64 tiles per XCD); with one block per tile (G = total, the default) block b takes tile
// b. Per-member fields make a member's result independent of the mix: it reads lattice parity0 ^ (it & 1) through the
// stream table of that parity (M.s, host-built per chunk) and accumulates means from launch acc_from on. BS (64, 128,
// 256) and the grid are the batch tuner's choice.
// One copy of the cell body with the accumulation switched at run time, one level per thread, and (PERSIST = false)
// no loop at all: a compile-time ACC / non-ACC pair of copies in one kernel, or loops over a persistent grid's tiles and
// a member's levels, raised the kernel from the single step's 80-84 VGPRs to 90-126 and spilled 240-990 SGPRs into
// VGPR lanes (f16s at 600 x 600 x 64 batched alone: 0.66-0.81 ms against 0.45 ms now).
template <class S>
__device__ __forceinline__ void batch_tile(const Member &M, int p, int c, int k, uint32_t step, bool acc) {
  const Grid &g = M.g;
  const uint32_t ck = ld_tab(g.colk, (size_t)(c < (int)g.NXY ? c : 0));
  const double th = step_cell_body<S, true, false, true>(g, &M.s[p], M.f[p], M.f[1 - p], M.gq[p], M.gq[1 - p], M.acc,
                                                         nullptr, c, k, ck, step, acc);
  if (acc) content_add(g, M.acc, k, c - (int)(threadIdx.x & 63), th);
}

// the member of tile t: binary search over tile_base (ascending)
__device__ __forceinline__ int member_of(const Member *__restrict__ mem, int nmem, uint32_t t) {
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mem[mid].tile_base <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// PERSIST = false: one block per tile; PERSIST = true: a grid of G < total blocks walking the tiles
template <class S, int BS, bool PERSIST>
__global__ __launch_bounds__(BS) void batch_step_kernel(const Member *__restrict__ mem, int nmem, uint32_t total,
                                                        int it) {
  if constexpr (!PERSIST) {
    const uint32_t t = blockIdx.x;
    if (t >= total) return;
    const Member &M = mem[member_of(mem, nmem, t)];
    if (it >= M.nsteps) return;                                  // not reached: the launch covers stepping members
    uint32_t cb;
    const uint32_t kk = div_by(t - M.tile_base, M.bx, M.inv_bx, cb);
    batch_tile<S>(M, M.parity0 ^ (it & 1), (int)(cb * BS + threadIdx.x), M.g.kmin + (int)kk, M.step0 + (uint32_t)it,
                  it >= M.acc_from);
  } else {
    const uint32_t G = gridDim.x, tb = tile_of_block(blockIdx.x, G, 64);
    // member of the block's first tile, then advanced as t grows
    int d = tb < total ? member_of(mem, nmem, tb) : 0;
    for (uint32_t base = 0; base < total; base += G) {
      const uint32_t t = base + tb;
      if (t >= total) continue;                                   // partial last window
      while (d + 1 < nmem && mem[d + 1].tile_base <= t) d++;      // t grows with base: the member index only advances
      const Member &M = mem[d];
      if (it >= M.nsteps) continue;                               // not reached: the launch covers stepping members
      uint32_t cb;
      const uint32_t kk = div_by(t - M.tile_base, M.bx, M.inv_bx, cb);
      batch_tile<S>(M, M.parity0 ^ (it & 1), (int)(cb * BS + threadIdx.x), M.g.kmin + (int)kk,
                    M.step0 + (uint32_t)it, it >= M.acc_from);
    }
  }
}

// near-surface read (genSurfWGSL, js/live/kernel.js:435-471): out3 = [3][NXY] north-up theta, u*conv, v*conv
template <class S>
__global__ __launch_bounds__(256) void surface_kernel(Grid g, const void *fA, const void *gA, float conv,
                                                      float *__restrict__ out3) {
  using R = typename S::real;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int)g.NXY) return;
  int i, j;
  split_column(g, c, i, j);
  const uint32_t ck = g.colk[c];
  const int ktc = col_kt(ck), kg = col_ktop(ck), k = kg;
  const uint32_t o = (uint32_t)(g.ny - 1 - j) * g.nx + i;
  if (k >= g.nz) {
    out3[o] = 0.0f; out3[g.NXY + o] = 0.0f; out3[2 * g.NXY + o] = 0.0f;
    return;
  }
  const uint32_t n = (uint32_t)k * g.NXY + (uint32_t)c;
  const float4 pr = g.prof[max(k - ktc, 0)];
  R f[19], gg[7];
  stream_cell<S, false>(g, nullptr, fA, gA, n, c, i, j, k, kg, 0, false, false, R(pr.x), R(pr.y), f, gg);
  const Moments<R> m = moments(f, gg);
  const Forces<R> F = forces<true, true>(g.P, g.rp, g.nz, k, kg, ktc, R(g.colcq[c].x), R(pr.x), R(pr.y), m);
  out3[o] = (float)m.th;
  out3[g.NXY + o] = (float)(F.ux * R(conv));
  out3[2 * g.NXY + o] = (float)(F.uy * R(conv));
}

// equilibrium of the inflow profile at the local height above terrain, both lattices (genInitWGSL :524-553),
// evaluated operation by operation as the reference (no contraction): the f32 state is bit-identical to the closed form
template <class S>
__global__ __launch_bounds__(256) void init_kernel(Grid g, void *f0, void *f1) {
#pragma clang fp contract(off)
  using R = typename S::real;
  using E = typename S::elem;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (c >= (int)g.NXY) return;
  const int ktc = col_kt(g.colk[c]);
  const int zagl = min(max(k - ktc, 0), g.nz - 1);
  const float4 pr = g.prof[zagl];
  const R U = R(pr.x), V = R(pr.y), usq = U * U + V * V;
  const size_t N = g.N;
  const uint32_t n = (uint32_t)k * g.NXY + (uint32_t)c;
#pragma unroll
  for (int q = 0; q < 19; q++) {
    const R cu = R(CX[q]) * U + R(CY[q]) * V;
    const float wf = q == 0 ? (float)(1.0 / 3.0) : (q < 7 ? (float)(1.0 / 18.0) : (float)(1.0 / 36.0));
    const R fq = R(wf) * (((R(1) + R(3) * cu) + R(4.5) * cu * cu) - R(1.5) * usq);
    const E e = S::enc_f(fq, wf);
    ((E *)f0)[lat_index<S>(19, N, q, n)] = e;
    ((E *)f1)[lat_index<S>(19, N, q, n)] = e;
  }
}

// cells that turned from solid to fluid: rest state f = w_q, g = 0 (genReinitWGSL :557-575)
template <class S>
__global__ __launch_bounds__(256) void reinit_kernel(Grid g, const uint32_t *__restrict__ oldk, void *f, void *gq) {
  using R = typename S::real;
  using E = typename S::elem;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int)g.NXY) return;
  const int k1 = min((int)oldk[c], g.nz);
  const size_t N = g.N;
  for (int k = (int)g.ktop[c]; k < k1; k++) {
    const uint32_t n = (uint32_t)k * g.NXY + (uint32_t)c;
#pragma unroll
    for (int q = 0; q < 19; q++) {
      const float wf = q == 0 ? (float)(1.0 / 3.0) : (q < 7 ? (float)(1.0 / 18.0) : (float)(1.0 / 36.0));
      ((E *)f)[lat_index<S>(19, N, q, n)] = S::enc_f(R(wf), wf);
    }
#pragma unroll
    for (int q = 0; q < 7; q++) ((E *)gq)[lat_index<S>(7, N, q, n)] = S::enc_g(R(0), w7q(q), R(0));
    if constexpr (S::theta_ref) ((float *)((char *)gq + g.th_off))[n] = 0.0f;
  }
}

// u8 theta volume, north-up [k][row][col], 4 cells per u32, and max theta (genVolWGSL :475-520)
template <class S>
__global__ __launch_bounds__(256) void volume_kernel(Grid g, const void *gA, int nzv, float volscale,
                                                     uint32_t *__restrict__ vout, uint32_t *__restrict__ tmax_bits) {
  using R = typename S::real;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nvol = nzv * (int)g.NXY;
  const int nw = (nvol + 3) / 4;
  float tmax = 0.0f;
  if (t < nw) {
    uint32_t packed = 0u;
    for (int e = 0; e < 4; e++) {
      const int p = 4 * t + e;
      if (p < nvol) {
        const int k = p / (int)g.NXY, rem = p % (int)g.NXY;
        const int i = rem % g.nx, j = g.ny - 1 - rem / g.nx;
        const int c = j * g.nx + i;
        const int kg = (int)g.ktop[c];
        float th = 0.0f;
        if (k >= kg) {
          int kn[3][3];
          load_neighbours(g, i, j, kg, kn);
          R gg[7];
          Pull7<0>::template run<S>(gg, g, gA, (uint32_t)k * g.NXY + (uint32_t)c, i, j, k, kn);
          th = (float)(gg[0] + gg[1] + gg[2] + gg[3] + gg[4] + gg[5] + gg[6]);
        }
        const uint32_t qv = (uint32_t)fminf(fmaxf(rintf(th * volscale), 0.0f), 255.0f);
        packed |= qv << (8u * (uint32_t)e);
        tmax = fmaxf(tmax, th);
      }
    }
    vout[t] = packed;
  }
  // max over the block, then one atomic per block (max is order-independent, so the result is deterministic)
  __shared__ float red[256];
  red[threadIdx.x] = fmaxf(tmax, 0.0f);
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(tmax_bits, __float_as_uint(red[0]));
}

// diagnostics over the stored state (genStatsWGSL :580-633) in FP64, two levels, deterministic: a grid of at most
// STATS_MAX_BLOCKS blocks strides over the cells (a thread takes cells t, t + G * 256, ... in order), each block
// reduces its threads and writes its partials part[s * G + block] for s = 0..6 (sum theta, sum ke, sum (rho - 1),
// fluid count, max theta, max |u|, non-finite count); then one block per statistic reduces the G partials in a
// fixed order (stats_finish). The partials take 7 * G doubles whatever the grid size.
template <class S>
__global__ __launch_bounds__(256) void stats_kernel(Grid g, const void *fA, const void *gA, double *__restrict__ part) {
  using R = typename S::real;
  using E = typename S::elem;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < g.N; n += stride) {
    const uint32_t c = n % g.NXY, k = n / g.NXY;
    if ((int)k < (int)g.ktop[c]) continue;
    const size_t N = g.N;
    R a[19];
    OwnPops<0>::template run<S>(a, (const E *)fA, N, n);
    R rho = a[0], jx = R(0), jy = R(0), jz = R(0);
#pragma unroll
    for (int q = 1; q < 19; q++) {
      rho += a[q];
      jx += R(CX[q]) * a[q];
      jy += R(CY[q]) * a[q];
      jz += R(CZ[q]) * a[q];
    }
    R th = R(0);
#pragma unroll
    for (int q = 0; q < 7; q++) th += ldg<S>(g, gA, q, n);
    if (!isfinite((double)rho) || !isfinite((double)th)) {
      v[6] += 1.0;
    } else {
      const double jj = (double)jx * jx + (double)jy * jy + (double)jz * jz;
      v[0] += (double)th;
      v[1] += 0.5 * jj / (double)rho;
      v[2] += (double)rho - 1.0;
      v[3] += 1.0;
      v[4] = fmax(v[4], (double)th);
      v[5] = fmax(v[5], sqrt(jj) / (double)rho);
    }
  }
  __shared__ double red[7][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < 7; s++) {
    double x = v[s];
    if (s == 4 || s == 5) {
      for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
    } else {
      x = wave_sum(x);
    }
    if (lane == 0) red[s][wv] = x;
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    const int s = threadIdx.x;
    const int nwv = blockDim.x >> 6;
    double x = red[s][0];
    for (int w = 1; w < nwv; w++) x = (s == 4 || s == 5) ? fmax(x, red[s][w]) : x + red[s][w];
    part[(size_t)s * gridDim.x + blockIdx.x] = x;
  }
}

// block s: out8[s] = fixed-order reduction (sum, or max for s = 4, 5) of the nb partials part[s * nb + b]
__global__ __launch_bounds__(256) void stats_finish(const double *__restrict__ part, uint32_t nb, double *__restrict__ out8) {
  __shared__ double red[256];
  const int s = blockIdx.x;
  const bool mx = (s == 4 || s == 5);
  const double *p = part + (size_t)s * nb;
  double x = 0.0;
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) x = mx ? fmax(x, p[b]) : x + p[b];
  red[threadIdx.x] = x;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = mx ? fmax(red[threadIdx.x], red[threadIdx.x + h]) : red[threadIdx.x] + red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out8[s] = red[0];
}

// one block: *out = fixed-order sum of x[0..n)
__global__ __launch_bounds__(256) void sum_f64(const double *__restrict__ x, uint32_t n, double *__restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (uint32_t b = threadIdx.x; b < n; b += blockDim.x) s += x[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// storage <-> f32 / f64 conversion of one population plane (read_state / write_state); w = the plane's weight (f: the
// D3Q19 weight of its deviation; g: the D3Q7 weight); th = theta plane of a theta-referenced D3Q7 lattice
// population q of a lattice (is_f: the 19 f populations, else the 7 theta populations) of n cells
template <class S, typename T>
__global__ __launch_bounds__(256) void decode_plane(const void *lat, int q, T *__restrict__ dst, uint32_t n, float w,
                                                    int is_f, const float *__restrict__ th) {
  using E = typename S::elem;
  using R = typename S::real;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const E e = ((const E *)lat)[lat_index<S>(is_f ? 19 : 7, n, q, t)];
  dst[t] = (T)(is_f ? S::dec_f(e, w) : S::dec_g(e, w, th ? R(th[t]) : R(0)));
}
template <class S, typename T>
__global__ __launch_bounds__(256) void encode_plane(const T *__restrict__ src, void *lat, int q, uint32_t n, float w,
                                                    int is_f, const float *__restrict__ th) {
  using E = typename S::elem;
  using R = typename S::real;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  ((E *)lat)[lat_index<S>(is_f ? 19 : 7, n, q, t)] = is_f ? S::enc_f(R(src[t]), w) : S::enc_g(R(src[t]), w, th ? R(th[t]) : R(0));
}

// ---- launchers ------------------------------------------------------------------------------------------------
size_t face_len(int face, int nx, int ny, int nz) { return face < 2 ? 2 * (size_t)nz * ny : 2 * (size_t)nz * nx; }
size_t face_offset(int face, int nx, int ny, int nz) {
  size_t o = 0;
  for (int f = 0; f < face; f++) o += face_len(f, nx, ny, nz);
  return o;
}

uint32_t content_slots(uint32_t nxy, int nz) { return ((nxy + 63u) >> 6) * (uint32_t)nz; }

int storage_bytes(int storage) { return storage == 0 ? 4 : (storage == 3 ? 8 : 2); }
bool storage_theta_ref(int storage) { return storage == 1 || storage == 2; }
static bool storage_16bit(int storage) { return storage == 1 || storage == 2 || storage == 4; }
size_t theta_offset(int storage, int lb, size_t N) {
  if (!storage_theta_ref(storage)) return 0;
  if (lb)   // after the blocked pair's array, from the D3Q7 base (19 populations into it)
    return ((size_t)26 * lattice_cells(lb, N) * (size_t)storage_bytes(storage) + 255) / 256 * 256 -
           ((size_t)19 << lb) * (size_t)storage_bytes(storage);
  return (7 * N * (size_t)storage_bytes(storage) + 255) / 256 * 256;
}
size_t lattice_cells(int lb, size_t N) { return lb ? (N + ((size_t)1 << lb) - 1) >> lb << lb : N; }
size_t lattice_guard_bytes(int storage, int lb, int nx) { return pull_guard_bytes(lb, storage_bytes(storage), nx); }
// a blocked pair is one array of QB populations (the D3Q19 "lattice" covers it, the D3Q7 one lies inside it)
size_t f_lattice_bytes(int storage, int lb, size_t N) {
  if (!lb) return 19 * N * (size_t)storage_bytes(storage);
  const size_t pair = (26 * lattice_cells(lb, N) * (size_t)storage_bytes(storage) + 255) / 256 * 256;
  return pair + (storage_theta_ref(storage) ? 4 * N : 0);   // the theta plane of the 16-bit formats after the array
}
size_t g_lattice_bytes(int storage, int lb, size_t N) {
  if (lb) return 0;
  return storage_theta_ref(storage) ? theta_offset(storage, 0, N) + 4 * N : 7 * N * (size_t)storage_bytes(storage);
}
int population_layout(int storage, size_t N) {
  // blocked when every byte offset of the pair's array (and theta plane) fits 32 bits (cell_off), else [q][cell]
  const int lb = storage == 0 ? StoreF32B::LB : (storage_16bit(storage) ? StoreF16B::LB : 0);
  // (a pulled offset reaches pull_dbytes beyond the lane's own cell offset)
  return lb && f_lattice_bytes(storage, lb, N) + pull_dbytes(lb, storage_bytes(storage)) < ((size_t)1 << 32) ? lb : 0;
}
size_t g_lattice_offset(int storage, int lb) {
  return lb ? ((size_t)19 << lb) * (size_t)storage_bytes(storage) : 0;
}

// the D3Q7 populations of a blocked pair set to 0 (the D3Q19 populations between them untouched)
template <class S>
__global__ __launch_bounds__(256) void zero_g_blocked_kernel(typename S::elem *gq, uint32_t cells) {
  const uint32_t n = blockIdx.x * 256 + threadIdx.x;
  if (n >= cells) return;
  gq[lat_index<S>(7, cells, blockIdx.y, n)] = typename S::elem{};
}

#define AHT_DISPATCH(storage, lb, CALL)                                                \
  switch (storage) {                                                                   \
    case 0:                                                                            \
      if (lb) { using S = StoreF32B; CALL; } else { using S = StoreF32; CALL; }        \
      break;                                                                           \
    case 1:                                                                            \
      if (lb) { using S = StoreF16B; CALL; } else { using S = StoreF16; CALL; }        \
      break;                                                                           \
    case 2:                                                                            \
      if (lb) { using S = StoreBF16B; CALL; } else { using S = StoreBF16; CALL; }      \
      break;                                                                           \
    case 3: { using S = StoreF64; CALL; } break;                                        \
    case 4:                                                                            \
      if (lb) { using S = StoreF16SB; CALL; } else { using S = StoreF16S; CALL; }      \
      break;                                                                           \
    default: return hipErrorInvalidValue;                                              \
  }

template <class S, int KB>
static void launch_step_kb(const Grid &g, const Grid *gd, const Streams *sd, const void *fA, void *fB, const void *gA,
                           void *gB, const Acc *acc, float *out4, int block, uint32_t step, hipStream_t st,
                           uint32_t lds) {
  const uint32_t bx = (g.NXY + block - 1) / block;
  const uint32_t nk = (uint32_t)(g.nz - g.kmin + KB - 1) / KB;
  const uint32_t sw = g.win == WIN_PLANE_SLABS ? (bx + NUM_XCD - 1) / NUM_XCD : 0;   // plane slabs: blocks per slab
  const dim3 grid(sw ? NUM_XCD * sw * nk : bx * nk);
  const Acc a = acc ? *acc : Acc{nullptr, nullptr, nullptr};
  const float ib = 1.0f / (float)bx, isw = sw ? 1.0f / (float)sw : 0.0f;
  if (acc && out4) hipLaunchKernelGGL((step_kernel<S, true, true, KB>), grid, dim3(block), lds, st, gd, sd, fA, fB, gA, gB, a, out4, bx, ib, sw, isw, step);
  else if (acc) hipLaunchKernelGGL((step_kernel<S, true, false, KB>), grid, dim3(block), lds, st, gd, sd, fA, fB, gA, gB, a, out4, bx, ib, sw, isw, step);
  else if (out4) hipLaunchKernelGGL((step_kernel<S, false, true, KB>), grid, dim3(block), lds, st, gd, sd, fA, fB, gA, gB, a, out4, bx, ib, sw, isw, step);
  else hipLaunchKernelGGL((step_kernel<S, false, false, KB>), grid, dim3(block), lds, st, gd, sd, fA, fB, gA, gB, a, out4, bx, ib, sw, isw, step);
}

hipError_t launch_step(int storage, const Grid &g, const Grid *gd, const Streams *sd, const void *fA, void *fB,
                       const void *gA, void *gB, const Acc *acc, float *out4, int block, int levels, uint32_t step,
                       hipStream_t st, uint32_t lds) {
  if (g.kmin >= g.nz) return hipSuccess;
  AHT_DISPATCH(storage, g.lb, {
    if (levels == 4) (launch_step_kb<S, 4>)(g, gd, sd, fA, fB, gA, gB, acc, out4, block, step, st, lds);
    else if (levels == 2) (launch_step_kb<S, 2>)(g, gd, sd, fA, fB, gA, gB, acc, out4, block, step, st, lds);
    else (launch_step_kb<S, 1>)(g, gd, sd, fA, fB, gA, gB, acc, out4, block, step, st, lds);
  })
  return hipGetLastError();
}

// A CU's 4 SIMDs take blocks of block / 64 wavefronts each; at most `waves` per SIMD means at most 256 waves / block
// blocks per CU, which a reservation of lds_cu / blocks bytes (rounded down to 1 KiB, the allocation step assumed
// here) enforces: one more block would need more LDS than the CU has. The step kernel uses no LDS itself.
uint32_t occupancy_lds(int block, int waves, uint32_t lds_cu, uint32_t lds_block) {
  if (waves <= 0 || block <= 0) return 0;
  const uint32_t blocks = (uint32_t)(256 * waves / block);
  if (blocks == 0) return 0;
  const uint32_t lds = lds_cu / blocks / 1024u * 1024u;
  if (lds == 0 || lds > lds_block || (uint64_t)(blocks + 1) * lds <= lds_cu) return 0;
  return lds;
}

bool sweep2_supported(int storage, int lb) { return storage == 0 && lb == StoreF32B::LB; }

hipError_t launch_sweep2(int storage, const Grid &g, const Grid *gd, const void *fA, void *fB, const void *gA,
                         void *gB, int tx, int kc, uint32_t win, uint32_t step, hipStream_t st) {
  if (!sweep2_supported(storage, g.lb) || (tx != 16 && tx != 32) || kc < 1) return hipErrorInvalidValue;
  if (g.kmin >= g.nz) return hipSuccess;
  const int ty = SWEEP_THREADS / tx;
  const uint32_t tiles_x = (uint32_t)(g.nx + tx - 3) / (uint32_t)(tx - 2);
  const uint32_t tiles_y = (uint32_t)(g.ny + ty - 3) / (uint32_t)(ty - 2);
  const uint32_t chunks = (uint32_t)(g.nz - g.kmin + kc - 1) / (uint32_t)kc;
  const dim3 grid(tiles_x * tiles_y * chunks);
  if (win == WIN_PLANE_SLABS) win = 1;
  if (tx == 16)
    hipLaunchKernelGGL((sweep2_kernel<StoreF32BC, 16>), grid, dim3(SWEEP_THREADS), 0, st, gd, fA, fB, gA, gB,
                       tiles_x, tiles_x * tiles_y, kc, win, step);
  else
    hipLaunchKernelGGL((sweep2_kernel<StoreF32BC, 32>), grid, dim3(SWEEP_THREADS), 0, st, gd, fA, fB, gA, gB,
                       tiles_x, tiles_x * tiles_y, kc, win, step);
  return hipGetLastError();
}

// the stream bases of one lattice parity, host-built (the same the device forms with TAB = false)
Streams make_streams(int storage, const Grid &g, const void *fA, void *fB, const void *gA, void *gB) {
  static const int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  static const int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
  static const int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
  const int64_t b = storage_bytes(storage), N = g.N;
  Streams s{};
  if (g.lb) {   // blocked layout: population bases minus the pull's uniform part, the remainder for the lane's compare
    for (int q = 0; q < 19; q++) {
      const int64_t off = cx[q] + (int64_t)cy[q] * g.nx + (int64_t)cz[q] * g.NXY;
      const int64_t sh = pull_shift(g.lb, (int)b, off);
      s.lo[q] = pull_rem(g.lb, off);
      s.rd[q] = (uint64_t)((intptr_t)fA + (intptr_t)(((size_t)q << g.lb) * b) - (intptr_t)sh);
      s.wr[q] = (uint64_t)((uintptr_t)fB + ((size_t)q << g.lb) * b);
      if (q < 7) {
        s.rd[19 + q] = (uint64_t)((intptr_t)gA + (intptr_t)(((size_t)q << g.lb) * b) - (intptr_t)sh);
        s.wr[19 + q] = (uint64_t)((uintptr_t)gB + ((size_t)q << g.lb) * b);
        if (storage_theta_ref(storage)) {   // the theta plane stays [cell]: pulled bases as in [q][cell]
          const int64_t o7 = cx[q] + (int64_t)cy[q] * g.nx + (int64_t)cz[q] * g.NXY;
          s.rd[26 + q] = (uint64_t)((uintptr_t)gA + g.th_off - o7 * 4);
        }
      }
    }
    if (storage_theta_ref(storage)) s.wr[26] = (uint64_t)((uintptr_t)gB + g.th_off);
    return s;
  }
  for (int q = 0; q < 19; q++) {
    const int64_t off = cx[q] + (int64_t)cy[q] * g.nx + (int64_t)cz[q] * g.NXY;   // D3Q7 q < 7 shares c_q
    s.rd[q] = (uint64_t)((uintptr_t)fA + (q * N - off) * b);
    s.wr[q] = (uint64_t)((uintptr_t)fB + q * N * b);
    if (q < 7) {
      s.rd[19 + q] = (uint64_t)((uintptr_t)gA + (q * N - off) * b);
      s.wr[19 + q] = (uint64_t)((uintptr_t)gB + q * N * b);
      if (storage_theta_ref(storage)) s.rd[26 + q] = (uint64_t)((uintptr_t)gA + g.th_off - off * 4);
    }
  }
  if (storage_theta_ref(storage)) s.wr[26] = (uint64_t)((uintptr_t)gB + g.th_off);
  return s;
}

hipError_t launch_zero_g_blocked(int storage, const Grid &g, void *gq, hipStream_t st) {
  if (!g.lb) return hipErrorInvalidValue;
  const uint32_t cells = (uint32_t)lattice_cells(g.lb, g.N);
  AHT_DISPATCH(storage, g.lb, hipLaunchKernelGGL(zero_g_blocked_kernel<S>, dim3((cells + 255) / 256, 7), dim3(256), 0,
                                                 st, (typename S::elem *)gq, cells))
  return hipGetLastError();
}

hipError_t launch_batch_step(int storage, int lb, int block, const Member *mem, int nmem, uint32_t total_tiles, int it,
                             int grid, hipStream_t st, uint32_t lds) {
  if (total_tiles == 0 || nmem == 0) return hipSuccess;
  const uint32_t G = std::min<uint32_t>((uint32_t)grid, total_tiles);
  const bool persist = G < total_tiles;
#define AHT_BATCH_LAUNCH(BS)                                                                                          \
  if (persist) hipLaunchKernelGGL((batch_step_kernel<S, BS, true>), dim3(G), dim3(BS), 0, st, mem, nmem, total_tiles, it); \
  else hipLaunchKernelGGL((batch_step_kernel<S, BS, false>), dim3(G), dim3(BS), lds, st, mem, nmem, total_tiles, it);
  AHT_DISPATCH(storage, lb, {
    if (block == 64) { AHT_BATCH_LAUNCH(64) }
    else if (block == 128) { AHT_BATCH_LAUNCH(128) }
    else if (block == 256) { AHT_BATCH_LAUNCH(256) }
    else return hipErrorInvalidValue;
  })
#undef AHT_BATCH_LAUNCH
  return hipGetLastError();
}

hipError_t launch_init(int storage, const Grid &g, void *f0, void *f1, hipStream_t st) {
  const dim3 grid((g.NXY + 255) / 256, g.nz);
  AHT_DISPATCH(storage, g.lb, hipLaunchKernelGGL(init_kernel<S>, grid, dim3(256), 0, st, g, f0, f1))
  return hipGetLastError();
}

hipError_t launch_reinit(int storage, const Grid &g, const uint32_t *old_ktop, void *f, void *gq, hipStream_t st) {
  AHT_DISPATCH(storage, g.lb, hipLaunchKernelGGL(reinit_kernel<S>, dim3((g.NXY + 255) / 256), dim3(256), 0, st, g, old_ktop, f, gq))
  return hipGetLastError();
}

hipError_t launch_surface(int storage, const Grid &g, const void *fA, const void *gA, float conv, float *out3,
                          hipStream_t st) {
  AHT_DISPATCH(storage, g.lb, hipLaunchKernelGGL(surface_kernel<S>, dim3((g.NXY + 255) / 256), dim3(256), 0, st, g, fA, gA, conv, out3))
  return hipGetLastError();
}

hipError_t launch_volume(int storage, const Grid &g, const void *gA, int nzv, float volscale, uint32_t *vout,
                         uint32_t *tmax_bits, hipStream_t st) {
  const uint32_t nw = ((uint32_t)nzv * g.NXY + 3) / 4;
  AHT_DISPATCH(storage, g.lb, hipLaunchKernelGGL(volume_kernel<S>, dim3((nw + 255) / 256), dim3(256), 0, st, g, gA, nzv, volscale, vout, tmax_bits))
  return hipGetLastError();
}

uint32_t stats_blocks(const Grid &g) { return std::min<uint32_t>((g.N + 255) / 256, STATS_MAX_BLOCKS); }

hipError_t launch_stats(int storage, const Grid &g, const void *fA, const void *gA, double *part, double *out8,
                        hipStream_t st) {
  const uint32_t nb = stats_blocks(g);
  AHT_DISPATCH(storage, g.lb, hipLaunchKernelGGL(stats_kernel<S>, dim3(nb), dim3(256), 0, st, g, fA, gA, part))
  hipLaunchKernelGGL(stats_finish, dim3(7), dim3(256), 0, st, part, nb, out8);
  return hipGetLastError();
}

hipError_t launch_sum(const double *x, uint32_t n, double *out, hipStream_t st) {
  hipLaunchKernelGGL(sum_f64, dim3(1), dim3(256), 0, st, x, n, out);
  return hipGetLastError();
}

hipError_t launch_decode(int storage, int lb, const void *lat, int q, void *dst, int dst_f64, uint32_t n, float w,
                         int is_f, const float *th, hipStream_t st) {
  const dim3 grid((n + 255) / 256);
  AHT_DISPATCH(storage, lb, {
    if (dst_f64) hipLaunchKernelGGL((decode_plane<S, double>), grid, dim3(256), 0, st, lat, q, (double *)dst, n, w, is_f, th);
    else hipLaunchKernelGGL((decode_plane<S, float>), grid, dim3(256), 0, st, lat, q, (float *)dst, n, w, is_f, th);
  })
  return hipGetLastError();
}

hipError_t launch_encode(int storage, int lb, const void *src, int src_f64, void *lat, int q, uint32_t n, float w,
                         int is_f, const float *th, hipStream_t st) {
  const dim3 grid((n + 255) / 256);
  AHT_DISPATCH(storage, lb, {
    if (src_f64) hipLaunchKernelGGL((encode_plane<S, double>), grid, dim3(256), 0, st, (const double *)src, lat, q, n, w, is_f, th);
    else hipLaunchKernelGGL((encode_plane<S, float>), grid, dim3(256), 0, st, (const float *)src, lat, q, n, w, is_f, th);
  })
  return hipGetLastError();
}

}  // namespace aht
