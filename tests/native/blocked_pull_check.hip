// Host-side check of the blocked layouts' pulled addresses (storage.h: pull_shift, pull_rem, pull_voff,
// pull_guard_bytes) for every cell the fast path can step (levels 1 .. nz - 2, every column) and every population:
//  * the address a lane forms (stream base + pull_voff) equals the source cell's element in the blocked layout,
//    including sources outside [0, N), whose blocks are numbered on below 0 and above the last block;
//  * the lane's offset fits 32 bits;
//  * every address lies inside the allocation aht_create makes: a guard part, pair A, pair B, a guard part, each part
//    rounded up to 2 MiB (pair = the blocked array of 26 populations, then the f32 theta plane of the theta-referenced
//    16-bit formats).
// Built for the host only (no GPU). Usage: blocked_pull_check nx ny nz [plane stride] ... ; prints one line per grid
// and format, exits non-zero on the first mismatch.
#include <cstdio>
#include <cstdlib>
#include <cstdint>

#include "storage.h"

using namespace aht;

static const int CXq[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int CYq[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int CZq[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};

static int64_t up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// element (population q of 26, cell src) of a blocked pair, blocks numbered on outside [0, N)
template <class S>
static int64_t elem_byte(int q, int64_t src) {
  const int64_t blk = src >> S::LB, lo = src & ((1 << S::LB) - 1);
  return ((blk * S::QB + q) * ((int64_t)1 << S::LB) + lo) * S::bytes;
}

template <class S>
static int check(const char *name, int nx, int ny, int nz, int stride, bool theta_plane) {
  const int64_t NXY = (int64_t)nx * ny, N = NXY * nz;
  const int64_t cells = (N + ((int64_t)1 << S::LB) - 1) >> S::LB << S::LB;
  const int64_t pair = up((int64_t)PAIR_QB * cells * S::bytes, 256) + (theta_plane ? 4 * N : 0);
  if (pair + pull_dbytes(S::LB, S::bytes) >= ((int64_t)1 << 32)) {
    std::printf("%s %dx%dx%d: not blocked (pair of %lld bytes)\n", name, nx, ny, nz, (long long)pair);
    return 0;
  }
  const int64_t A = (int64_t)2 << 20, fb = up(pair, A), guard = up((int64_t)pull_guard_bytes(S::LB, S::bytes, nx), A);
  const int64_t lo_ok = 0, hi_ok = 2 * guard + 2 * fb;
  int64_t checked = 0, outside = 0;
  for (int p = 0; p < 2; p++) {
    const int64_t fA = guard + p * fb, gA = fA + ((int64_t)19 << S::LB) * S::bytes;
    for (int k = 1; k <= nz - 2; k++) {
      if (stride > 1 && k > 2 && k < nz - 3 && k % stride) continue;
      for (int64_t c = 0; c < NXY; c++) {
        const int64_t n = (int64_t)k * NXY + c;
        for (int q = 0; q < 19; q++) {
          const int64_t off = CXq[q] + (int64_t)CYq[q] * nx + (int64_t)CZq[q] * NXY;
          const int64_t sh = pull_shift(S::LB, S::bytes, off);
          const uint32_t o = pull_rem(S::LB, off);
          const uint64_t own = (uint64_t)(n >> S::LB) * ((uint64_t)S::QB << S::LB) * S::bytes +
                               (uint64_t)(n & ((1 << S::LB) - 1)) * S::bytes;
          if (own + pull_dbytes(S::LB, S::bytes) >= ((uint64_t)1 << 32)) {
            std::printf("FAIL %s %dx%dx%d: lane offset of cell %lld exceeds 32 bits\n", name, nx, ny, nz, (long long)n);
            return 1;
          }
          const uint32_t voff = pull_voff<S, 19>((uint32_t)n, o);
          const int64_t src = n - off;
          if (src < 0 || src >= N) outside++;
          for (int lat = 0; lat < (q < 7 ? 2 : 1); lat++) {   // D3Q19 population q, then D3Q7 population q
            const int64_t plane = (lat ? gA : fA) + ((int64_t)q << S::LB) * S::bytes;
            const int64_t addr = plane - sh + (int64_t)voff;
            const int64_t want = fA + elem_byte<S>(lat ? 19 + q : q, src);
            if (addr != want || addr < lo_ok || addr + S::bytes > hi_ok) {
              std::printf("FAIL %s %dx%dx%d pair %d cell %lld q %d lattice %d: address %lld, want %lld, allocation "
                          "[%lld, %lld)\n", name, nx, ny, nz, p, (long long)n, q, lat, (long long)addr, (long long)want,
                          (long long)lo_ok, (long long)hi_ok);
              return 1;
            }
            checked++;
          }
        }
      }
    }
  }
  std::printf("%s %dx%dx%d: %lld pulled addresses checked (%lld from outside the lattice), guard %lld bytes\n", name,
              nx, ny, nz, (long long)checked, (long long)outside, (long long)guard);
  return 0;
}

int main(int argc, char **argv) {
  for (int a = 1; a + 3 <= argc; a += 4) {
    const int nx = atoi(argv[a]), ny = atoi(argv[a + 1]), nz = atoi(argv[a + 2]), st = atoi(argv[a + 3]);
    if (check<StoreF32B>("f32", nx, ny, nz, st, false) || check<StoreF16SB>("f16s", nx, ny, nz, st, false) ||
        check<StoreF16B>("f16", nx, ny, nz, st, true))
      return 1;
  }
  return 0;
}
