// Query boxes over a uniform target grid level, their enumeration by groups
// of lanes, and the selection of the K smallest keys among a box's targets.
// Shared by the association passes (capped rows, K nearest) and the
// anisotropic association.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cvo_device.h"
#include "cvo_internal.h"

namespace cvo {

// floor(v) clamped to [-1, n] in float before the conversion, so extreme or
// non-finite query coordinates cannot overflow the int cell index (NaN -> -1).
__device__ __forceinline__ int cell_floor(float v, int n) {
  v = fminf(fmaxf(v, -1.0f), (float)n);
  return (int)floorf(v);
}

// Orders this wave's LDS accesses before and after it (no other wave shares a
// row's histogram).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Query box of a row in a grid level: the cell runs (x-rows of cells) it spans.
struct CellBox {
  int lo0, lo1, lo2, nx_run, span1, nruns;
};

// The cells of [q - r, q + r] per axis (no cell when the box misses the grid).
__device__ __forceinline__ CellBox cell_box3(const GridLevel& GL, float q0, float q1, float q2, float r0, float r1,
                                             float r2) {
  CellBox B{0, 0, 0, 0, 0, 0};
  int lo0 = cell_floor((q0 - r0 - GL.ox) * GL.inv_h, GL.nx), hi0 = cell_floor((q0 + r0 - GL.ox) * GL.inv_h, GL.nx);
  int lo1 = cell_floor((q1 - r1 - GL.oy) * GL.inv_h, GL.ny), hi1 = cell_floor((q1 + r1 - GL.oy) * GL.inv_h, GL.ny);
  int lo2 = cell_floor((q2 - r2 - GL.oz) * GL.inv_h, GL.nz), hi2 = cell_floor((q2 + r2 - GL.oz) * GL.inv_h, GL.nz);
  if (hi0 < 0 || lo0 >= GL.nx || hi1 < 0 || lo1 >= GL.ny || hi2 < 0 || lo2 >= GL.nz) return B;
  lo0 = max(lo0, 0); hi0 = min(hi0, GL.nx - 1);
  lo1 = max(lo1, 0); hi1 = min(hi1, GL.ny - 1);
  lo2 = max(lo2, 0); hi2 = min(hi2, GL.nz - 1);
  B.lo0 = lo0;
  B.lo1 = lo1;
  B.lo2 = lo2;
  B.nx_run = hi0 - lo0 + 1;
  B.span1 = hi1 - lo1 + 1;
  B.nruns = B.span1 * (hi2 - lo2 + 1);
  return B;
}

__device__ __forceinline__ CellBox cell_box(const GridLevel& GL, float q0, float q1, float q2, float rq) {
  return cell_box3(GL, q0, q1, q2, rq, rq, rq);
}

// Calls fn(valid, sorted position) for every target of the box, G at a time:
// all G lanes of the group call together (valid false past the end), so fn
// may use group-wide ballots and shuffles.
template <int G, class Fn>
__device__ __forceinline__ void for_each_in_box(const GridLevel& GL, const CellBox& B, int g, Fn&& fn) {
  for (int rbase = 0; rbase < B.nruns; rbase += G) {
    const int myrun = rbase + g;
    uint32_t rs = 0, rl = 0;
    if (myrun < B.nruns) {
      const int iz = myrun / B.span1, iy = myrun - iz * B.span1;
      const uint32_t c0 =
          ((uint32_t)(B.lo2 + iz) * (uint32_t)GL.ny + (uint32_t)(B.lo1 + iy)) * (uint32_t)GL.nx + (uint32_t)B.lo0;
      rs = ld_global(GL.cell_start + c0);
      rl = ld_global(GL.cell_start + c0 + (uint32_t)B.nx_run) - rs;
    }
    uint32_t incl = rl;
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
      uint32_t t = __shfl_up(incl, o, G);
      if (g >= o) incl += t;
    }
    const uint32_t total = __shfl(incl, G - 1, G);
    const uint32_t excl = incl - rl;
    for (uint32_t base = 0; base < total; base += G) {
      const uint32_t idx = base + (uint32_t)g;
      int kk = 0;
#pragma unroll
      for (int s = G / 2; s > 0; s >>= 1) {
        uint32_t e = __shfl(excl, kk + s, G);
        if (e <= idx) kk += s;
      }
      const uint32_t run_start = __shfl(rs, kk, G);
      const uint32_t run_excl = __shfl(excl, kk, G);
      const bool valid = idx < total;
      fn(valid, valid ? run_start + (idx - run_excl) : 0u);
    }
  }
}

// The group's bits of a wave ballot.
template <int G>
__device__ __forceinline__ uint32_t group_bits(uint64_t m) {
  if constexpr (G == 64) {
    return 0u;  // use the 64-bit ballot directly
  } else {
    constexpr uint32_t kGroupMask = (uint32_t)((1ull << G) - 1ull);
    return (uint32_t)(m >> (G * ((threadIdx.x & 63) / G))) & kGroupMask;
  }
}

// Selection of the K smallest keys among a set of targets.  each(fn) calls
// fn(valid, sorted position) for every target of the set, all G lanes of the
// group together (for_each_in_box and its like); key_of(pidx) gives a
// target's key, ~0 for a target that does not take part; the keys that take
// part are distinct.  The group counts them and their range; with more than
// K, a most-significant-digit radix select (8 bits a pass, one 256-bin LDS
// histogram per group, enumerating the set again each pass) finds the K-th
// smallest.  On return a target is among the K smallest exactly when its key
// k != ~0 satisfies (k >> tsh) <= (thr >> tsh); returns the count.
template <int G, class EachFn, class KeyFn>
__device__ __forceinline__ uint32_t select_smallest(EachFn&& each, int g, uint32_t K, uint32_t* hist,
                                                    KeyFn&& key_of, uint64_t& thr, int& tsh) {
  uint32_t cnt = 0;
  uint64_t kmin = ~0ull, kmax = 0;
  each([&](bool valid, uint32_t pidx) {
    if (!valid) return;
    const uint64_t key = key_of(pidx);
    if (key == ~0ull) return;
    cnt++;
    kmin = min(kmin, key);
    kmax = max(kmax, key);
  });
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) {
    cnt += __shfl_xor(cnt, o, G);
    kmin = min(kmin, (uint64_t)__shfl_xor((unsigned long long)kmin, o, G));
    kmax = max(kmax, (uint64_t)__shfl_xor((unsigned long long)kmax, o, G));
  }
  thr = ~0ull;
  tsh = 0;
  if (cnt <= K || K == 0) {
    if (K == 0) tsh = 64;  // nothing is taken: see the test below
    return cnt;
  }
  // keys are distinct, so kmin != kmax and plen < 64
  uint32_t need = K;
  int plen = __clzll((long long)(kmin ^ kmax));
  uint64_t prefix = plen > 0 ? (kmin & ~(~0ull >> plen)) : 0ull;
  constexpr int CH = 256 / G;
  for (int pass = 0; pass < 8 && plen < 64; ++pass) {
    const int db = min(8, 64 - plen);
    const int sh = 64 - plen - db;
#pragma unroll
    for (int i = 0; i < CH; ++i) hist[g * CH + i] = 0u;
    wave_lds_sync();
    each([&](bool valid, uint32_t pidx) {
      if (!valid) return;
      const uint64_t key = key_of(pidx);
      if (key == ~0ull) return;
      if (plen > 0 && (key >> (64 - plen)) != (prefix >> (64 - plen))) return;
      atomicAdd(&hist[(uint32_t)(key >> sh) & 0xffu], 1u);
    });
    wave_lds_sync();
    // the bucket holding the need-th key: chunk sums, group prefix, then a
    // sequential scan inside the one chunk that crosses need
    uint32_t csum = 0;
#pragma unroll
    for (int i = 0; i < CH; ++i) csum += hist[g * CH + i];
    uint32_t incl = csum;
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
      uint32_t t = __shfl_up(incl, o, G);
      if (g >= o) incl += t;
    }
    const uint32_t excl = incl - csum;
    const bool mine = excl < need && incl >= need;
    uint32_t b = 0, less = 0, cb = 0;
    if (mine) {
      uint32_t run = excl;
      for (int i = 0; i < CH; ++i) {
        const uint32_t h = hist[g * CH + i];
        if (run + h >= need) {
          b = (uint32_t)(g * CH + i);
          less = run;
          cb = h;
          break;
        }
        run += h;
      }
    }
    const uint64_t wm = __ballot(mine);
    const uint64_t gm = G == 64 ? wm : (uint64_t)group_bits<G>(wm);
    wave_lds_sync();  // histogram reads done before the next pass clears it
    if (gm == 0ull) break;  // cannot happen while the counts are consistent
    const int owner = __ffsll((unsigned long long)gm) - 1;
    b = __shfl(b, owner, G);
    less = __shfl(less, owner, G);
    cb = __shfl(cb, owner, G);
    prefix |= (uint64_t)b << sh;
    thr = prefix;
    tsh = sh;
    if (less + cb == need) break;
    // continue inside bucket b; thr stays its upper bound meanwhile
    need -= less;
    plen += db;
  }
  return cnt;
}

// Whether a key is among the selected (see select_smallest).
__device__ __forceinline__ bool selected(uint64_t key, uint64_t thr, int tsh) {
  return key != ~0ull && tsh < 64 && (key >> tsh) <= (thr >> tsh);
}

}  // namespace cvo
