// Device-side layout of a loaded index and wave-level helpers shared by the
// gfx950 kernels. One wavefront = 64 lanes.
//
// Codes are stored shifted: s = u - 128 for every byte of the 16-bit code
// u = 256 h + l (so c - 32896 = 256 (h - 128) + (l - 128)). Squared L2 is
// unchanged by a common shift, so every integer distance equals the reference's
// unsigned one, and gfx950's signed int8 MFMA applies without bias terms.
// Padded dimensions hold 0 on both sides and add nothing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace katoann_dev {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;
constexpr uint64_t kNoKey = ~0ull;   // empty slot; above every real key

// f(l2) = float(double(l2) * step_sq), round to nearest even (reference
// kernels.cpp:166-168, 209-210).
__device__ __forceinline__ float fscale(int64_t l2, double step_sq) {
  return static_cast<float>(static_cast<double>(l2) * step_sq);
}

// Ranking / result key: float bits (non-negative floats order like their bits)
// above the 32-bit id, so the u64 order is the (float, id) order.
__device__ __forceinline__ uint64_t make_key(float f, uint32_t id) {
  return (static_cast<uint64_t>(__float_as_uint(f)) << 32) | id;
}

__device__ __forceinline__ float key_float(uint64_t k) {
  return __uint_as_float(static_cast<uint32_t>(k >> 32));
}

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x & 63); }

__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint64_t umax64(uint64_t a, uint64_t b) { return a < b ? b : a; }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Largest l in [0, lmax] with fscale(l) <= t (f is monotone non-decreasing);
// -1 if none. Candidates whose integer l2 exceeds it cannot reach key t.
__device__ inline int64_t l2_bound(float t, double step_sq, int64_t lmax) {
  if (!(fscale(0, step_sq) <= t)) return -1;
  if (fscale(lmax, step_sq) <= t) return lmax;
  double g = floor(static_cast<double>(t) / step_sq);
  int64_t lo = 0, hi = lmax;  // invariant: f(lo) <= t < f(hi)
  if (g >= 0.0 && g < static_cast<double>(lmax)) {
    const int64_t gi = static_cast<int64_t>(g);
    if (fscale(gi, step_sq) <= t) {
      lo = gi;
      if (gi + 1 <= lmax && !(fscale(gi + 1, step_sq) <= t)) return gi;
    } else {
      hi = gi;
    }
  }
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (fscale(mid, step_sq) <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ---- wave-wide sorting of 64-bit keys (one key per lane, ascending) --------

__device__ __forceinline__ uint64_t cmpx(uint64_t v, int j, bool keep_min) {
  const uint64_t o = __shfl_xor(v, j);
  return keep_min ? umin64(v, o) : umax64(v, o);
}

__device__ inline uint64_t wave_sort64(uint64_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const bool up = (lane & k) == 0;
      const bool lower = (lane & j) == 0;
      v = cmpx(v, j, lower == up);
    }
  }
  return v;
}

// Sorts a bitonic sequence held one per lane into ascending order.
__device__ inline uint64_t wave_bitonic_merge64(uint64_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int j = 32; j > 0; j >>= 1) v = cmpx(v, j, (lane & j) == 0);
  return v;
}

// Sorted list of up to SC keys, SC = 64 * R (R a power of two): register r
// of lane l holds entry 64 r + l. Empty slots are kNoKey.
template <int SC>
struct TopList {
  uint64_t e[SC / 64];
};

// Bitonic merge of a bitonic sequence of SC entries into ascending order:
// strides >= 64 compare registers, strides < 64 compare lanes.
template <int SC>
__device__ inline void list_bitonic_merge(TopList<SC> &t) {
  constexpr int R = SC / 64;
#pragma unroll
  for (int j = R / 2; j >= 1; j >>= 1)
#pragma unroll
    for (int r = 0; r < R; ++r)
      if ((r & j) == 0) {
        const uint64_t lo = umin64(t.e[r], t.e[r + j]), hi = umax64(t.e[r], t.e[r + j]);
        t.e[r] = lo;
        t.e[r + j] = hi;
      }
#pragma unroll
  for (int r = 0; r < R; ++r) t.e[r] = wave_bitonic_merge64(t.e[r]);
}

// Full ascending sort of the SC entries: the bitonic network over index
// i = 64 r + lane (stride >= 64: registers; stride < 64: lanes).
template <int SC>
__device__ inline void list_sort(TopList<SC> &t) {
  constexpr int R = SC / 64;
  const int lane = lane_id();
#pragma unroll
  for (int k = 2; k <= SC; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
        const int jr = j / 64;
#pragma unroll
        for (int r = 0; r < R; ++r)
          if ((r & jr) == 0) {
            const int p = r | jr;
            const bool up = ((64 * r) & k) == 0;
            const uint64_t lo = umin64(t.e[r], t.e[p]), hi = umax64(t.e[r], t.e[p]);
            t.e[r] = up ? lo : hi;
            t.e[p] = up ? hi : lo;
          }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const bool up = ((64 * r + lane) & k) == 0;
          const bool lower = (lane & j) == 0;
          t.e[r] = cmpx(t.e[r], j, lower == up);
        }
      }
    }
  }
}

// Keeps the S smallest of (list, c), where c holds at most one key per lane
// (kNoKey where none). Keys are unique.
template <int SC>
__device__ inline void top_merge(TopList<SC> &t, uint64_t c, uint32_t S) {
  const int lane = lane_id();
  const uint64_t cs = wave_sort64(c);
  const uint64_t crev = __shfl(cs, 63 - lane);
  t.e[SC / 64 - 1] = umin64(t.e[SC / 64 - 1], crev);
  list_bitonic_merge<SC>(t);
#pragma unroll
  for (int r = 0; r < SC / 64; ++r)
    if (static_cast<uint32_t>(lane + 64 * r) >= S) t.e[r] = kNoKey;
}

// Inserts one key (wave-uniform) into the sorted list, keeping S entries.
template <int SC>
__device__ inline void top_insert(TopList<SC> &t, uint64_t k, uint32_t S) {
  const int lane = lane_id();
  uint32_t pos = 0;
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) pos += __popcll(ballot(t.e[r] < k));
  if (pos >= S) return;
  uint64_t carry = 0;
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) {
    const uint64_t old = t.e[r];
    uint64_t prev = __shfl_up(old, 1);
    if (lane == 0) prev = carry;
    carry = __shfl(old, 63);
    const uint32_t g = static_cast<uint32_t>(lane + 64 * r);
    t.e[r] = g < pos ? old : (g == pos ? k : prev);
    if (g >= S) t.e[r] = kNoKey;
  }
}

// Entry i of the list; i may differ per lane (both registers are read from the
// source lane, then the caller's register is picked).
template <int SC>
__device__ __forceinline__ uint64_t top_get(const TopList<SC> &t, uint32_t i) {
  uint64_t v = 0;
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) {
    const uint64_t x = __shfl(t.e[r], static_cast<int>(i % 64));
    if (i / 64 == static_cast<uint32_t>(r)) v = x;
  }
  return v;
}

template <int SC>
__device__ __forceinline__ uint32_t top_count(const TopList<SC> &t) {
  uint32_t n = 0;
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) n += __popcll(ballot(t.e[r] != kNoKey));
  return n;
}

// The S-th smallest (1-based) of the valid keys held CH per lane (kNoKey
// slots ignored); at least S valid keys. Bitwise search on the float bits,
// then on the ids among keys sharing the float of the S-th.
template <int CH>
__device__ inline uint64_t kth_key(const uint64_t (&e)[CH], uint32_t S) {
  uint32_t hi = 0;
  for (int b = 31; b >= 0; --b) {
    const uint32_t t = hi | (1u << b);
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < CH; ++i) c += __popcll(ballot(static_cast<uint32_t>(e[i] >> 32) < t));
    if (c < S) hi = t;
  }
  uint32_t lt = 0, eq = 0;
  uint32_t mlo = 0;
#pragma unroll
  for (int i = 0; i < CH; ++i) {
    const uint32_t h = static_cast<uint32_t>(e[i] >> 32);
    lt += __popcll(ballot(h < hi));
    eq += __popcll(ballot(h == hi));
    if (h == hi) mlo = max(mlo, static_cast<uint32_t>(e[i]));
  }
  const uint32_t need = S - lt;
  uint32_t lo = 0;
  if (need == eq) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mlo = max(mlo, static_cast<uint32_t>(__shfl_xor(static_cast<int>(mlo), o)));
    lo = mlo;
  } else {
    for (int b = 31; b >= 0; --b) {
      const uint32_t t = lo | (1u << b);
      uint32_t c = 0;
#pragma unroll
      for (int i = 0; i < CH; ++i)
        c += __popcll(ballot(static_cast<uint32_t>(e[i] >> 32) == hi && static_cast<uint32_t>(e[i]) < t));
      if (c < need) lo = t;
    }
  }
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Block-wide exclusive scan of one value per thread (256 threads); tmp holds
// 4 words.
__device__ inline uint32_t block_exscan256(uint32_t v, uint32_t *tmp) {
  const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) tmp[w] = x;
  __syncthreads();
  uint32_t base = 0;
  for (int k = 0; k < w; ++k) base += tmp[k];
  __syncthreads();
  return base + x - v;
}

template <int W>
__device__ inline void wave_sort_run(uint64_t *run) {
  const int lane = lane_id();
  TopList<W> t;
#pragma unroll
  for (int r = 0; r < W / 64; ++r) t.e[r] = run[64 * r + lane];
  list_sort<W>(t);
#pragma unroll
  for (int r = 0; r < W / 64; ++r) run[64 * r + lane] = t.e[r];
}

__device__ inline void sort_runs(uint64_t *buf, uint32_t W, uint32_t nruns) {
  const int w = threadIdx.x >> 6;
  if (static_cast<uint32_t>(w) >= nruns) return;
  uint64_t *run = buf + static_cast<size_t>(w) * W;
  switch (W) {
    case 64: wave_sort_run<64>(run); break;
    case 128: wave_sort_run<128>(run); break;
    case 256: wave_sort_run<256>(run); break;
    case 512: wave_sort_run<512>(run); break;
    default: wave_sort_run<1024>(run); break;
  }
}

// number of keys in sorted run a[0, n) below x (upper: at or below x)
__device__ __forceinline__ uint32_t count_below(const uint64_t *a, uint32_t n, uint64_t x, bool upper) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (upper ? a[mid] <= x : a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}


// Sorts keys[0, M) ascending (M a power of two, 64..4096, 256 threads), with
// keys[M, 2 M) as the second buffer: runs of M / 4 (or M) keys sorted by the
// waves in registers, then merge levels by rank (merge path); returns the
// buffer that holds the result.
__device__ inline uint64_t *block_sort_keys(uint64_t *keys, uint32_t M) {
  const int tid = threadIdx.x;
  const uint32_t W = M >= 256 ? M / 4 : M, nruns = M / W;
  sort_runs(keys, W, nruns);
  __syncthreads();
  uint64_t *src = keys, *dst = keys + M;
  for (uint32_t L = W; L < M; L <<= 1) {
    for (uint32_t e = tid; e < M; e += 256) {
      const uint32_t pair = e / (2 * L) * (2 * L), off = e - pair;
      const uint64_t x = src[e];
      const bool inA = off < L;
      const uint64_t *other = src + pair + (inA ? L : 0);
      const uint32_t pos = (inA ? off : off - L) + count_below(other, L, x, !inA);
      dst[pair + pos] = x;
    }
    __syncthreads();
    uint64_t *t = src;
    src = dst;
    dst = t;
  }
  return src;
}

}  // namespace katoann_dev
