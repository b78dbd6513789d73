// Ranking pass and the survivors / refinement / top-k kernel of the gfx950
// search (reference src/index.cpp:998-1127 with kernels.cpp:296-316).
//
// Ranking key of a candidate x of a probed list for query q:
//   r = f(qn + xn - 2 sum_{k<P} qh_k xh_k)  over the shifted high bytes of the
//   leading P rotated dims; key = (float bits of r) << 32 | id, so the u64
//   order is the (float, id) order. Each query keeps the S smallest keys.
//
// A query's threshold is always the S-th smallest key of some subset of its
// candidates, hence never below its final S-th key: filtering by it never drops
// a key of the final S, whatever the order in which work runs.
//
// 1. Probe ranks are split into bands whose width grows geometrically (the
//    first band is each query's nearest list). For each band:
//    k_rank_band (list-major, persistent) takes items (list, up to QMAX of the
//    band's queries of that list, up to a row range) from a device counter. The
//    item's query codes are staged once into LDS in MFMA fragment order; each
//    wave streams chunks of the list's rows straight from HBM into registers as
//    MFMA operands (the next chunk in flight while the current one is ranked),
//    so every row is read once per item and the loads of all waves overlap.
//    A candidate whose integer l2 is within its query's bound is appended to
//    the query's pool. k_update then keeps each pool's S smallest keys and
//    lowers the threshold (and its integer bound) to the S-th: nearer lists
//    come first, so the bound is tight when the far lists are ranked.
// 2. k_refine (one wave per query): the S smallest keys of the pool, exact
//    16-bit distances of those survivors on MFMA, the reference's arrival
//    rule, ascending output.
#include <float.h>

#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"

using namespace katoann_dev;

namespace {

__device__ __forceinline__ v4i load16(const int8_t *p) { return *reinterpret_cast<const v4i *>(p); }

__device__ __forceinline__ v4i mfma(v4i a, v4i b, v4i c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// --------------------------------------------------------------------------
// Ranking pass of one probe-rank band.
struct BandArgs {
  const int8_t *plane;       // N x P, row-major, shifted high bytes
  const int32_t *xnorm;      // N: shifted norm of those bytes
  const uint32_t *ids;       // N
  uint32_t nlist;
  uint32_t band;
  const int8_t *qhi;         // B x dc
  uint32_t dc;
  const int32_t *qnorm;      // B
  const int32_t *lbound;     // B: largest integer l2 whose key can reach the query's threshold
  const uint32_t *qlist;
  const uint32_t *item_off;  // per (band, list) slot, nslots + 1
  const uint4 *items;        // item descriptors: first query slot, queries, first row, end row
  uint32_t stride;           // pool slots per query
  uint64_t *pool;            // B x stride
  uint32_t *pool_cnt;        // B, kCntStride apart
  uint32_t *state;           // [0] overflow, [1] kept, [3 + band] queue head
  double step_sq;
};

// Shape of the ranking kernel for P = 64 KC: T row fragments (16 rows each) per
// wave unit, at most 16 operand registers of 16 bytes per unit. Each of the two
// query buffers holds QB queries (codes, norms, bounds, ids); both buffers and
// the waves' CAP staged keys fit two blocks per CU.
template <int KC>
struct RankShape {
  static constexpr int T = KC == 1 ? 8 : KC <= 4 ? 4 : KC == 5 ? 3 : KC <= 8 ? 2 : 1;
  static constexpr int RC = 16 * T;
  static constexpr int CAP = 256;
  static constexpr int QFIT = ((69 * 1024 / 2) / (64 * KC + 12)) & ~15;
  static constexpr int QB = QFIT > 448 ? 448 : QFIT;    // queries per item
  static constexpr int QA = QB < 64 ? 64 : QB;          // length of the per-query arrays
  static constexpr size_t BUF = static_cast<size_t>(QB) * 64 * KC + 12 * QA;
  static constexpr size_t SMEM = 2 * BUF + 4 * CAP * 10;
};

// Descriptor of every ranking item: thread per item, binary search of its
// (band, list) slot in item_off.
__global__ void k_items(const uint32_t *__restrict__ item_off, uint32_t nslots, uint32_t nlist,
                        const uint64_t *__restrict__ list_off, const uint32_t *__restrict__ gcnt,
                        const uint32_t *__restrict__ qoff, uint32_t qc, uint32_t rt, uint32_t cap, uint4 *items) {
  const uint32_t it = blockIdx.x * 256 + threadIdx.x;
  const uint32_t total = item_off[nslots];
  if (it >= total || it >= cap) return;
  uint32_t lo = 0, hi = nslots;  // item_off[lo] <= it < item_off[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (item_off[mid] <= it) lo = mid;
    else hi = mid;
  }
  const uint32_t slot = lo, c = slot % nlist;
  const uint64_t n = list_off[c + 1] - list_off[c];
  const uint32_t tiles = static_cast<uint32_t>((n + rt - 1) / rt);
  const uint32_t local = it - item_off[slot];
  const uint32_t qch = local / tiles, rtile = local % tiles;
  uint4 d;
  d.x = qoff[slot] + qch * qc;
  d.y = min(qc, gcnt[slot] - qch * qc);
  d.z = static_cast<uint32_t>(list_off[c] + static_cast<uint64_t>(rtile) * rt);
  d.w = static_cast<uint32_t>(list_off[c] + min(n, static_cast<uint64_t>(rtile + 1) * rt));
  items[it] = d;
}

typedef __attribute__((address_space(3))) void lds_void;

// Asynchronous copies from global memory straight into LDS: lane l's `size`
// bytes land at l * size past the wave-uniform LDS address.
__device__ __forceinline__ void dma16(const void *g, void *l) {
  __builtin_amdgcn_global_load_lds(g, (lds_void *)l, 16, 0, 0);
}
__device__ __forceinline__ void dma4(const void *g, void *l) {
  __builtin_amdgcn_global_load_lds(g, (lds_void *)l, 4, 0, 0);
}

// LDS byte address of a pointer into shared memory.
__device__ __forceinline__ uint32_t lds_addr(const void *p) {
  return static_cast<uint32_t>(reinterpret_cast<size_t>((const __attribute__((address_space(3))) char *)p));
}

// LDS accesses of the ranking loop written as asm, each block waiting for its
// own results: the compiler makes every LDS access it can see (and every asm
// declared to touch memory) wait for all copies into LDS in flight, and with
// them for the row prefetch, which would serialise the pipeline. Callers only
// read LDS written before the last block barrier and the wave's own staging
// area, which only these volatile blocks touch (they keep their order).
__device__ __forceinline__ void lds_read4(uint32_t a, v4i &x0, v4i &x1, v4i &x2, v4i &x3) {
  asm volatile(
      "ds_read_b128 %0, %4\n\t"
      "ds_read_b128 %1, %4 offset:1024\n\t"
      "ds_read_b128 %2, %4 offset:2048\n\t"
      "ds_read_b128 %3, %4 offset:3072\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(x0), "=&v"(x1), "=&v"(x2), "=&v"(x3)
      : "v"(a)
);
}
__device__ __forceinline__ void lds_read1(uint32_t a, v4i &x0) {
  asm volatile(
      "ds_read_b128 %0, %1\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(x0)
      : "v"(a)
);
}
__device__ __forceinline__ void lds_read2(uint32_t a, uint32_t b, v4i &x0, v4i &x1) {
  asm volatile(
      "ds_read_b128 %0, %2\n\t"
      "ds_read_b128 %1, %3\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(x0), "=&v"(x1)
      : "v"(a), "v"(b)
);
}
__device__ __forceinline__ void lds_read_key(uint32_t ak, uint32_t aq, uint64_t &key, uint32_t &q) {
  asm volatile(
      "ds_read_b64 %0, %2\n\t"
      "ds_read_u16 %1, %3\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(key), "=&v"(q)
      : "v"(ak), "v"(aq)
);
}
__device__ __forceinline__ uint32_t lds_read_u32(uint32_t a) {
  uint32_t x;
  asm volatile(
      "ds_read_b32 %0, %1\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(x)
      : "v"(a)
);
  return x;
}
__device__ __forceinline__ void lds_store_key(uint32_t ak, uint64_t key, uint32_t aq, uint32_t q) {
  asm volatile(
      "ds_write_b64 %0, %1\n\t"
      "ds_write_b16 %2, %3\n\t"
      "s_waitcnt lgkmcnt(0)"
      :
      : "v"(ak), "v"(key), "v"(aq), "v"(q));
}

// The KC operand pieces of query fragment f (1 KB apart from a).
template <int KC>
__device__ __forceinline__ void read_frag(uint32_t a, v4i (&A)[KC]) {
#pragma unroll
  for (int kc = 0; kc + 4 <= KC; kc += 4) lds_read4(a + 1024 * kc, A[kc], A[kc + 1], A[kc + 2], A[kc + 3]);
#pragma unroll
  for (int kc = KC & ~3; kc < KC; ++kc) lds_read1(a + 1024 * kc, A[kc]);
}

// Work split of one item: row units of RC rows; with fewer row units than
// waves, the query fragments are split too.
struct ItemSplit {
  uint32_t nqf, qsplit, units;
};
template <int RC>
__device__ __forceinline__ ItemSplit split_of(uint4 d) {
  ItemSplit s;
  s.nqf = (d.y + 15) / 16;
  const uint32_t nrc = (d.w - d.z + RC - 1) / RC;
  s.qsplit = nrc >= 4 ? 1u : min(s.nqf, 4u / nrc);
  s.units = nrc * s.qsplit;
  return s;
}

// Persistent blocks of 4 waves, one item at a time, software-pipelined over
// items: while item k is ranked from query buffer k % 2, the next item's query
// ids, norms, bounds and codes are copied into the other buffer by LDS DMA, and
// each wave's first row unit of the next item is loaded into registers during
// its last unit of item k. Item indices come from the band's counter two items
// ahead.
//   A (16 queries x 64 bytes) from LDS, B (16 rows x 64 bytes) in registers:
//   C[i][j] = dot(query i, row j) sits in lane 16 (i / 4) + j, element i % 4.
//   Keep iff qn + xn - 2 dot <= L, i.e. 2 dot - xn >= qn - L.
// Kept keys are staged per wave in LDS with the item-local query and flushed to
// the queries' pools with one returning atomic per key, issued by the whole
// wave at once, so no atomic round trip sits inside the MFMA loop.
template <int KC>
__global__ __launch_bounds__(256, 2) void k_rank_band(BandArgs a) {
  using Sh = RankShape<KC>;
  constexpr int T = Sh::T, RC = Sh::RC, NB = T * KC, CAP = Sh::CAP, QB = Sh::QB, QA = Sh::QA;
  constexpr uint32_t P = 64 * KC;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  uint64_t *skey = reinterpret_cast<uint64_t *>(smem + 2 * Sh::BUF);       // 4 x CAP staged keys
  uint16_t *sq = reinterpret_cast<uint16_t *>(skey + 4 * CAP);              // 4 x CAP their item-local query
  __shared__ uint4 s_desc[3];
  __shared__ uint32_t s_idx[3];
  const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6, g16 = lane >> 4, c16 = lane & 15;
  const uint32_t ib = a.item_off[a.band * a.nlist], ie = a.item_off[(a.band + 1) * a.nlist];
  uint64_t *wkey = skey + w * CAP;
  uint16_t *wq = sq + w * CAP;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  uint32_t n_kept = 0, nst = 0;  // wave-uniform: keys kept, keys staged

  auto bufA = [&](uint32_t k) { return reinterpret_cast<int8_t *>(smem + (k & 1) * Sh::BUF); };
  auto bufN = [&](uint32_t k) { return reinterpret_cast<int32_t *>(smem + (k & 1) * Sh::BUF + QB * P); };
  auto bufL = [&](uint32_t k) { return bufN(k) + QA; };
  auto bufQ = [&](uint32_t k) { return reinterpret_cast<uint32_t *>(bufN(k) + 2 * QA); };

  // staging of an item's queries into buffer k: phase 1 reads the query ids
  // (64 per group, group j by wave j % 4; ids past the item repeat its last),
  // phase 2 copies norms, bounds and codes by those ids
  uint32_t qidr[(QB + 255) / 256];
  auto stage1 = [&](uint4 d, uint32_t k) {
    uint32_t *qd = bufQ(k);
#pragma unroll
    for (int jj = 0; jj < (QB + 255) / 256; ++jj) {
      const uint32_t j = w + 4 * jj;
      if (64 * j < d.y) {
        const uint32_t e = d.x + min(64 * j + lane, d.y - 1);
        qidr[jj] = a.qlist[e];
        dma4(a.qlist + e, qd + 64 * j);
      }
    }
  };
  auto stage2 = [&](uint4 d, uint32_t k) {
    int8_t *A = bufA(k);
    int32_t *N = bufN(k), *L = bufL(k);
    const uint32_t nqf = (d.y + 15) / 16;
#pragma unroll
    for (int jj = 0; jj < (QB + 255) / 256; ++jj) {
      const uint32_t j = w + 4 * jj;
      if (64 * j < d.y) {
        const uint32_t q = qidr[jj];
        dma4(a.qnorm + q, N + 64 * j);
        dma4(a.lbound + q, L + 64 * j);
#pragma unroll
        for (int fr = 0; fr < 4; ++fr) {
          const uint32_t f = 4 * j + fr;
          if (f < nqf) {
            const uint32_t qf = __shfl(q, 16 * fr + c16);
            const int8_t *src = a.qhi + static_cast<size_t>(qf) * a.dc + 16 * g16;
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) dma16(src + 64 * kc, A + (f * KC + kc) * 1024);
          }
        }
      }
    }
  };

  auto load_unit = [&](uint32_t u, uint4 d, uint32_t qsplit, v4i(&b)[NB], int32_t(&x)[T], uint32_t(&id)[T]) {
    const uint32_t r0 = d.z + (u / qsplit) * RC;
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const uint32_t row = min(r0 + 16 * t + c16, d.w - 1);
      const int8_t *rp = a.plane + static_cast<size_t>(row) * P + 16 * g16;
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) b[t * KC + kc] = load16(rp + 64 * kc);
      x[t] = a.xnorm[row];
      id[t] = a.ids[row];
    }
  };

  // staged keys -> pools (query ids of buffer k), 64 at a time: every run of
  // lanes holding the same query reserves its slots with one atomic
  const uint32_t aK = lds_addr(wkey), aQ = lds_addr(wq);
  auto flush = [&](uint32_t k) {
    const uint32_t aI = lds_addr(bufQ(k));
    for (uint32_t i0 = 0; i0 < nst; i0 += 64) {
      const uint32_t i = i0 + lane, nv = min(64u, nst - i0);
      const bool v = i < nst;
      uint64_t key;
      uint32_t ql;
      lds_read_key(aK + 8 * i, aQ + 2 * i, key, ql);
      const uint32_t q = lds_read_u32(aI + 4 * min(ql, static_cast<uint32_t>(QB - 1)));
      const uint32_t qq = v ? q : 0xffffffffu;
      const uint32_t prev = __shfl_up(qq, 1);
      const bool head = v && (lane == 0 || prev != qq);
      const uint64_t hm = ballot(head);
      uint32_t base = 0;
      if (head) {
        const uint64_t above = lane == 63 ? 0ull : hm >> (lane + 1);
        const uint32_t nxt = above ? lane + 1 + static_cast<uint32_t>(__builtin_ctzll(above)) : nv;
        base = atomicAdd(&a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride], nxt - lane);
      }
      const uint64_t le = lane == 63 ? ~0ull : (2ull << lane) - 1ull;
      const int h = 63 - __builtin_clzll(hm & le | 1ull);
      const uint32_t pos = __shfl(base, h) + static_cast<uint32_t>(lane - h);
      if (v) {
        if (pos < a.stride) a.pool[static_cast<size_t>(q) * a.stride + pos] = key;
        else atomicOr(&a.state[0], 1u);
      }
    }
    nst = 0;
  };

  if (tid == 0) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const uint32_t x = ib + atomicAdd(&a.state[3 + a.band], 1u);
      s_idx[i] = x;
      s_desc[i] = x < ie ? a.items[x] : uint4{0, 1, 0, 1};
    }
  }
  __syncthreads();
  if (s_idx[0] >= ie) return;
  v4i bc[NB], bn[NB];
  int32_t xc[T], xn2[T];
  uint32_t ic[T], in2[T];
  {
    const uint4 d0 = s_desc[0];
    stage1(d0, 0);
    stage2(d0, 0);
    const ItemSplit sp = split_of<RC>(d0);
    if (static_cast<uint32_t>(w) < sp.units) load_unit(w, d0, sp.qsplit, bc, xc, ic);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }

  for (uint32_t k = 0;; ++k) {
    const uint4 d = s_desc[k % 3], nd = s_desc[(k + 1) % 3];
    const bool nvalid = s_idx[(k + 1) % 3] < ie;
    const ItemSplit sp = split_of<RC>(d), nsp = split_of<RC>(nd);
    uint32_t i2 = 0;
    uint4 d2 = {0, 1, 0, 1};
    if (tid == 0) i2 = ib + atomicAdd(&a.state[3 + a.band], 1u);
    if (nvalid) stage1(nd, k + 1);

    const uint32_t aA = lds_addr(bufA(k)), aN = lds_addr(bufN(k)), aL = lds_addr(bufL(k));
    bool had = false;
    for (uint32_t u = w; u < sp.units; u += 4) {
      had = true;
      // this wave's following unit (same item, else its first of the next item,
      // else the current one again): loaded unconditionally so that the waits
      // for it stay exact
      const bool same = u + 4 < sp.units;
      const bool cross = !same && nvalid && static_cast<uint32_t>(w) < nsp.units;
      load_unit(same ? u + 4 : cross ? static_cast<uint32_t>(w) : u, cross ? nd : d,
                cross ? nsp.qsplit : sp.qsplit, bn, xn2, in2);
      const uint32_t rc = u / sp.qsplit, qs = u % sp.qsplit;
      const uint32_t r0 = d.z + rc * RC;
      const uint32_t fa = qs * sp.nqf / sp.qsplit, fb = (qs + 1) * sp.nqf / sp.qsplit;
      // hot pass: distances of every fragment, and which fragments keep any
      uint32_t fmask = 0;
      for (uint32_t f = fa; f < fb; ++f) {
        v4i Af[KC], qn4, L4;
        read_frag<KC>(aA + (f * KC * 64 + lane) * 16, Af);
        lds_read2(aN + (16 * f + 4 * g16) * 4, aL + (16 * f + 4 * g16) * 4, qn4, L4);
        const v4i h4 = qn4 - L4;
        v4i acc[T];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int kc = 0; kc < KC; ++kc)
#pragma unroll
          for (int t = 0; t < T; ++t) acc[t] = mfma(Af[kc], bc[t * KC + kc], acc[t]);
        bool any = false;
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) any |= 2 * acc[t][r] - xc[t] >= h4[r];
        if (ballot(any)) fmask |= 1u << (f - fa);
      }
      // cold pass over the fragments that keep candidates: recompute, stage
      // the kept keys, flush when the stage is full
      while (fmask) {
        const uint32_t f = fa + static_cast<uint32_t>(__builtin_ctz(fmask));
        fmask &= fmask - 1;
        v4i Af[KC], qn4, L4;
        read_frag<KC>(aA + (f * KC * 64 + lane) * 16, Af);
        lds_read2(aN + (16 * f + 4 * g16) * 4, aL + (16 * f + 4 * g16) * 4, qn4, L4);
        const v4i h4 = qn4 - L4;
        v4i acc[T];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int kc = 0; kc < KC; ++kc)
#pragma unroll
          for (int t = 0; t < T; ++t) acc[t] = mfma(Af[kc], bc[t * KC + kc], acc[t]);
#pragma unroll 1
        for (int r = 0; r < 4; ++r) {
          const uint32_t ql = 16 * f + 4 * g16 + r;
          const bool qok = ql < d.y;
          const int32_t hr = h4[r], qnr = qn4[r];
#pragma unroll
          for (int t = 0; t < T; ++t) {
            const int32_t dv = acc[t][r];
            const bool bit = qok && r0 + 16 * t + c16 < d.w && 2 * dv - xc[t] >= hr;
            const uint64_t m = ballot(bit);
            if (!m) continue;
            if (nst + 64 > CAP) flush(k);
            if (bit) {
              const uint32_t pos = nst + static_cast<uint32_t>(__popcll(m & lt_mask));
              lds_store_key(aK + 8 * pos, make_key(fscale(qnr + xc[t] - 2 * dv, a.step_sq), ic[t]), aQ + 2 * pos,
                            ql);
            }
            nst += static_cast<uint32_t>(__popcll(m));
            n_kept += static_cast<uint32_t>(__popcll(m));
          }
        }
      }
#pragma unroll
      for (int i = 0; i < NB; ++i) bc[i] = bn[i];
#pragma unroll
      for (int t = 0; t < T; ++t) {
        xc[t] = xn2[t];
        ic[t] = in2[t];
      }
    }
    if (nvalid) {
      stage2(nd, k + 1);
      if (!had && static_cast<uint32_t>(w) < nsp.units) load_unit(w, nd, nsp.qsplit, bc, xc, ic);
    }
    if (nst) flush(k);
    if (tid == 0) {
      if (i2 < ie) d2 = a.items[i2];
      s_idx[(k + 2) % 3] = i2;
      s_desc[(k + 2) % 3] = d2;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (!nvalid) break;
  }
  if (lane == 0 && n_kept) atomicAdd(&a.state[1], n_kept);
}

// After a band: each query's pool back to its S smallest keys, ascending, and
// the query's threshold to the S-th of them (an upper bound of its final S-th
// key), with the integer l2 bound of that threshold for the next band's
// filter. Entries [0, sorted) are the previous update's sorted list.
struct UpdateArgs {
  uint64_t *pool;
  uint32_t *pool_cnt, *pool_sorted;
  uint64_t *thr;
  int32_t *lbound;
  uint32_t B, S, stride;
  int64_t lmax;  // P * 255^2: largest possible l2
  double step_sq;
};

template <int SC>
__global__ __launch_bounds__(256) void k_update(UpdateArgs a) {
  const int lane = lane_id();
  const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= a.B) return;
  const uint32_t n = min(a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride], a.stride),
                 sorted = a.pool_sorted[q];
  if (n == sorted) return;
  uint64_t *pq = a.pool + static_cast<size_t>(q) * a.stride;
  TopList<SC> t;
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) {
    const uint32_t i = lane + 64 * r;
    t.e[r] = i < sorted ? pq[i] : kNoKey;
  }
  for (uint32_t b = sorted; b < n; b += 64) {
    const uint64_t k = b + lane < n ? pq[b + lane] : kNoKey;
    const uint64_t mk = ballot(k != kNoKey);
    if (__popcll(mk) <= 4) {
      uint64_t mm = mk;
      while (mm) {
        const int bb = __ffsll(static_cast<long long>(mm)) - 1;
        mm &= mm - 1;
        top_insert<SC>(t, __shfl(k, bb), a.S);
      }
    } else {
      top_merge<SC>(t, k, a.S);
    }
  }
  const uint32_t cnt = top_count<SC>(t);
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) {
    const uint32_t i = lane + 64 * r;
    if (i < cnt) pq[i] = t.e[r];
  }
  if (lane == 0) {
    a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride] = cnt;
    a.pool_sorted[q] = cnt;
  }
  if (cnt == a.S) {
    const uint64_t kth = top_get<SC>(t, a.S - 1);
    if (lane == 0 && kth < a.thr[q]) {
      a.thr[q] = kth;
      a.lbound[q] = static_cast<int32_t>(l2_bound(key_float(kth), a.step_sq, a.lmax));
    }
  }
}

// --------------------------------------------------------------------------
struct RefineArgs {
  const uint64_t *pool;      // B x stride
  uint32_t stride;
  const uint32_t *pool_cnt;
  const uint64_t *thr;
  uint32_t B, S, K;
  const int8_t *hi, *lo;     // N x dc
  const int64_t *norm;       // N
  const uint32_t *id_row;    // npts
  uint32_t npts, dc;
  const int8_t *qhi, *qlo;   // B x dc
  const int64_t *qnorm;      // B
  double step_sq;
  uint32_t *out_ids;
  float *out_dists;
  uint64_t *surv;            // B x S or null
  uint32_t *nsurv;           // B or null
};

template <int SC>
__device__ inline uint64_t wave_max_list(const TopList<SC> &t) {
  uint64_t m = 0;
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) m = umax64(m, t.e[r] == kNoKey ? 0ull : t.e[r]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = umax64(m, __shfl_xor(m, o));
  return m;
}

// Final survivors, exact distances, arrival rule, output (src/index.cpp:1046-1127):
//  1. the S smallest keys of the query's pool (exactly the S smallest ranking
//     keys of all its candidates);
//  2. exact distances of the survivors: rows gathered by id, the four byte
//     products on MFMA (A = 16 survivor rows of one plane, B = the query's high
//     and low bytes as columns 0 and 1), combined in int64;
//  3. the reference's arrival rule: survivors in descending (rank, id) order
//     enter a k-entry max-heap of (distance, id) while it has room, then
//     replace its maximum iff their distance is strictly smaller;
//  4. output ascending by (distance, id), padded with FLT_MAX / UINT32_MAX
//     (src/index.cpp:384-398).
template <int SC>
__global__ __launch_bounds__(256) void k_refine(RefineArgs a) {
  __shared__ uint64_t ex[4][SC];
  const int lane = lane_id(), w = threadIdx.x >> 6;
  const uint32_t q = blockIdx.x * 4 + w;
  if (q >= a.B) return;
  TopList<SC> sv;
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) sv.e[r] = kNoKey;
  {
    const uint64_t base = static_cast<uint64_t>(q) * a.stride;
    const uint32_t m = min(a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride], a.stride);
    const uint64_t tq = a.thr[q];
    for (uint32_t b = 0; b < m; b += 64) {
      uint64_t k = b + lane < m ? a.pool[base + b + lane] : kNoKey;
      if (k > tq) k = kNoKey;
      const uint64_t mk = ballot(k != kNoKey);
      if (!mk) continue;
      if (__popcll(mk) <= 4) {
        uint64_t mm = mk;
        while (mm) {
          const int bb = __ffsll(static_cast<long long>(mm)) - 1;
          mm &= mm - 1;
          top_insert<SC>(sv, __shfl(k, bb), a.S);
        }
      } else {
        top_merge<SC>(sv, k, a.S);
      }
    }
  }
  const uint32_t ns = top_count<SC>(sv);
  if (a.surv) {
#pragma unroll
    for (int r = 0; r < SC / 64; ++r) {
      const uint32_t i = lane + 64 * r;
      if (i < a.S) a.surv[static_cast<size_t>(q) * a.S + i] = sv.e[r];
    }
  }
  if (a.nsurv && lane == 0) a.nsurv[q] = ns;

  const int8_t *qh = a.qhi + static_cast<size_t>(q) * a.dc;
  const int8_t *ql = a.qlo + static_cast<size_t>(q) * a.dc;
  const int64_t qn = a.qnorm[q];
  const int koff = 16 * (lane >> 4), col = lane & 15;
  for (uint32_t t0 = 0; t0 < ns; t0 += 16) {
    const uint32_t s = t0 + (lane & 15);
    const uint64_t key = top_get<SC>(sv, min(s, static_cast<uint32_t>(SC - 1)));
    const uint32_t id = static_cast<uint32_t>(key);
    const bool sok = s < ns && id < a.npts;
    const uint32_t row = sok ? a.id_row[id] : 0xffffffffu;
    const bool rok = sok && row != 0xffffffffu;
    v4i ah = {0, 0, 0, 0}, al = {0, 0, 0, 0};
    for (uint32_t kc = 0; kc < a.dc; kc += 64) {
      const v4i xh = rok ? load16(a.hi + static_cast<size_t>(row) * a.dc + kc + koff) : v4i{0, 0, 0, 0};
      const v4i xl = rok ? load16(a.lo + static_cast<size_t>(row) * a.dc + kc + koff) : v4i{0, 0, 0, 0};
      const v4i b = col == 0 ? load16(qh + kc + koff) : (col == 1 ? load16(ql + kc + koff) : v4i{0, 0, 0, 0});
      ah = mfma(xh, b, ah);
      al = mfma(xl, b, al);
    }
    // survivor i = lane (< 16): C[i][0] in lane 16 (i/4), element i % 4; C[i][1] one lane up
    const int i = lane & 15;
    const int src = 16 * (i >> 2);
    int32_t hh = 0, hl = 0, lh = 0, ll = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int32_t a0 = __shfl(ah[r], src), a1 = __shfl(ah[r], src + 1);
      const int32_t b0 = __shfl(al[r], src), b1 = __shfl(al[r], src + 1);
      if ((i & 3) == r) {
        hh = a0;
        hl = a1;
        lh = b0;
        ll = b1;
      }
    }
    const uint32_t si = t0 + i;
    const bool mine_ok = __shfl(static_cast<int>(rok), i) != 0;
    const uint32_t myrow = __shfl(row, i);
    const uint32_t myid = __shfl(id, i);
    if (lane < 16 && si < ns) {
      uint64_t e = kNoKey;
      if (mine_ok) {
        const int64_t dot = static_cast<int64_t>(hh) * 65536 + (static_cast<int64_t>(hl) + lh) * 256 +
                            static_cast<int64_t>(ll);
        e = make_key(fscale(qn + a.norm[myrow] - 2 * dot, a.step_sq), myid);
      }
      ex[w][si] = e;
    }
  }
  wave_fence();

  TopList<SC> h;
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) h.e[r] = kNoKey;
  uint32_t nh = 0;
  uint64_t mx = 0;
  for (int t = static_cast<int>(ns) - 1; t >= 0; --t) {
    const uint64_t c = ex[w][t];
    if (c == kNoKey) continue;
    if (nh < a.K) {
#pragma unroll
      for (int r = 0; r < SC / 64; ++r)
        if (static_cast<uint32_t>(r) == nh / 64 && lane == static_cast<int>(nh % 64)) h.e[r] = c;
      ++nh;
      if (nh == a.K) mx = wave_max_list<SC>(h);
    } else if (key_float(c) < key_float(mx)) {
      bool done = false;
#pragma unroll
      for (int r = 0; r < SC / 64; ++r) {
        const uint64_t hit = ballot(h.e[r] == mx);
        if (!done && hit) {
          const int hl2 = __ffsll(static_cast<long long>(hit)) - 1;
          if (lane == hl2) h.e[r] = c;
          done = true;
        }
      }
      mx = wave_max_list<SC>(h);
    }
  }
  list_sort<SC>(h);
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) {
    const uint32_t g = lane + 64 * r;
    if (g < a.K) {
      const bool ok = g < nh;
      a.out_ids[static_cast<size_t>(q) * a.K + g] = ok ? static_cast<uint32_t>(h.e[r]) : 0xffffffffu;
      a.out_dists[static_cast<size_t>(q) * a.K + g] = ok ? key_float(h.e[r]) : FLT_MAX;
    }
  }
}

}  // namespace

namespace katoann_hip {

uint32_t refine_cap(uint32_t S) {
  return S <= 64 ? 64 : S <= 128 ? 128 : S <= 256 ? 256 : S <= 512 ? 512 : S <= 1024 ? 1024 : 0;
}

uint32_t rank_queries_per_item(uint32_t P) {
  switch (P / 64) {
#define QM_CASE(n) \
  case n:          \
    return RankShape<n>::QB;
    QM_CASE(1) QM_CASE(2) QM_CASE(3) QM_CASE(4) QM_CASE(5) QM_CASE(6) QM_CASE(7) QM_CASE(8)
    QM_CASE(9) QM_CASE(10) QM_CASE(11) QM_CASE(12) QM_CASE(13) QM_CASE(14) QM_CASE(15) QM_CASE(16)
#undef QM_CASE
    default:
      return 0;
  }
}

void launch_items(const uint32_t *item_off, uint32_t nslots, uint32_t nlist, const uint64_t *list_off,
                  const uint32_t *gcnt, const uint32_t *qoff, uint32_t qc, uint32_t rt, uint32_t cap, uint4 *items,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_items, dim3((cap + 255) / 256), dim3(256), 0, s, item_off, nslots, nlist, list_off, gcnt, qoff,
                     qc, rt, cap, items);
}

template <int KC>
static void band_launch(const BandArgs &a, uint32_t grid, hipStream_t s) {
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_rank_band<KC>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(RankShape<KC>::SMEM));
    attr = true;
  }
  hipLaunchKernelGGL(k_rank_band<KC>, dim3(grid), dim3(256), RankShape<KC>::SMEM, s, a);
}

bool launch_rank_band(uint32_t band, const int8_t *plane, const int32_t *xnorm, const uint32_t *ids, uint32_t nlist,
                      uint32_t P, const int8_t *qhi, uint32_t dc, const int32_t *qnorm, const int32_t *lbound,
                      const uint32_t *qlist, const uint32_t *item_off, const uint4 *items, uint32_t stride,
                      double step_sq, uint64_t *pool, uint32_t *pool_cnt, uint32_t *state, uint32_t grid,
                      hipStream_t s) {
  BandArgs a{plane, xnorm, ids, nlist, band, qhi, dc, qnorm, lbound, qlist, item_off, items, stride,
             pool, pool_cnt, state, step_sq};
  switch (P / 64) {
#define KC_CASE(n) \
  case n:          \
    band_launch<n>(a, grid, s); \
    return true;
    KC_CASE(1) KC_CASE(2) KC_CASE(3) KC_CASE(4) KC_CASE(5) KC_CASE(6) KC_CASE(7) KC_CASE(8)
    KC_CASE(9) KC_CASE(10) KC_CASE(11) KC_CASE(12) KC_CASE(13) KC_CASE(14) KC_CASE(15) KC_CASE(16)
#undef KC_CASE
    default:
      return false;
  }
}

bool launch_update(uint64_t *pool, uint32_t *pool_cnt, uint32_t *pool_sorted, uint64_t *thr, int32_t *lbound,
                   uint32_t B, uint32_t S, uint32_t stride, uint32_t P, double step_sq, hipStream_t s) {
  UpdateArgs a{pool, pool_cnt, pool_sorted, thr, lbound, B, S, stride, static_cast<int64_t>(P) * 65025, step_sq};
  const dim3 grid((B + 3) / 4);
  switch (refine_cap(S)) {
    case 64: hipLaunchKernelGGL(k_update<64>, grid, dim3(256), 0, s, a); return true;
    case 128: hipLaunchKernelGGL(k_update<128>, grid, dim3(256), 0, s, a); return true;
    case 256: hipLaunchKernelGGL(k_update<256>, grid, dim3(256), 0, s, a); return true;
    case 512: hipLaunchKernelGGL(k_update<512>, grid, dim3(256), 0, s, a); return true;
    case 1024: hipLaunchKernelGGL(k_update<1024>, grid, dim3(256), 0, s, a); return true;
    default: return false;
  }
}


bool launch_refine(const uint64_t *pool, uint32_t stride, const uint32_t *pool_cnt, const uint64_t *thr,
                   uint32_t B, uint32_t S, uint32_t K, const int8_t *hi, const int8_t *lo, const int64_t *norm,
                   const uint32_t *id_row, uint32_t npts, uint32_t dc, const int8_t *qhi, const int8_t *qlo,
                   const int64_t *qnorm, double step_sq, uint32_t *out_ids, float *out_dists, uint64_t *surv,
                   uint32_t *nsurv, hipStream_t s) {
  RefineArgs a{pool, stride, pool_cnt, thr, B, S, K, hi, lo, norm, id_row, npts, dc, qhi, qlo, qnorm, step_sq,
               out_ids, out_dists, surv, nsurv};
  const dim3 grid((B + 3) / 4);
  switch (refine_cap(S)) {
    case 64: hipLaunchKernelGGL(k_refine<64>, grid, dim3(256), 0, s, a); return true;
    case 128: hipLaunchKernelGGL(k_refine<128>, grid, dim3(256), 0, s, a); return true;
    case 256: hipLaunchKernelGGL(k_refine<256>, grid, dim3(256), 0, s, a); return true;
    case 512: hipLaunchKernelGGL(k_refine<512>, grid, dim3(256), 0, s, a); return true;
    case 1024: hipLaunchKernelGGL(k_refine<1024>, grid, dim3(256), 0, s, a); return true;
    default: return false;
  }
}

}  // namespace katoann_hip
