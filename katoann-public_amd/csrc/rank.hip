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
//    the ranking kernel (list-major, persistent) runs items (list, up to QC of
//    the band's queries of that list, up to a row range), dealt round-robin:
//    k_rank_wave (default) keeps an item's rows in a wave's registers and
//    streams its queries through; k_rank_band stages an item's queries in LDS
//    and streams its rows through the block's waves.
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
  const int8_t *qhi;         // B rows dc bytes apart, the ranked bytes first
  uint32_t dc;
  const int32_t *qnorm;      // B
  const int32_t *lbound;     // B: largest integer l2 whose key can reach the query's threshold
  const uint32_t *qlist;
  const uint32_t *item_off;  // per (band, list) slot, nslots + 1
  const uint4 *items;        // item descriptors: first query slot, queries, first row, end row
  uint32_t stride;           // pool slots per query
  uint64_t *pool;            // B x stride
  uint32_t *pool_cnt;        // B, kCntStride apart
  uint32_t *state;           // [0] overflow, [1] kept
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

// Descriptor of every ranking item: thread per item (slot order: binary search
// of its (band, list) slot in item_off), placed within its band by cost class
// (katoann_hip::item_class of its query count; ihist: the per (band, class)
// counts, then cursors). A block's items of one (band, class) take one run of
// positions (one atomic on the cursor), in item order within each wave, so a
// row tile's query chunks stay neighbours and are dealt to neighbouring waves
// of one round: they read the tile's rows from HBM about once per L2.
__global__ __launch_bounds__(256) void k_items(const uint32_t *__restrict__ item_off, uint32_t nslots,
                                               uint32_t nlist, const uint64_t *__restrict__ list_off,
                                               const uint32_t *__restrict__ gcnt, const uint32_t *__restrict__ qoff,
                                               uint32_t qc, uint32_t rt, uint32_t qsq, uint32_t qsrt, uint32_t cap,
                                               uint32_t seed, uint32_t join, const uint32_t *__restrict__ gbv,
                                               uint32_t *ihist, uint4 *items) {
  constexpr uint32_t NC = katoann_hip::kItemClasses, NH = katoann_hip::kPlanBands * NC;
  __shared__ uint32_t base[NH];   // first position of (band, class) in the items, then this block's run
  __shared__ uint32_t cnt[NH];    // this block's items per (band, class)
  __shared__ uint32_t wcnt[4][NH];
  const uint32_t total = min(item_off[nslots], cap);
  if (blockIdx.x * 256 >= total) return;
  const uint32_t it = blockIdx.x * 256 + threadIdx.x;
  const bool valid = it < total;
  const int lane = lane_id(), w = threadIdx.x >> 6;
  const uint32_t nb = (nslots + nlist - 1) / nlist;
  for (uint32_t i = threadIdx.x; i < NH; i += 256) {
    cnt[i] = 0;
    wcnt[0][i] = wcnt[1][i] = wcnt[2][i] = wcnt[3][i] = 0;
  }
  if (threadIdx.x < nb) {  // band b's classes start at its first item, most fragments first
    uint32_t at = item_off[threadIdx.x * nlist];
    for (uint32_t c = 0; c < NC; ++c) {
      base[threadIdx.x * NC + c] = at;
      at += ihist[threadIdx.x * NC + c];
    }
  }
  uint4 d{0, 0, 0, 0};
  uint32_t key = 0;  // band * NC + class
  if (valid) {
    uint32_t lo = 0, hi = nslots;  // item_off[lo] <= it < item_off[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (item_off[mid] <= it) lo = mid;
      else hi = mid;
    }
    const uint32_t slot = lo, c = slot % nlist;
    uint64_t first, end;
    katoann_hip::slot_rows(slot / nlist, list_off[c + 1] - list_off[c], seed, first, end);
    const uint64_t n = end - first;
    uint32_t local = it - item_off[slot];
    // a slot of the first probe-rank band holding its list's joined seed
    // remainder: part 1 (rows [0, n1), its own gb queries), then part 2 (rows
    // [n1, n), all gcnt); otherwise one part of gcnt queries over all rows
    uint32_t g = gcnt[slot];
    uint64_t r0 = first, r1 = end;
    const bool joined = join && slot / nlist == katoann_hip::kJoinBand && gbv[c] < g;
    if (joined) {
      const uint32_t gb = gbv[c], n1 = static_cast<uint32_t>(min(n, static_cast<uint64_t>(seed)));
      const uint32_t i1 = katoann_hip::chunk_items(gb, n1, qc, rt);
      if (local < i1) {
        g = gb;
        r1 = first + n1;
      } else {
        local -= i1;
        r0 = first + n1;
      }
    }
    const uint64_t nr = r1 - r0;
    if (!joined && katoann_hip::slot_qs(g, qsq)) {  // all the slot's queries, rows qsrt at a time
      d.x = qoff[slot];
      d.y = gcnt[slot] | katoann_hip::kQsItem;
      d.z = static_cast<uint32_t>(list_off[c] + first + static_cast<uint64_t>(local) * qsrt);
      d.w = static_cast<uint32_t>(list_off[c] + first + min(n, static_cast<uint64_t>(local + 1) * qsrt));
      key = (slot / nlist) * NC + katoann_hip::item_class(gcnt[slot]);
    } else {  // the query chunks of one row tile are consecutive items
      const uint32_t nch = (g + qc - 1) / qc;
      const uint32_t qch = local % nch, rtile = local / nch;
      d.x = qoff[slot] + qch * qc;
      d.y = min(qc, g - qch * qc);
      d.z = static_cast<uint32_t>(list_off[c] + r0 + static_cast<uint64_t>(rtile) * rt);
      d.w = static_cast<uint32_t>(list_off[c] + r0 + min(nr, static_cast<uint64_t>(rtile + 1) * rt));
      key = (slot / nlist) * NC + katoann_hip::item_class(d.y);
    }
  }
  __syncthreads();
  // rank within the wave among its items of the same key (lane order), the
  // wave's count per key, then the waves' offsets and the block's run
  uint32_t rank = 0;
  {
    uint64_t todo = ballot(valid);
    while (todo) {
      const uint32_t k = static_cast<uint32_t>(__shfl(static_cast<int>(key), __ffsll(static_cast<long long>(todo)) - 1));
      const uint64_t m = ballot(valid && key == k) & todo;
      if (valid && key == k) rank = static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
      if (lane == 0) wcnt[w][k] = static_cast<uint32_t>(__popcll(m));
      todo &= ~m;
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < NH; i += 256) {
    const uint32_t c0 = wcnt[0][i], c1 = wcnt[1][i], c2 = wcnt[2][i], c3 = wcnt[3][i];
    wcnt[0][i] = 0;
    wcnt[1][i] = c0;
    wcnt[2][i] = c0 + c1;
    wcnt[3][i] = c0 + c1 + c2;
    cnt[i] = c0 + c1 + c2 + c3;
    if (cnt[i]) base[i] += atomicAdd(&ihist[NH + i], cnt[i]);
  }
  __syncthreads();
  if (valid) {
    const uint32_t pos = base[key] + wcnt[w][key] + rank;
    if (pos < cap) items[pos] = d;
  }
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
      const int h = 63 - __builtin_clzll((hm & le) | 1ull);
      const uint32_t pos = __shfl(base, h) + static_cast<uint32_t>(lane - h);
      if (v) {
        if (pos < a.stride) a.pool[static_cast<size_t>(q) * a.stride + pos] = key;
        else atomicOr(&a.state[0], 1u);
      }
    }
    nst = 0;
  };

  // items dealt round-robin over the blocks (a shared queue counter
  // serialises every taker on one address): block k takes k, k + G, ...
  uint32_t next_item = ib + blockIdx.x;
  if (tid == 0) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const uint32_t x = next_item;
      next_item += gridDim.x;
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
    if (tid == 0) {
      i2 = next_item;
      next_item += gridDim.x;
    }
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

// --------------------------------------------------------------------------
// Ranking pass of one probe-rank band, wave-independent form. Each wave takes
// items (list, up to 16 T rows, up to QC of the band's queries of the list)
// from the band's counter, keeps the item's rows in registers as MFMA B
// operands, and streams the item's queries through 16 at a time as A operands
// loaded straight from the query codes (a few MB, L2-resident): the query ids
// two fragments ahead, codes, norms and bounds one fragment ahead. No LDS and
// no block barriers: the waves of a CU overlap one another's row loads, MFMAs
// and filters freely.
//   Keep  <=>  qn + xn - 2 dot <= L  <=>  2 dot - xn >= qn - L. The MFMA
//   accumulators start at -ceil(xn / 2), so the test of a fragment is a max
//   over the lane's rows and one compare per query slot; fragments in which a
//   row passes form the keys f(qn + xn - 2 dot) (keep_exact).
// Kept keys go to the pools through the wave's deferred stage.
// Items are dealt to the waves round-robin (no shared counter).
struct WaveArgs {
  const int8_t *plane;       // N x P, row-major, shifted high bytes
  const int32_t *xnorm;      // N
  const uint32_t *ids;       // N
  const int8_t *qhi;         // B rows dc bytes apart, the ranked bytes first
  uint32_t dc;
  const int32_t *qnorm;      // B
  const int32_t *lbound;     // B
  const uint32_t *qlist;
  const uint32_t *item_off;  // per (band, list) slot, nslots + 1
  const uint4 *items;        // first query slot, queries, first row, end row
  uint32_t nlist;
  uint32_t stride;
  uint64_t *pool;
  uint32_t *pool_cnt;
  uint32_t *state;           // [0] overflow, [1] kept
  uint32_t band;
  double step_sq;
};

// --------------------------------------------------------------------------
// Deferred pool appends of one wave. A fragment's kept keys are staged in the
// wave's LDS region with their query slot j (0..15) and their index within the
// slot's run; one returning atomic per query reserves the run in the query's
// pool, and the keys are written to the pool when the wave next keeps keys (or
// at its end), so the atomic's round trip overlaps the MFMAs in between. A
// fragment with more than KCAP keys (the first band keeps every candidate) is
// written at once instead.
constexpr int KCAP = 256;

struct KeepStage {
  uint64_t *key;       // LDS: 2 x KCAP per wave
  uint32_t *meta;      // LDS: 2 x KCAP per wave: slot j | index << 4
  uint32_t pending;    // keys staged in region `par` (wave-uniform)
  int par;
  uint32_t base, q;    // lane 16 g + r (r < 4): pool base and query of slot 4 g + r
};

// (A loop with a fixed trip count and at most one overflow atomic, so the
// compiler can count the wave's outstanding memory operations across it.)
__device__ __forceinline__ void keep_drain(KeepStage &ks, const WaveArgs &a) {
  const int lane = lane_id();
  bool ovf = false;
#pragma unroll
  for (int c = 0; c < KCAP / 64; ++c) {
    if (static_cast<uint32_t>(64 * c) >= ks.pending) break;
    const uint32_t e = 64 * c + lane;
    const bool v = e < ks.pending;
    const uint32_t m = v ? ks.meta[ks.par * KCAP + e] : 0u;
    const uint64_t key = v ? ks.key[ks.par * KCAP + e] : 0ull;
    const int src = 16 * ((m & 15) >> 2) + (m & 3);
    const uint32_t pos = __shfl(ks.base, src) + (m >> 4);
    const uint32_t q = __shfl(ks.q, src);
    if (v) {
      if (pos < a.stride) a.pool[static_cast<size_t>(q) * a.stride + pos] = key;
      else ovf = true;
    }
  }
  if (ovf) atomicOr(&a.state[0], 1u);
  ks.pending = 0;
}

// Inclusive prefix sums within each 16-lane row of the wave (DPP row shifts).
__device__ __forceinline__ uint32_t row_scan16(uint32_t x) {
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x111, 0xf, 0xf, true));
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x112, 0xf, 0xf, true));
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x114, 0xf, 0xf, true));
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x118, 0xf, 0xf, true));
  return x;
}

__device__ __forceinline__ uint32_t byte_sum(uint32_t x) {
  return (x & 0xffu) + ((x >> 8) & 0xffu) + ((x >> 16) & 0xffu) + (x >> 24);
}

// v[t] for a lane-varying t < T.
template <int T, class V>
__device__ __forceinline__ V pick(const V (&v)[T], int t) {
  V r = v[0];
#pragma unroll
  for (int i = 1; i < T; ++i) {
    r = t == i ? v[i] : r;
    asm volatile("" : "+v"(r));  // a select chain, not an indexed stack array
  }
  return r;
}

// The filter of a fragment. The accumulators are seeded with -ceil(xn / 2), so
// acc[t][r] = dot(row 16 t + c16, query slot 4 g16 + r) - ceil(xn / 2), and a
// row is kept iff acc >= H with H = floor((qn - L) / 2), i.e. iff
// 2 dot - xn >= 2 floor((qn - L) / 2): every row with l2 <= L, plus those with
// l2 = L + 1 when qn - L is odd. The extra rows are real candidates of the
// query, and a pool keeps its S smallest keys, so they change no result.
// keep_exact takes a fragment in which some row passed: the lane's kept rows
// as bit masks per query slot, their runs in the wave's stage (row-wise prefix
// sums of byte-packed counts), and the keys f(qn - 2 acc - (xn & 1)). H4[r]:
// H of slot 4 g16 + r; vmask / oddm: bit t set iff row 16 t + c16 is in the
// item / has an odd norm; qslot(j) / qnorm_slot(j): id and norm of query slot j.
// Returns the keys kept.
template <int T, class QS, class QN>
__device__ __forceinline__ uint32_t keep_exact(KeepStage &ks, const WaveArgs &a, const v4i (&acc)[T],
                                               uint32_t vmask, uint32_t oddm, const uint32_t (&xid)[T],
                                               const int32_t (&H4)[4], QS qslot, QN qnorm_slot) {
  const int lane = lane_id(), g16 = lane >> 4, c16 = lane & 15;
  // norms of the lane's query slots, fetched first so the shuffles' latency
  // overlaps the tests below
  int32_t qnr[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) qnr[r] = qnorm_slot(4 * g16 + r);
  uint32_t bits[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    uint32_t m = 0;
    // bit t <- sign of H - 1 - acc (set iff acc >= H), shifted in from t = T - 1 down
#pragma unroll
    for (int t = T - 1; t >= 0; --t)
      m = __builtin_amdgcn_alignbit(m, static_cast<uint32_t>(H4[r] - 1 - acc[t][r]), 31);
    bits[r] = m & vmask;
  }
  const uint32_t c = static_cast<uint32_t>(__builtin_popcount(bits[0])) |
                     static_cast<uint32_t>(__builtin_popcount(bits[1])) << 8 |
                     static_cast<uint32_t>(__builtin_popcount(bits[2])) << 16 |
                     static_cast<uint32_t>(__builtin_popcount(bits[3])) << 24;
  // per 16-lane group g, byte r: keys of slot 4 g + r (at most 16 T < 256)
  const uint32_t inc = row_scan16(c), ex = inc - c;
  const uint32_t G0 = __builtin_amdgcn_readlane(inc, 15), G1 = __builtin_amdgcn_readlane(inc, 31);
  const uint32_t G2 = __builtin_amdgcn_readlane(inc, 47), G3 = __builtin_amdgcn_readlane(inc, 63);
  const uint32_t S0 = byte_sum(G0), S1 = byte_sum(G1), S2 = byte_sum(G2), S3 = byte_sum(G3);
  const uint32_t total = S0 + S1 + S2 + S3;
  if (total == 0) return 0;
  const uint32_t Gm = g16 == 0 ? G0 : g16 == 1 ? G1 : g16 == 2 ? G2 : G3;
  uint32_t so[4];  // run of slot 4 g16 + r in the stage
  so[0] = (g16 > 0 ? S0 : 0u) + (g16 > 1 ? S1 : 0u) + (g16 > 2 ? S2 : 0u);
#pragma unroll
  for (int r = 1; r < 4; ++r) so[r] = so[r - 1] + ((Gm >> (8 * (r - 1))) & 0xffu);
  const uint32_t mine = c16 < 4 ? (Gm >> (8 * c16)) & 0xffu : 0u;
  const uint32_t qmine = qslot(4 * g16 + (c16 & 3));
  // a fragment with more than KCAP keys (the first band keeps every candidate)
  // is written to the pools at once
  const bool direct = total > static_cast<uint32_t>(KCAP);
  uint32_t base = 0;
  if (direct) {
    keep_drain(ks, a);
    if (c16 < 4 && mine) base = atomicAdd(&a.pool_cnt[static_cast<size_t>(qmine) * katoann_hip::kCntStride], mine);
  }
  const int reg = ks.par ^ 1;
  bool ovf = false;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int j = 4 * g16 + r;
    const uint32_t q = direct ? qslot(j) : 0u;
    const uint32_t boff = direct ? __shfl(base, 16 * g16 + r) : 0u;
    uint32_t idx = (ex >> (8 * r)) & 0xffu;  // the lane's first key within the slot's run
    int32_t av[T];
#pragma unroll
    for (int t = 0; t < T; ++t) av[t] = acc[t][r];
    uint32_t m = bits[r];
    while (m) {
      const int t = __builtin_ctz(m);
      m &= m - 1u;
      const int32_t l2 = qnr[r] - 2 * pick(av, t) - static_cast<int32_t>((oddm >> t) & 1u);
      const uint64_t key = make_key(fscale(l2, a.step_sq), pick(xid, t));
      if (direct) {
        const uint32_t pos = boff + idx;
        if (pos < a.stride) a.pool[static_cast<size_t>(q) * a.stride + pos] = key;
        else ovf = true;
      } else {
        ks.key[reg * KCAP + so[r] + idx] = key;
        ks.meta[reg * KCAP + so[r] + idx] = static_cast<uint32_t>(j) | (idx << 4);
      }
      ++idx;
    }
  }
  if (direct) {
    if (ovf) atomicOr(&a.state[0], 1u);
    return total;
  }
  if (c16 < 4 && mine) base = atomicAdd(&a.pool_cnt[static_cast<size_t>(qmine) * katoann_hip::kCntStride], mine);
  keep_drain(ks, a);
  ks.par = reg;
  ks.pending = total;
  ks.base = base;
  ks.q = qmine;
  return total;
}

// --------------------------------------------------------------------------
// Run-staged pool appends of one wave (the wave kernel). Kept keys stay in the
// wave's LDS stage across fragments and items, each with the index of its
// run: the keys one fragment keeps for one query, contiguous in the stage,
// recorded as (query, first stage entry, length). The stage goes to the pools
// only when it is full and at the wave's end: one returning atomic per run
// reserves the run in its query's pool, then every key is written. So the
// fragment loop issues no global store or atomic, and the compiler's waits
// for the query fragments in flight count only those loads; the flush ends by
// waiting for its own stores and atomics, so none is left for a later wait.
constexpr uint32_t RS_KEYS = 1024;  // staged keys per wave
constexpr uint32_t RS_RUNS = 256;   // staged runs per wave

constexpr int vm_wait_imm(int n) { return (n & 0xf) | (((n >> 4) & 3) << 14) | (0x7 << 4) | (0xf << 8); }

struct RunStage {
  uint64_t *key;     // RS_KEYS
  uint16_t *run;     // RS_KEYS: run of each key
  uint32_t *rq;      // RS_RUNS: query of each run
  uint32_t *rfirst;  // RS_RUNS: first stage entry (low 16 bits) | length << 16
  uint32_t *rbase;   // RS_RUNS: pool position reserved at the flush
  uint32_t n, nr;    // keys and runs staged (wave-uniform)
};

__device__ __forceinline__ void runs_flush(RunStage &rs, const WaveArgs &a) {
  const int lane = lane_id();
  wave_fence();
  for (uint32_t i = lane; i < rs.nr; i += 64) {
    const uint32_t f = rs.rfirst[i];
    rs.rbase[i] = atomicAdd(&a.pool_cnt[static_cast<size_t>(rs.rq[i]) * katoann_hip::kCntStride], f >> 16);
  }
  wave_fence();
  bool ovf = false;
  for (uint32_t e = lane; e < rs.n; e += 64) {
    const uint32_t r = rs.run[e];
    const uint32_t pos = rs.rbase[r] + (e - (rs.rfirst[r] & 0xffffu));
    if (pos < a.stride) a.pool[static_cast<size_t>(rs.rq[r]) * a.stride + pos] = rs.key[e];
    else ovf = true;
  }
  if (ovf) atomicOr(&a.state[0], 1u);
  rs.n = 0;
  rs.nr = 0;
  __builtin_amdgcn_s_waitcnt(vm_wait_imm(0));
  wave_fence();
}

// keep_exact's filter and keys (same bit masks, counts and key formula), with
// the keys staged as runs. A fragment with more keys than the stage holds (the
// first band keeps every candidate) is written to the pools at once, after the
// stage, and waited for.
template <int T, class QS, class QN>
__device__ __forceinline__ uint32_t keep_runs(RunStage &rs, const WaveArgs &a, const v4i (&acc)[T], uint32_t vmask,
                                              uint32_t oddm, const uint32_t (&xid)[T], const int32_t (&H4)[4],
                                              QS qslot, QN qnorm_slot) {
  const int lane = lane_id(), g16 = lane >> 4, c16 = lane & 15;
  int32_t qnr[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) qnr[r] = qnorm_slot(4 * g16 + r);
  uint32_t bits[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    uint32_t m = 0;
#pragma unroll
    for (int t = T - 1; t >= 0; --t)
      m = __builtin_amdgcn_alignbit(m, static_cast<uint32_t>(H4[r] - 1 - acc[t][r]), 31);
    bits[r] = m & vmask;
  }
  const uint32_t c = static_cast<uint32_t>(__builtin_popcount(bits[0])) |
                     static_cast<uint32_t>(__builtin_popcount(bits[1])) << 8 |
                     static_cast<uint32_t>(__builtin_popcount(bits[2])) << 16 |
                     static_cast<uint32_t>(__builtin_popcount(bits[3])) << 24;
  const uint32_t inc = row_scan16(c), ex = inc - c;
  const uint32_t G0 = __builtin_amdgcn_readlane(inc, 15), G1 = __builtin_amdgcn_readlane(inc, 31);
  const uint32_t G2 = __builtin_amdgcn_readlane(inc, 47), G3 = __builtin_amdgcn_readlane(inc, 63);
  const uint32_t S0 = byte_sum(G0), S1 = byte_sum(G1), S2 = byte_sum(G2), S3 = byte_sum(G3);
  const uint32_t total = S0 + S1 + S2 + S3;
  if (total == 0) return 0;
  const uint32_t Gm = g16 == 0 ? G0 : g16 == 1 ? G1 : g16 == 2 ? G2 : G3;
  uint32_t so[4];  // first entry of slot 4 g16 + r within the fragment's keys
  so[0] = (g16 > 0 ? S0 : 0u) + (g16 > 1 ? S1 : 0u) + (g16 > 2 ? S2 : 0u);
#pragma unroll
  for (int r = 1; r < 4; ++r) so[r] = so[r - 1] + ((Gm >> (8 * (r - 1))) & 0xffu);
  const uint32_t mine = c16 < 4 ? (Gm >> (8 * c16)) & 0xffu : 0u;
  const uint32_t qmine = qslot(4 * g16 + (c16 & 3));  // every lane takes part in the shuffle
  if (total > RS_KEYS) {
    if (rs.n) runs_flush(rs, a);
    uint32_t base = 0;
    if (c16 < 4 && mine) base = atomicAdd(&a.pool_cnt[static_cast<size_t>(qmine) * katoann_hip::kCntStride], mine);
    bool ovf = false;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint32_t q = qslot(4 * g16 + r);
      const uint32_t boff = __shfl(base, 16 * g16 + r);
      uint32_t idx = (ex >> (8 * r)) & 0xffu;
      int32_t av[T];
#pragma unroll
      for (int t = 0; t < T; ++t) av[t] = acc[t][r];
      uint32_t m = bits[r];
      while (m) {
        const int t = __builtin_ctz(m);
        m &= m - 1u;
        const int32_t l2 = qnr[r] - 2 * pick(av, t) - static_cast<int32_t>((oddm >> t) & 1u);
        const uint32_t pos = boff + idx;
        if (pos < a.stride) a.pool[static_cast<size_t>(q) * a.stride + pos] = make_key(fscale(l2, a.step_sq), pick(xid, t));
        else ovf = true;
        ++idx;
      }
    }
    if (ovf) atomicOr(&a.state[0], 1u);
    __builtin_amdgcn_s_waitcnt(vm_wait_imm(0));
    return total;
  }
  // slots with keys: bit 4 g + c of nz for slot 4 g + c; each gets a run
  const uint64_t nzb = ballot(c16 < 4 && mine > 0);
  uint32_t nz = 0;
#pragma unroll
  for (int g = 0; g < 4; ++g) nz |= static_cast<uint32_t>((nzb >> (16 * g)) & 0xfu) << (4 * g);
  const uint32_t nr = static_cast<uint32_t>(__builtin_popcount(nz));
  if (rs.n + total > RS_KEYS || rs.nr + nr > RS_RUNS) runs_flush(rs, a);
  if (c16 < 4 && mine) {
    const uint32_t j = 4u * g16 + c16;
    uint32_t sj = so[0];
#pragma unroll
    for (int r = 1; r < 4; ++r) sj = c16 == r ? so[r] : sj;
    const uint32_t ri = rs.nr + static_cast<uint32_t>(__builtin_popcount(nz & ((1u << j) - 1u)));
    rs.rq[ri] = qmine;
    rs.rfirst[ri] = (rs.n + sj) | (mine << 16);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t j = 4u * g16 + r;
    const uint32_t ri = rs.nr + static_cast<uint32_t>(__builtin_popcount(nz & ((1u << j) - 1u)));
    uint32_t e = rs.n + so[r] + ((ex >> (8 * r)) & 0xffu);
    int32_t av[T];
#pragma unroll
    for (int t = 0; t < T; ++t) av[t] = acc[t][r];
    uint32_t m = bits[r];
    while (m) {
      const int t = __builtin_ctz(m);
      m &= m - 1u;
      const int32_t l2 = qnr[r] - 2 * pick(av, t) - static_cast<int32_t>((oddm >> t) & 1u);
      rs.key[e] = make_key(fscale(l2, a.step_sq), pick(xid, t));
      rs.run[e] = static_cast<uint16_t>(ri);
      ++e;
    }
  }
  rs.n += total;
  rs.nr += nr;
  return total;
}

// Per item: the MFMA seeds -ceil(xn / 2) of the lane's rows, and the masks of
// rows in the item and of odd norms.
template <int T>
__device__ __forceinline__ void row_seeds(const int32_t (&xn)[T], uint32_t rows, v4i (&seed)[T], uint32_t &vmask,
                                          uint32_t &oddm) {
  const int c16 = lane_id() & 15;
  vmask = 0;
  oddm = 0;
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int32_t c = -((xn[t] + 1) >> 1);
    seed[t] = v4i{c, c, c, c};
    vmask |= static_cast<uint32_t>(16u * t + c16 < rows) << t;
    oddm |= static_cast<uint32_t>(xn[t] & 1) << t;
  }
}

// Whether any row of the fragment passes: max over the lane's rows per slot.
template <int T>
__device__ __forceinline__ bool fragment_passes(const v4i (&acc)[T], const int32_t (&H4)[4]) {
  bool any = false;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    int32_t m = acc[0][r];
#pragma unroll
    for (int t = 1; t < T; ++t) m = max(m, acc[t][r]);
    any |= m >= H4[r];
  }
  return ballot(any) != 0;
}

// One query fragment of the wave kernel: the lane's A operands, its query id,
// norm and bound.
template <int KC>
struct QFrag {
  v4i A[KC];
  int32_t qn, lb;
  uint32_t q;
};

// Item descriptor i through the scalar cache (wave-uniform i): the table was
// written by an earlier launch and is read-only here.
__device__ __forceinline__ uint4 item_desc(const uint4 *items, uint32_t i) {
  typedef const __attribute__((address_space(4))) uint32_t *cptr;
  const cptr q = (cptr)(reinterpret_cast<uintptr_t>(items + i));
  return uint4{q[0], q[1], q[2], q[3]};
}

// One item of the wave form (rows in registers, queries streamed).
template <int KC, int T, int NB>
__device__ __forceinline__ void wave_item(const uint4 dv, const WaveArgs &a, RunStage &rs, uint32_t &n_kept) {
  constexpr uint32_t P = 64 * KC;
  const int lane = lane_id(), g16 = lane >> 4, c16 = lane & 15;
  // the descriptor is the wave's (one item per wave): in scalar registers, the
  // fragment loop's trip count and tests are uniform branches
  const uint4 d{__builtin_amdgcn_readfirstlane(dv.x), __builtin_amdgcn_readfirstlane(dv.y),
                __builtin_amdgcn_readfirstlane(dv.z), __builtin_amdgcn_readfirstlane(dv.w)};
  const uint32_t nq = d.y, nf = (nq + 15) / 16, rows = d.w - d.z;
  const uint32_t *ql = a.qlist + d.x;
  // query of lane c16 in fragment f (clamped into the item)
  auto qid_of = [&](uint32_t f) { return ql[min(16 * f + c16, nq - 1)]; };
  // Query fragments stream through a ring of NB register buffers, codes
  // loaded NB - 1 fragments ahead and ids NB ahead; the loop is unrolled by
  // NB so no buffer is ever copied (a copy would wait for the loads just
  // issued). Ids are loaded before the loads that follow them (the rows,
  // a step's codes), so waiting for an id waits for nothing issued after
  // it; fragments past the item load its last query again (unconditional
  // loads leave no register copies). Norms and bounds stay raw until the
  // fragment is ranked.
  uint32_t qfirst[NB];
#pragma unroll
  for (int u = 0; u < NB; ++u) qfirst[u] = qid_of(u);
  // the item's rows: B operands, norms and ids
  v4i b[T * KC];
  int32_t xn[T];
  uint32_t xid[T];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const uint32_t row = min(d.z + 16 * t + c16, d.w - 1);
    const int8_t *rp = a.plane + static_cast<size_t>(row) * P + 16 * g16;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) b[t * KC + kc] = load16(rp + 64 * kc);
    xn[t] = a.xnorm[row];
    xid[t] = a.ids[row];
  }
  QFrag<KC> Q[NB];
  auto load_q = [&](QFrag<KC> &B, uint32_t q) {
    B.q = q;
    const int8_t *src = a.qhi + static_cast<size_t>(q) * a.dc + 16 * g16;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) B.A[kc] = load16(src + 64 * kc);
    B.qn = a.qnorm[q];
    B.lb = a.lbound[q];
  };
  uint32_t qnext = qfirst[NB - 1];
#pragma unroll
  for (int u = 0; u + 1 < NB; ++u) load_q(Q[u], qfirst[u]);
  v4i seed[T];
  uint32_t vmask, oddm;
  row_seeds<T>(xn, rows, seed, vmask, oddm);
  // fragment f from buffer C; fragment f + NB - 1's codes into buffer N
  auto step = [&](uint32_t f, const QFrag<KC> &C, QFrag<KC> &N) {
    const uint32_t qa = qnext;
    qnext = qid_of(f + NB);
    load_q(N, qa);
    // lane c16's query: H = floor((qn - L) / 2), or 2^30 (above every
    // accumulator) past the item
    const int32_t H = 16 * f + c16 < nq ? (C.qn - C.lb) >> 1 : (1 << 30);
    v4i acc[T];
#pragma unroll
    for (int t = 0; t < T; ++t) acc[t] = mfma(C.A[0], b[t * KC], seed[t]);
#pragma unroll
    for (int kc = 1; kc < KC; ++kc)
#pragma unroll
      for (int t = 0; t < T; ++t) acc[t] = mfma(C.A[kc], b[t * KC + kc], acc[t]);
    // H of the lane's 4 query slots 4 g16 + r
    int32_t H4[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) H4[r] = __shfl(H, 4 * g16 + r);
    if (!fragment_passes<T>(acc, H4)) return;
    const uint32_t qid = C.q;
    const int32_t qn = C.qn;
    n_kept += keep_runs<T>(
        rs, a, acc, vmask, oddm, xid, H4, [&](int j) { return static_cast<uint32_t>(__shfl(static_cast<int>(qid), j)); },
        [&](int j) { return __shfl(qn, j); });
  };
  for (uint32_t f0 = 0; f0 < nf; f0 += NB) {
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      if (f0 + u >= nf) break;
      step(f0 + u, Q[u], Q[(u + NB - 1) % NB]);
    }
  }
}

template <int KC, int T, int NB>
__global__ __launch_bounds__(256, (T * KC > 32 ? 1 : 2)) void k_rank_wave(WaveArgs a) {
  __shared__ uint64_t s_key[4][RS_KEYS];
  __shared__ uint16_t s_run[4][RS_KEYS];
  __shared__ uint32_t s_rq[4][RS_RUNS], s_rfirst[4][RS_RUNS], s_rbase[4][RS_RUNS];
  const int lane = lane_id(), w = threadIdx.x >> 6;
  RunStage rs{s_key[w], s_run[w], s_rq[w], s_rfirst[w], s_rbase[w], 0u, 0u};
  const uint32_t ib = a.item_off[a.band * a.nlist], ie = a.item_off[(a.band + 1) * a.nlist];
  uint32_t n_kept = 0;
  // The band's items come in cost order (most query fragments first) and are
  // dealt to the grid's waves in rounds of one item per wave, alternately
  // forward and backward (round r: item r nw + j to wave j, or to wave nw - 1 -
  // j), so every wave draws about the same work (a shared queue counter
  // serialises every taker on one address). Each wave reads its next item's
  // descriptor through the scalar cache while it ranks the current one.
  const uint32_t nw = gridDim.x * 4, wg = blockIdx.x * 4 + static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(w));
  auto item_at = [&](uint32_t r) { return ib + r * nw + ((r & 1u) ? nw - 1u - wg : wg); };
  uint32_t r = 0, it = item_at(0);
  uint4 dn = it < ie ? item_desc(a.items, it) : uint4{0, 0, 0, 0};
  while (it < ie) {
    const uint4 d = dn;
    const uint32_t nx = item_at(r + 1);
    if (nx < ie) dn = item_desc(a.items, nx);
    wave_item<KC, T, NB>(d, a, rs, n_kept);
    ++r;
    it = nx;
  }
  if (rs.n) runs_flush(rs, a);
  if (lane == 0 && n_kept) atomicAdd(&a.state[1], n_kept);
}

// --------------------------------------------------------------------------
// Ranking pass of one probe-rank band, query-stationary form. A wave takes
// items (list, up to 16 QF of the band's queries of the list, a row range),
// dealt round-robin, keeps the item's queries in registers as MFMA A operands
// (QF fragments of 16) and streams the item's rows through 16 at a time as B
// operands, NR row tiles in flight: every row byte comes from HBM once per
// item, and one item holds all the band's queries of a list that has at most
// 16 QF of them. The filter, exact test and pool appends are k_rank_wave's
// (one 16-row tile per step).
template <int KC>
struct RowTile {
  v4i b[KC];
  int32_t xn;
  uint32_t id;
};

// Stage of a query-stationary wave: the kept keys of its current item in LDS,
// each with its query slot (0 .. 16 QF - 1, slot s = lane s's query) and its
// index among the slot's staged keys, and per slot the count staged so far.
// The queries stay the same for the whole item, so the stage is flushed (one
// returning atomic per slot reserves its run in the query's pool, then the
// keys are written) only when a fragment's keys would not fit, and at the end
// of the item: the row stream carries no atomic whose return it waits for.
constexpr uint32_t QS_CAP = 512;

struct QsStage {
  uint64_t *key;   // QS_CAP
  uint32_t *meta;  // QS_CAP: slot | index << 6
  uint32_t *cnt;   // 64: keys staged per slot
  uint32_t n;      // keys staged (wave-uniform)
};

// Lane s (< 64) holds the query of slot s (slotq) and whether the slot is in
// the item.
__device__ __forceinline__ void qs_flush(QsStage &st, const WaveArgs &a, uint32_t slotq, bool slotok) {
  const int lane = lane_id();
  wave_fence();
  const uint32_t c = st.cnt[lane];
  uint32_t base = 0;
  if (slotok && c) base = atomicAdd(&a.pool_cnt[static_cast<size_t>(slotq) * katoann_hip::kCntStride], c);
  bool ovf = false;
  for (uint32_t e0 = 0; e0 < st.n; e0 += 64) {
    const uint32_t e = e0 + lane;
    const bool v = e < st.n;
    const uint32_t m = v ? st.meta[e] : 0u;
    const uint64_t key = v ? st.key[e] : 0ull;
    const int slot = static_cast<int>(m & 63u);
    const uint32_t pos = __shfl(base, slot) + (m >> 6);
    const uint32_t q = __shfl(slotq, slot);
    if (v) {
      if (pos < a.stride) a.pool[static_cast<size_t>(q) * a.stride + pos] = key;
      else ovf = true;
    }
  }
  if (ovf) atomicOr(&a.state[0], 1u);
  st.cnt[lane] = 0;
  st.n = 0;
  wave_fence();
}

// Stages the kept keys of fragment f (query slots 16 f .. 16 f + 15) of one
// 16-row tile: lane c16's row, slots 4 g16 + r. Returns the keys staged.
__device__ __forceinline__ uint32_t qs_keep(QsStage &st, const WaveArgs &a, const v4i &acc, int f, bool rowok,
                                            uint32_t odd, uint32_t id, const int32_t (&H4)[4], int32_t qnf,
                                            uint32_t slotq, bool slotok) {
  const int lane = lane_id(), g16 = lane >> 4, c16 = lane & 15;
  int32_t qnr[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) qnr[r] = __shfl(qnf, 4 * g16 + r);
  uint32_t c = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) c |= static_cast<uint32_t>(rowok && acc[r] >= H4[r]) << (8 * r);
  // per 16-lane group g, byte r: keys of slot 4 g + r up to this lane
  const uint32_t inc = row_scan16(c), ex = inc - c;
  const uint32_t G0 = __builtin_amdgcn_readlane(inc, 15), G1 = __builtin_amdgcn_readlane(inc, 31);
  const uint32_t G2 = __builtin_amdgcn_readlane(inc, 47), G3 = __builtin_amdgcn_readlane(inc, 63);
  const uint32_t S0 = byte_sum(G0), S1 = byte_sum(G1), S2 = byte_sum(G2), S3 = byte_sum(G3);
  const uint32_t total = S0 + S1 + S2 + S3;
  if (total == 0) return 0;
  if (st.n + total > QS_CAP) qs_flush(st, a, slotq, slotok);
  const uint32_t Gm = g16 == 0 ? G0 : g16 == 1 ? G1 : g16 == 2 ? G2 : G3;
  uint32_t so = st.n + (g16 > 0 ? S0 : 0u) + (g16 > 1 ? S1 : 0u) + (g16 > 2 ? S2 : 0u);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t slot = 16u * f + 4 * g16 + r;
    const uint32_t prev = st.cnt[slot];  // keys of the slot staged before this fragment
    if ((c >> (8 * r)) & 1u) {
      const uint32_t idx = (ex >> (8 * r)) & 0xffu;
      const int32_t l2 = qnr[r] - 2 * acc[r] - static_cast<int32_t>(odd);
      st.key[so + idx] = make_key(fscale(l2, a.step_sq), id);
      st.meta[so + idx] = slot | ((prev + idx) << 6);
    }
    so += (Gm >> (8 * r)) & 0xffu;
  }
  wave_fence();
  // slot counts: lane c16 < 4 of group g16 adds slot 4 g16 + c16's keys
  if (c16 < 4) {
    const uint32_t slot = 16u * f + 4 * g16 + c16;
    st.cnt[slot] += (Gm >> (8 * c16)) & 0xffu;
  }
  st.n += total;
  wave_fence();
  return total;
}

// One item of the query-stationary form (queries in registers, rows streamed).
template <int KC, int QF, int NR>
__device__ __forceinline__ void qs_item(const uint4 d, const WaveArgs &a, QsStage &st, uint32_t &n_kept) {
  constexpr uint32_t P = 64 * KC;
  const int lane = lane_id(), g16 = lane >> 4, c16 = lane & 15;
  const uint32_t nq = d.y, rows = d.w - d.z, ntile = (rows + 15) / 16, nf = (nq + 15) / 16;
  const uint32_t *ql = a.qlist + d.x;
  // tiles past the item load its last row again (unconditional loads)
  auto load_rows = [&](RowTile<KC> &r, uint32_t t) {
    const uint32_t row = d.z + min(16 * t + c16, rows - 1);
    const int8_t *rp = a.plane + static_cast<size_t>(row) * P + 16 * g16;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) r.b[kc] = load16(rp + 64 * kc);
    r.xn = a.xnorm[row];
    r.id = a.ids[row];
  };
  // the item's queries: lane c16 of fragment f holds query 16 f + c16 (the
  // last one again past the item). Their ids are loaded first, then the
  // first row tiles, then the codes that need the ids: the ids' round trip
  // overlaps the rows', and the codes' wait does not wait for the rows.
  uint32_t qid[QF];
#pragma unroll
  for (int f = 0; f < QF; ++f) qid[f] = ql[min(16u * f + c16, nq - 1)];
  // lane s: the query of slot s (and whether slot s is in the item)
  const uint32_t slotq = ql[min(static_cast<uint32_t>(lane), nq - 1)];
  const bool slotok = static_cast<uint32_t>(lane) < min(nq, 16u * QF);
  RowTile<KC> R[NR];
#pragma unroll
  for (int u = 0; u + 1 < NR; ++u) load_rows(R[u], u);
  v4i A[QF][KC];
  int32_t qnv[QF], lbv[QF], H4[QF][4];
#pragma unroll
  for (int f = 0; f < QF; ++f) {
    const int8_t *src = a.qhi + static_cast<size_t>(qid[f]) * a.dc + 16 * g16;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) A[f][kc] = load16(src + 64 * kc);
    qnv[f] = a.qnorm[qid[f]];
    lbv[f] = a.lbound[qid[f]];
  }
#pragma unroll
  for (int f = 0; f < QF; ++f) {
    // H = floor((qn - L) / 2), or 2^30 (above every accumulator) past the item
    const int32_t H = 16u * f + c16 < nq ? (qnv[f] - lbv[f]) >> 1 : (1 << 30);
#pragma unroll
    for (int r = 0; r < 4; ++r) H4[f][r] = __shfl(H, 4 * g16 + r);
  }
  for (uint32_t t0 = 0; t0 < ntile; t0 += NR) {
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      const uint32_t t = t0 + u;
      if (t >= ntile) break;
      load_rows(R[(u + NR - 1) % NR], t + NR - 1);
      const RowTile<KC> &C = R[u];
      const int32_t sd = -((C.xn + 1) >> 1);
      const v4i seed = {sd, sd, sd, sd};
      const bool rowok = 16 * t + c16 < rows;
      const uint32_t odd = static_cast<uint32_t>(C.xn & 1);
      v4i acc[QF][1];
#pragma unroll
      for (int f = 0; f < QF; ++f) {
        if (static_cast<uint32_t>(f) >= nf) break;  // fragments past the item's queries
        acc[f][0] = mfma(A[f][0], C.b[0], seed);
#pragma unroll
        for (int kc = 1; kc < KC; ++kc) acc[f][0] = mfma(A[f][kc], C.b[kc], acc[f][0]);
      }
#pragma unroll
      for (int f = 0; f < QF; ++f) {
        if (static_cast<uint32_t>(f) >= nf) break;
        if (!fragment_passes<1>(acc[f], H4[f])) continue;
        n_kept += qs_keep(st, a, acc[f][0], f, rowok, odd, C.id, H4[f], qnv[f], slotq, slotok);
      }
    }
  }
  if (st.n) qs_flush(st, a, slotq, slotok);
}

template <int KC, int QF, int NR>
__global__ __launch_bounds__(256, 2) void k_rank_qs(WaveArgs a) {
  __shared__ uint64_t s_key[4][QS_CAP];
  __shared__ uint32_t s_meta[4][QS_CAP];
  __shared__ uint32_t s_cnt[4][64];
  const int lane = lane_id(), w = threadIdx.x >> 6;
  QsStage st{s_key[w], s_meta[w], s_cnt[w], 0u};
  st.cnt[lane] = 0;
  wave_fence();
  const uint32_t ib = a.item_off[a.band * a.nlist], ie = a.item_off[(a.band + 1) * a.nlist];
  uint32_t n_kept = 0;
  const uint32_t nw = gridDim.x * 4;
  for (uint32_t it = ib + blockIdx.x * 4 + w; it < ie; it += nw) qs_item<KC, QF, NR>(a.items[it], a, st, n_kept);
  if (lane == 0 && n_kept) atomicAdd(&a.state[1], n_kept);
}

// Both forms in one launch: an item whose query count carries kQsItem (its
// list has at most 16 QF of the band's queries) runs query-stationary, the
// others with the list's rows in registers and its queries streamed, so a
// list's rows come from HBM once either way.
template <int KC, int T, int NB, int QF, int NR>
__global__ __launch_bounds__(256, 2) void k_rank_hyb(WaveArgs a) {
  __shared__ uint64_t s_key[4][RS_KEYS];
  __shared__ uint16_t s_run[4][RS_KEYS];
  __shared__ uint32_t s_rq[4][RS_RUNS], s_rfirst[4][RS_RUNS], s_rbase[4][RS_RUNS];
  __shared__ uint64_t q_key[4][QS_CAP];
  __shared__ uint32_t q_meta[4][QS_CAP];
  __shared__ uint32_t q_cnt[4][64];
  const int lane = lane_id(), w = threadIdx.x >> 6;
  RunStage rs{s_key[w], s_run[w], s_rq[w], s_rfirst[w], s_rbase[w], 0u, 0u};
  QsStage st{q_key[w], q_meta[w], q_cnt[w], 0u};
  st.cnt[lane] = 0;
  wave_fence();
  const uint32_t ib = a.item_off[a.band * a.nlist], ie = a.item_off[(a.band + 1) * a.nlist];
  uint32_t n_kept = 0;
  const uint32_t nw = gridDim.x * 4;
  for (uint32_t it = ib + blockIdx.x * 4 + w; it < ie; it += nw) {
    uint4 d = a.items[it];
    if (d.y & katoann_hip::kQsItem) {
      d.y &= ~katoann_hip::kQsItem;
      qs_item<KC, QF, NR>(d, a, st, n_kept);
    } else {
      wave_item<KC, T, NB>(d, a, rs, n_kept);
    }
  }
  if (rs.n) runs_flush(rs, a);
  if (lane == 0 && n_kept) atomicAdd(&a.state[1], n_kept);
}

// --------------------------------------------------------------------------
// Ranking pass of one probe-rank band, block form with a query ring. A block
// of 4 waves takes items (list, up to 4 x 16 T rows, up to QC of the band's
// queries of the list), dealt round-robin; wave w keeps rows [16 T w, 16 T (w +
// 1)) of the item in registers as MFMA B operands. The item's queries stream
// through a two-stage LDS buffer, 16 per step: the codes of the fragment LA
// steps ahead are loaded into registers (each wave one 1 KB piece of every
// fragment, gathered by query id, ids loaded a further step ahead), written to
// the stage the step before use, and read by all four waves, so every query
// byte crosses L2 once per block and the gather latency hides behind LA - 1
// steps of MFMAs. The hot and exact tests and the pool appends are those of
// k_rank_wave.
template <int KC, int T>
__global__ __launch_bounds__(256, 2) void k_rank_ring(WaveArgs a) {
  constexpr uint32_t P = 64 * KC;
  constexpr int LA = 4;              // fragments loaded ahead of use
  constexpr int NR = LA - 1;         // register ring: fragments g + 1 .. g + LA - 1
  constexpr int NPW = (KC + 3) / 4;  // 1 KB pieces per wave per fragment
  __shared__ __attribute__((aligned(16))) int8_t sA[2][KC * 1024];
  __shared__ __attribute__((aligned(16))) int32_t sH[2][16];
  __shared__ int32_t sN[2][16];
  __shared__ uint32_t sQ[2][16];
  __shared__ uint64_t s_key[4][2 * KCAP];
  __shared__ uint32_t s_meta[4][2 * KCAP];
  const int lane = lane_id(), w = threadIdx.x >> 6, g16 = lane >> 4, c16 = lane & 15;
  KeepStage ks{s_key[w], s_meta[w], 0u, 0, 0u, 0u};
  const uint32_t ib = a.item_off[a.band * a.nlist], ie = a.item_off[(a.band + 1) * a.nlist];
  uint32_t n_kept = 0;
  for (uint32_t it = ib + blockIdx.x; it < ie; it += gridDim.x) {
    const uint4 d = a.items[it];
    const uint32_t nq = d.y, nf = (nq + 15) / 16;
    const uint32_t *ql = a.qlist + d.x;
    const uint32_t r0 = d.z + static_cast<uint32_t>(w) * 16 * T;
    const uint32_t rows = r0 < d.w ? min(16u * T, d.w - r0) : 0u;
    // the wave's rows: B operands, norms and ids (a wave past the list's end
    // reads the last row and keeps nothing)
    v4i b[T * KC];
    int32_t xn[T];
    uint32_t xid[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const uint32_t row = min(r0 + 16 * t + c16, d.w - 1);
      const int8_t *rp = a.plane + static_cast<size_t>(row) * P + 16 * g16;
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) b[t * KC + kc] = load16(rp + 64 * kc);
      xn[t] = a.xnorm[row];
      xid[t] = a.ids[row];
    }
    v4i seed[T];
    uint32_t vmask, oddm;
    row_seeds<T>(xn, rows, seed, vmask, oddm);
    // producer: this wave's pieces of fragment f (query id q of lane c16),
    // and for wave 0 the fragment's norms and bounds
    auto qid_of = [&](uint32_t f) { return ql[min(16 * f + c16, nq - 1)]; };
    struct Frag {
      v4i p[NPW];
      int32_t qn, lb;
      uint32_t q;
    };
    auto load_frag = [&](uint32_t q) {
      Frag r;
      const int8_t *src = a.qhi + static_cast<size_t>(q) * a.dc + 16 * g16;
#pragma unroll
      for (int j = 0; j < NPW; ++j) {
        const int pc = w + 4 * j;
        r.p[j] = pc < KC ? load16(src + 64 * pc) : v4i{0, 0, 0, 0};
      }
      r.q = q;
      r.qn = w == 0 ? a.qnorm[q] : 0;
      r.lb = w == 0 ? a.lbound[q] : 0;
      return r;
    };
    auto store_frag = [&](const Frag &r, int st) {
#pragma unroll
      for (int j = 0; j < NPW; ++j) {
        const int pc = w + 4 * j;
        if (pc < KC) *reinterpret_cast<v4i *>(&sA[st][pc * 1024 + lane * 16]) = r.p[j];
      }
      if (w == 0 && lane < 16) {
        sH[st][lane] = r.qn - r.lb;
        sN[st][lane] = r.qn;
        sQ[st][lane] = r.q;
      }
    };
    // prologue: fragment 0 to stage 0, fragments 1 .. LA - 1 into the ring
    Frag ring[NR];
    uint32_t qnext = 0;
    {
      uint32_t q0s[LA];
#pragma unroll
      for (int i = 0; i < LA; ++i) q0s[i] = qid_of(min(static_cast<uint32_t>(i), nf - 1));
      qnext = qid_of(min(static_cast<uint32_t>(LA), nf - 1));
      const Frag f0 = load_frag(q0s[0]);
#pragma unroll
      for (int i = 0; i < NR; ++i) ring[i] = load_frag(q0s[i + 1]);
      store_frag(f0, 0);
    }
    __syncthreads();
    for (uint32_t g = 0; g < nf; ++g) {
      const int st = g & 1;
      // ahead: fragment g + LA's codes (ids loaded last step), g + LA + 1's ids
      Frag fn = ring[NR - 1];
      if (g + LA < nf) fn = load_frag(qnext);
      if (g + LA + 1 < nf) qnext = qid_of(g + LA + 1);
      if (rows) {
        v4i A[KC];
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) A[kc] = *reinterpret_cast<const v4i *>(&sA[st][kc * 1024 + lane * 16]);
        const v4i hv = *reinterpret_cast<const v4i *>(&sH[st][4 * g16]);
        v4i acc[T];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = mfma(A[0], b[t * KC], seed[t]);
#pragma unroll
        for (int kc = 1; kc < KC; ++kc)
#pragma unroll
          for (int t = 0; t < T; ++t) acc[t] = mfma(A[kc], b[t * KC + kc], acc[t]);
        int32_t H4[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) H4[r] = 16 * g + 4 * g16 + r < nq ? hv[r] >> 1 : (1 << 30);
        if (fragment_passes<T>(acc, H4))
          n_kept += keep_exact<T>(ks, a, acc, vmask, oddm, xid, H4, [&](int j) { return sQ[st][j]; },
                                  [&](int j) { return sN[st][j]; });
      }
      // fragment g + 1 into the other stage, then the ring moves on
      if (g + 1 < nf) store_frag(ring[0], st ^ 1);
#pragma unroll
      for (int i = 0; i + 1 < NR; ++i) ring[i] = ring[i + 1];
      ring[NR - 1] = fn;
      __syncthreads();
    }
  }
  keep_drain(ks, a);
  if (lane == 0 && n_kept) atomicAdd(&a.state[1], n_kept);
}

// --------------------------------------------------------------------------
// Ranking over the full 16-bit codes (Ranking::kFull, src/index.cpp:1130-1219):
// the exact distance of every candidate, the four byte products on MFMA,
//   dot' = 65536 (sh . sh) + 256 (sh . sl + sl . sh) + sl . sl,
// in three int32 accumulators (each |sum| <= 2 dc 128^2 < 2^31), combined in
// int64; key = (float bits of f(qn + xn - 2 dot')) << 32 | id. A key enters
// its query's pool iff it is below the query's threshold key (the S-th
// smallest of the pool so far, S = k), so each pool ends with the k smallest
// (float, id) keys of all the query's candidates.
// One wave per item (FULL_T x 16 rows, the item's queries 16 at a time), both
// operands streamed per 64-byte k-slice: the item's rows through L2 (read
// again for each query fragment), the queries from the batch's codes. Kept
// keys reserve their pool slots with one atomic per 16-lane group (the lanes
// of one query).
struct FullArgs {
  const int8_t *hi, *lo;     // N x dc, shifted bytes of the full codes
  const int64_t *xnorm;      // N: shifted norms of the full codes
  const uint32_t *ids;       // N
  const int8_t *qhi, *qlo;   // B x dc
  const int64_t *qnorm;      // B
  const uint64_t *thr;       // B: threshold key (all ones while the pool holds fewer than S)
  const uint32_t *qlist, *item_off;
  const uint4 *items;
  uint32_t nlist, band, dc, stride;
  uint64_t *pool;
  uint32_t *pool_cnt, *state;
  double step_sq;
};

constexpr int FULL_T = 4;

__global__ __launch_bounds__(256, 2) void k_rank_full(FullArgs a) {
  constexpr int T = FULL_T;
  const int lane = lane_id(), g16 = lane >> 4, c16 = lane & 15;
  const uint32_t ib = a.item_off[a.band * a.nlist], ie = a.item_off[(a.band + 1) * a.nlist];
  const uint32_t nw = gridDim.x * 4;
  const size_t dc = a.dc;
  uint32_t n_kept = 0;  // wave-uniform
  for (uint32_t it = ib + blockIdx.x * 4 + (threadIdx.x >> 6); it < ie; it += nw) {
    const uint4 d = a.items[it];
    const int8_t *xh[T], *xl[T];
    int64_t xn[T];
    uint32_t xid[T];
    bool rv[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const uint32_t row = d.z + 16 * t + c16;
      rv[t] = row < d.w;
      const size_t rr = min(row, d.w - 1);
      xh[t] = a.hi + rr * dc + 16 * g16;
      xl[t] = a.lo + rr * dc + 16 * g16;
      xn[t] = a.xnorm[rr];
      xid[t] = a.ids[rr];
    }
    for (uint32_t f = 0; 16 * f < d.y; ++f) {
      // A operand: query 16 f + (lane % 16), bytes 16 (lane / 16) of each k-slice
      const uint32_t qa = a.qlist[d.x + min(16 * f + c16, d.y - 1)];
      const int8_t *ph = a.qhi + qa * dc + 16 * g16, *pl = a.qlo + qa * dc + 16 * g16;
      v4i hh[T], cr[T], ll[T];
#pragma unroll
      for (int t = 0; t < T; ++t) hh[t] = cr[t] = ll[t] = v4i{0, 0, 0, 0};
#pragma unroll 2
      for (uint32_t k = 0; k < dc; k += 64) {
        const v4i qh = load16(ph + k), ql = load16(pl + k);
        v4i bh[T], bl[T];
#pragma unroll
        for (int t = 0; t < T; ++t) {
          bh[t] = load16(xh[t] + k);
          bl[t] = load16(xl[t] + k);
        }
#pragma unroll
        for (int t = 0; t < T; ++t) {
          hh[t] = mfma(qh, bh[t], hh[t]);
          cr[t] = mfma(qh, bl[t], cr[t]);
          cr[t] = mfma(ql, bh[t], cr[t]);
          ll[t] = mfma(ql, bl[t], ll[t]);
        }
      }
      // C[i][j] (query 16 f + i, row 16 t + j) sits in lane 16 (i / 4) + j,
      // element i % 4: the 16 lanes of a group share one query
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t qi = 16 * f + 4 * g16 + r;
        const uint32_t q = static_cast<uint32_t>(__shfl(static_cast<int>(qa), 4 * g16 + r));
        const bool qok = qi < d.y;
        const int64_t qn = a.qnorm[q];
        const uint64_t th = a.thr[q];
#pragma unroll
        for (int t = 0; t < T; ++t) {
          const int64_t dot = static_cast<int64_t>(hh[t][r]) * 65536 + static_cast<int64_t>(cr[t][r]) * 256 +
                              static_cast<int64_t>(ll[t][r]);
          const uint64_t key = make_key(fscale(qn + xn[t] - 2 * dot, a.step_sq), xid[t]);
          const bool keep = qok && rv[t] && key < th;
          const uint64_t m = ballot(keep);
          if (!m) continue;
          const uint32_t gm = static_cast<uint32_t>(m >> (16 * g16)) & 0xffffu;
          uint32_t base = 0;
          if (c16 == 0 && gm)
            base = atomicAdd(&a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride],
                             static_cast<uint32_t>(__popc(gm)));
          base = static_cast<uint32_t>(__shfl(static_cast<int>(base), 16 * g16));
          if (keep) {
            const uint32_t pos = base + static_cast<uint32_t>(__popc(gm & ((1u << c16) - 1u)));
            if (pos < a.stride) a.pool[static_cast<size_t>(q) * a.stride + pos] = key;
            else atomicOr(&a.state[0], 1u);
          }
          n_kept += static_cast<uint32_t>(__popcll(m));
        }
      }
    }
  }
  if (lane == 0 && n_kept) atomicAdd(&a.state[1], n_kept);
}

// --------------------------------------------------------------------------
// Ranking in bf16 precision (Precision::kBf16, src/index.cpp:1221-1293 with
// kernels.cpp:384-414), in the reference's AVX-512 arithmetic per
// (query, row): eight fp32 accumulators, pair k2 (bf16 elements 2 k2 and
// 2 k2 + 1) into accumulator k2 % 8 as _mm512_dpbf16_ps adds it (the odd
// element's exact product first, then the even one's, each sum rounded), then
// acc[g] += acc[g + s] for s = 4, 2, 1, and l2 = (qn + xn) - (dot + dot).
// Lane = (accumulator g = lane % 8, row slot s = lane / 8): a wave ranks 8 rows
// x 8 queries at a time, each lane running chain g of its row against the 8
// queries (the 8 slots read the same query dwords: one access); the butterfly
// over g (xor 4, 2, 1) is the reduction tree, after which lane (g, s) takes
// query g x row s. Keys are order-preserving float bits << 32 | id (distances
// can come out slightly negative); a key is kept iff below its query's
// threshold key, so each pool ends with the k nearest by (distance, id).
struct Bf16Args {
  const uint32_t *rows;   // N x dp / 2 dwords (bf16 pairs)
  const float *xnorm;     // N
  const uint32_t *ids;    // N
  const uint32_t *qbf;    // B x dp / 2
  const float *qnorm;     // B
  const uint64_t *thr;    // B
  const uint32_t *qlist, *item_off;
  const uint4 *items;
  uint32_t nlist, band, dp2, stride;
  uint64_t *pool;
  uint32_t *pool_cnt, *state;
};

constexpr uint32_t BF16_ROWS = 32;  // rows per item

__device__ __forceinline__ uint32_t ord_bits(float f) {
  const uint32_t b = __float_as_uint(f == 0.0f ? 0.0f : f);  // -0 ranks as +0
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float from_ord(uint32_t o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

__global__ __launch_bounds__(256) void k_rank_bf16(Bf16Args a) {
  const int lane = lane_id(), g = lane & 7, sl = lane >> 3;
  const uint32_t ib = a.item_off[a.band * a.nlist], ie = a.item_off[(a.band + 1) * a.nlist];
  const uint32_t nw = gridDim.x * 4;
  uint32_t n_kept = 0;  // wave-uniform
  for (uint32_t it = ib + blockIdx.x * 4 + (threadIdx.x >> 6); it < ie; it += nw) {
    const uint4 d = a.items[it];
    for (uint32_t r0 = d.z; r0 < d.w; r0 += 8) {
      const uint32_t row = r0 + sl;
      const bool rv = row < d.w;
      const uint32_t rr = min(row, d.w - 1);
      const uint32_t *xr = a.rows + static_cast<size_t>(rr) * a.dp2 + g;
      const float xn = a.xnorm[rr];
      const uint32_t xid = a.ids[rr];
      for (uint32_t q0 = 0; q0 < d.y; q0 += 8) {
        uint32_t qid[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) qid[i] = a.qlist[d.x + min(q0 + i, d.y - 1)];
        float acc[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = 0.0f;
        for (uint32_t k = 0; k < a.dp2; k += 8) {
          const uint32_t xw = xr[k];
          const float x0 = __uint_as_float(xw << 16), x1 = __uint_as_float(xw & 0xffff0000u);
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const uint32_t qw = a.qbf[static_cast<size_t>(qid[i]) * a.dp2 + g + k];
            acc[i] = fmaf(x1, __uint_as_float(qw & 0xffff0000u), acc[i]);
            acc[i] = fmaf(x0, __uint_as_float(qw << 16), acc[i]);
          }
        }
        float dot = 0.0f;
        uint32_t q = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          float v = acc[i];
          v = v + __shfl_xor(v, 4);
          v = v + __shfl_xor(v, 2);
          v = v + __shfl_xor(v, 1);
          dot = g == i ? v : dot;
          q = g == i ? qid[i] : q;
        }
        const float l2 = (a.qnorm[q] + xn) - (dot + dot);
        const uint64_t key = (static_cast<uint64_t>(ord_bits(l2)) << 32) | xid;
        const bool keep = rv && q0 + g < d.y && key < a.thr[q];
        if (keep) {
          const uint32_t pos = atomicAdd(&a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride], 1u);
          if (pos < a.stride) a.pool[static_cast<size_t>(q) * a.stride + pos] = key;
          else atomicOr(&a.state[0], 1u);
        }
        n_kept += static_cast<uint32_t>(__popcll(ballot(keep)));
      }
    }
  }
  if (lane == 0 && n_kept) atomicAdd(&a.state[1], n_kept);
}

// The k nearest of each pool after the last update (ascending, at most S = k
// keys): ids and distances, padded with UINT32_MAX / FLT_MAX
// (src/index.cpp:384-398); the survivors output as plain float bits << 32 | id.
__global__ __launch_bounds__(64) void k_emit_bf16(const uint64_t *__restrict__ pool, uint32_t stride,
                                                  const uint32_t *__restrict__ pool_cnt, uint32_t K,
                                                  uint32_t *__restrict__ ids, float *__restrict__ dists,
                                                  uint64_t *__restrict__ surv, uint32_t *__restrict__ nsurv) {
  const uint32_t q = blockIdx.x;
  const uint32_t n = min(pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride], min(stride, K));
  for (uint32_t i = threadIdx.x; i < K; i += 64) {
    const bool ok = i < n;
    const uint64_t key = ok ? pool[static_cast<size_t>(q) * stride + i] : ~0ull;
    const float f = from_ord(static_cast<uint32_t>(key >> 32));
    ids[static_cast<size_t>(q) * K + i] = ok ? static_cast<uint32_t>(key) : 0xffffffffu;
    dists[static_cast<size_t>(q) * K + i] = ok ? f : FLT_MAX;
    if (surv)
      surv[static_cast<size_t>(q) * K + i] =
          ok ? (static_cast<uint64_t>(__float_as_uint(f)) << 32) | static_cast<uint32_t>(key) : ~0ull;
  }
  if (nsurv && threadIdx.x == 0) nsurv[q] = n;
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

// One block of 4 waves per query: wave w merges batches w, w + 4, ... of the
// new keys into its own sorted list (wave 0 starts from the previous list);
// once a wave's list holds S keys, new keys above its S-th are dropped before
// any sorting. Wave 0 then merges the other three lists.
template <int SC>
__device__ inline void list_merge_from(TopList<SC> &t, const uint64_t *u, uint32_t S) {
  // t ascending, u ascending in LDS: t[i] = min(t[i], u[SC - 1 - i]) is a
  // bitonic sequence holding the SC smallest of both
  const int lane = lane_id();
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) t.e[r] = umin64(t.e[r], u[SC - 1 - (64 * r + lane)]);
  list_bitonic_merge<SC>(t);
#pragma unroll
  for (int r = 0; r < SC / 64; ++r)
    if (static_cast<uint32_t>(lane + 64 * r) >= S) t.e[r] = kNoKey;
}

constexpr uint32_t UPD_CAP = 1024;  // candidates of the selection path at most
constexpr int UPD_R = 16;           // pool keys per thread on the selection path (pools up to 4,096)

// Pools large enough for the selection path (k_update_sel); k_update merges
// the others (small pools merge faster: the batches past the first S keys are
// mostly dropped by the cap before any sorting).
__device__ __forceinline__ bool update_by_selection(uint32_t S, uint32_t n) {
  return S <= 256 && n > max(2 * S, 256u) && n <= 256u * UPD_R;
}

// Selection path (pools over max(2 S, 256) keys): at least S keys lie at or
// below the largest of the waves' ceil(S / 4)-th smallest thread minimum (one
// per lane; each lane's two smallest when ceil(S / 4) > 32, so that the wave
// picks among twice as many), and for pools in no particular order not many
// more;
// those candidates, sorted by the four waves, give the S smallest. A query
// with more than UPD_CAP candidates is left to k_update (run after it).
// Each thread's keys are loaded at once (independent loads) into registers;
// real keys are below kNoKey.
template <bool TWO>  // TWO: S > 128 (each lane's two smallest)
__global__ __launch_bounds__(256) void k_update_sel(UpdateArgs a) {
  __shared__ uint64_t skeys[2 * UPD_CAP];
  __shared__ uint64_t swk[4];
  __shared__ uint32_t stmp[4], stotal;
  const int lane = lane_id(), w = threadIdx.x >> 6, tid = threadIdx.x;
  const uint32_t q = blockIdx.x;
  const uint32_t n = min(a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride], a.stride);
  if (n == a.pool_sorted[q] || !update_by_selection(a.S, n)) return;
  uint64_t *pq = a.pool + static_cast<size_t>(q) * a.stride;
  uint64_t v[UPD_R];
#pragma unroll
  for (int r = 0; r < UPD_R; ++r) {
    const uint32_t i = tid + 256u * r;
    v[r] = i < n ? pq[i] : kNoKey;
  }
  // the wave's ceil(S / 4)-th smallest thread minimum (distance bits; among
  // the lanes' two smallest for K > 32), found bit by bit from the top with
  // one ballot per bit and value
  const uint32_t K = (a.S + 3) / 4;
  uint32_t mv = 0xffffffffu, mv2 = 0xffffffffu;
#pragma unroll
  for (int r = 0; r < UPD_R; ++r) {
    const uint32_t dv = static_cast<uint32_t>(v[r] >> 32);
    if (TWO) mv2 = min(mv2, max(mv, dv));
    mv = min(mv, dv);
  }
  uint32_t T = 0;
  for (int b = 31; b >= 0; --b) {
    const uint32_t hi0 = T | ((1u << b) - 1u);
    uint32_t c = static_cast<uint32_t>(__popcll(ballot(mv <= hi0)));
    if (TWO) c += static_cast<uint32_t>(__popcll(ballot(mv2 <= hi0)));
    if (c < K) T |= 1u << b;
  }
  const uint64_t wk = (static_cast<uint64_t>(T) << 32) | 0xffffffffull;
  if (lane == 0) swk[w] = wk;
  __syncthreads();
  const uint64_t tk = umax64(umax64(swk[0], swk[1]), umax64(swk[2], swk[3]));
  uint32_t cnt = 0;
#pragma unroll
  for (int r = 0; r < UPD_R; ++r) cnt += v[r] != kNoKey && v[r] <= tk ? 1u : 0u;
  uint32_t off = block_exscan256(cnt, stmp);
  if (tid == 255) stotal = off + cnt;
  __syncthreads();
  const uint32_t total = stotal;
  if (total > UPD_CAP) {
    // left for k_update: mark the query by a flag bit in pool_sorted
    if (tid == 0) a.pool_sorted[q] |= 0x80000000u;
    return;
  }
#pragma unroll
  for (int r = 0; r < UPD_R; ++r)
    if (v[r] != kNoKey && v[r] <= tk) skeys[off++] = v[r];
  uint32_t M = 64;
  while (M < total) M <<= 1;
  for (uint32_t i = total + tid; i < M; i += 256) skeys[i] = kNoKey;
  __syncthreads();
  const uint64_t *srt = block_sort_keys(skeys, M);
  for (uint32_t i = tid; i < a.S; i += 256) pq[i] = srt[i];
  if (tid == 0) {
    a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride] = a.S;
    a.pool_sorted[q] = a.S;
    const uint64_t kth = srt[a.S - 1];
    if (kth < a.thr[q]) {
      a.thr[q] = kth;
      a.lbound[q] = static_cast<int32_t>(l2_bound(key_float(kth), a.step_sq, a.lmax));
    }
  }
}

template <int SC>
__global__ __launch_bounds__(256, (SC == 512 ? 5 : 1)) void k_update(UpdateArgs a) {
  __shared__ uint64_t part[3][SC];
  const int lane = lane_id(), w = threadIdx.x >> 6;
  const uint32_t q = blockIdx.x;
  const uint32_t n = min(a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride], a.stride);
  uint32_t sorted = a.pool_sorted[q];
  if (sorted & 0x80000000u) {  // the selection kernel found too many candidates
    sorted &= 0x7fffffffu;
  } else if (n == sorted || update_by_selection(a.S, n)) {
    return;
  }
  uint64_t *pq = a.pool + static_cast<size_t>(q) * a.stride;
  TopList<SC> t;
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) {
    const uint32_t i = lane + 64 * r;
    t.e[r] = w == 0 && i < sorted ? pq[i] : kNoKey;
  }
  uint64_t cap = w == 0 && sorted >= a.S ? top_get<SC>(t, a.S - 1) : kNoKey;  // keys above it cannot stay
  // this wave's batches, the next one in flight
  uint32_t b = sorted + 64 * w;
  uint64_t kn = b + lane < n ? pq[b + lane] : kNoKey;
  for (; b < n; b += 256) {
    uint64_t k = kn;
    kn = b + 256 + lane < n ? pq[b + 256 + lane] : kNoKey;
    if (k >= cap) k = kNoKey;
    const uint64_t mk = ballot(k != kNoKey);
    if (!mk) continue;
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
    const uint64_t last = top_get<SC>(t, a.S - 1);
    if (last != kNoKey) cap = last;
  }
  if (w > 0) {
#pragma unroll
    for (int r = 0; r < SC / 64; ++r) part[w - 1][lane + 64 * r] = t.e[r];
  }
  __syncthreads();
  if (w > 0) return;
#pragma unroll
  for (int v = 0; v < 3; ++v) list_merge_from<SC>(t, part[v], a.S);
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


// Exact keys (f(l2) bits << 32 | id) of 16 survivors: lane i < 16 returns
// survivor i's, where every lane passes the key of survivor lane % 16 and
// whether it exists. The survivors' rows are gathered by id (A = 16 rows of
// one plane), the query's high and low bytes are the B operand's columns 0 and
// 1; the four byte products are combined in int64. A survivor whose id has no
// row gets kNoKey.
__device__ inline uint64_t exact_key16(const RefineArgs &a, uint64_t key, bool sok, const int8_t *qh, const int8_t *ql,
                                       int64_t qn) {
  const int lane = lane_id();
  const int koff = 16 * (lane >> 4), col = lane & 15;
  const uint32_t id = static_cast<uint32_t>(key);
  sok = sok && id < a.npts;
  const uint32_t row = sok ? a.id_row[id] : 0xffffffffu;
  const bool rok = sok && row != 0xffffffffu;
  v4i ah = {0, 0, 0, 0}, al = {0, 0, 0, 0};
  // RB chunks of 64 bytes in flight at once: the rows are random gathers
  constexpr int RB = 16;
  for (uint32_t kc0 = 0; kc0 < a.dc; kc0 += 64 * RB) {
    v4i xh[RB], xl[RB], bq[RB];
#pragma unroll
    for (int u = 0; u < RB; ++u) {
      const uint32_t kc = kc0 + 64 * u;
      const bool in = kc < a.dc;
      xh[u] = rok && in ? load16(a.hi + static_cast<size_t>(row) * a.dc + kc + koff) : v4i{0, 0, 0, 0};
      xl[u] = rok && in ? load16(a.lo + static_cast<size_t>(row) * a.dc + kc + koff) : v4i{0, 0, 0, 0};
      bq[u] = in && col == 0 ? load16(qh + kc + koff) : (in && col == 1 ? load16(ql + kc + koff) : v4i{0, 0, 0, 0});
    }
#pragma unroll
    for (int u = 0; u < RB; ++u) {
      ah = mfma(xh[u], bq[u], ah);
      al = mfma(xl[u], bq[u], al);
    }
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
  const bool mine_ok = __shfl(static_cast<int>(rok), i) != 0;
  const uint32_t myrow = __shfl(row, i);
  const uint32_t myid = __shfl(id, i);
  if (!mine_ok) return kNoKey;
  const int64_t dot =
      static_cast<int64_t>(hh) * 65536 + (static_cast<int64_t>(hl) + lh) * 256 + static_cast<int64_t>(ll);
  return make_key(fscale(qn + a.norm[myrow] - 2 * dot, a.step_sq), myid);
}

// Exact keys of one query's survivors ex[0, ns), one wave (the same MFMA
// operands and int64 combine as exact_key16). On entry ex[i] = row << 32 | id
// (row 0xffffffff: the id has no row), on return the exact key (kNoKey for no
// row). The rows stream through as steps of 16 survivors x RB chunks of 64
// bytes (both planes), RB dividing dc / 64, each step's gathers issued while
// the step before is multiplied, so a wave keeps a step of loads in flight
// from its first gather to its last; the survivors' norms travel with their
// steps. qb: the query's high then low bytes in LDS (dc each), zb: 16 zero
// bytes in LDS (the B operand's columns 2..15).
constexpr int RF_AHEAD = 1;  // gather steps in flight while one is multiplied

template <int RB>
__device__ inline void exact_keys(const RefineArgs &a, uint64_t *ex, uint32_t ns, const int8_t *qb, const int8_t *zb,
                                  int64_t qn) {
  const int lane = lane_id(), kq = lane >> 4, col = lane & 15;
  const uint32_t dc = a.dc, ncb = dc / (64 * RB), ng = (ns + 15) / 16, nsteps = ng * ncb;
  if (nsteps == 0) return;
  struct Step {
    v4i xh[RB], xl[RB];
    int64_t nrm;
    uint32_t id;
    bool ok;
  };
  // loads of step s (clamped: past the end it loads the last step again, so
  // every iteration issues the same loads)
  auto load = [&](Step &X, uint32_t s) {
    s = min(s, nsteps - 1);
    const uint32_t g = s / ncb, cb = s - g * ncb, i = 16 * g + col;
    const uint64_t e = ex[min(i, ns - 1)];
    const uint32_t row = static_cast<uint32_t>(e >> 32);
    X.ok = i < ns && row != 0xffffffffu;
    X.id = static_cast<uint32_t>(e);
    const size_t base = static_cast<size_t>(X.ok ? row : 0u) * dc + cb * (64u * RB) + 16u * kq;
#pragma unroll
    for (int u = 0; u < RB; ++u) {
      X.xh[u] = load16(a.hi + base + 64 * u);
      X.xl[u] = load16(a.lo + base + 64 * u);
    }
    X.nrm = a.norm[X.ok ? row : 0u];
  };
  v4i ah = {0, 0, 0, 0}, al = {0, 0, 0, 0};
  auto compute = [&](const Step &X, uint32_t s) {
    const uint32_t g = s / ncb, cb = s - g * ncb;
    if (cb == 0) ah = al = v4i{0, 0, 0, 0};
#pragma unroll
    for (int u = 0; u < RB; ++u) {
      const uint32_t kc = (cb * RB + u) * 64 + 16 * kq;
      const int8_t *bp = col == 0 ? qb + kc : col == 1 ? qb + dc + kc : zb;
      const v4i bq = *reinterpret_cast<const v4i *>(bp);
      ah = mfma(X.xh[u], bq, ah);
      al = mfma(X.xl[u], bq, al);
    }
    if (cb + 1 < ncb) return;
    // survivor i = lane (< 16): C[i][0] in lane 16 (i / 4), element i % 4; C[i][1] one lane up
    const int i = lane & 15, src = 16 * (i >> 2);
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
    // lane i < 16 holds survivor 16 g + i's id, norm and row flag (column i)
    const int64_t dot =
        static_cast<int64_t>(hh) * 65536 + (static_cast<int64_t>(hl) + lh) * 256 + static_cast<int64_t>(ll);
    const uint64_t key = X.ok ? make_key(fscale(qn + X.nrm - 2 * dot, a.step_sq), X.id) : kNoKey;
    if (lane < 16 && 16 * g + lane < ns) ex[16 * g + lane] = key;
  };
  Step X0, X1;
  load(X0, 0);
  if constexpr (RF_AHEAD == 1) {
    for (uint32_t s = 0; s < nsteps; s += 2) {
      load(X1, s + 1);
      compute(X0, s);
      if (s + 1 >= nsteps) break;
      load(X0, s + 2);
      compute(X1, s + 1);
    }
  } else {  // two steps ahead
    Step X2;
    load(X1, 1);
    for (uint32_t s = 0; s < nsteps; s += 3) {
      load(X2, s + 2);
      compute(X0, s);
      if (s + 1 >= nsteps) break;
      load(X0, s + 3);
      compute(X1, s + 1);
      if (s + 2 >= nsteps) break;
      load(X1, s + 4);
      compute(X2, s + 2);
    }
  }
}

// Final survivors, exact distances, arrival rule, output (src/index.cpp:1046-1127):
//  1. the S smallest keys of the query's pool (exactly the S smallest ranking
//     keys of all its candidates);
//  2. exact distances of the survivors: rows gathered by id, the four byte
//     products on MFMA (A = 16 survivor rows of one plane, B = the query's high
//     and low bytes as columns 0 and 1), combined in int64;
//  3. the reference's arrival rule: survivors in descending (rank, id) order
//     enter a k-entry max-heap of (distance, id) while it has room, then
//     replace its maximum iff their distance is strictly smaller (evaluated in
//     closed form, below);
//  4. output ascending by (distance, id), padded with FLT_MAX / UINT32_MAX
//     (src/index.cpp:384-398).
constexpr uint32_t RF_MAXDC = 4096;  // code bytes per row at most (dimensions up to 4,096)

template <int SC, int RB>
__global__ __launch_bounds__(256, 2) void k_refine(RefineArgs a) {
  __shared__ uint64_t ex[4][SC];
  __shared__ __attribute__((aligned(16))) int8_t qb[4][2 * RF_MAXDC + 16];
  const int lane = lane_id(), w = threadIdx.x >> 6;
  const uint32_t q = blockIdx.x * 4 + w;
  if (q >= a.B) return;
  TopList<SC> sv;
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) sv.e[r] = kNoKey;
  {
    const uint64_t base = static_cast<uint64_t>(q) * a.stride;
    const uint32_t m = min(a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride], a.stride);
    // keys above the threshold of the last update, or above the list's S-th
    // once it is full, cannot be among the S smallest (the pool is not
    // updated after the last band: it holds the S sorted survivors of the
    // earlier bands, then the last band's keys)
    uint64_t tq = a.thr[q];
    // Keys below the filter wait in the wave's LDS area (ex[w], free until the
    // survivors are known) until 64 have gathered, and are merged into the list
    // 64 at a time: a merge is a chain of dependent lane exchanges, and most
    // chunks pass only a few keys once the filter has tightened.
    uint64_t *pend = ex[w];
    uint32_t pc = 0;  // keys waiting (wave-uniform)
    const uint64_t below = (1ull << lane) - 1ull;
    auto chunk = [&](uint64_t k) {
      if (k > tq) k = kNoKey;
      const uint64_t mk = ballot(k != kNoKey);
      if (!mk) return;
      const uint32_t nk = static_cast<uint32_t>(__popcll(mk));
      const uint32_t at = pc + static_cast<uint32_t>(__popcll(mk & below));
      if (k != kNoKey && at < 64) pend[at] = k;
      if (pc + nk < 64) {
        pc += nk;
        return;
      }
      wave_fence();
      top_merge<SC>(sv, pend[lane], a.S);
      tq = umin64(tq, top_get<SC>(sv, a.S - 1));
      wave_fence();
      if (k != kNoKey && at >= 64) pend[at - 64] = k;
      pc = pc + nk - 64;
      wave_fence();
    };
    // The chunks holding the earlier survivors first, then the last band's
    // keys in four interleaved passes (every fourth chunk): the filter
    // tightens on a sample spread over the whole band, whatever order the
    // ranking kernel appended its keys in. Chunks are loaded eight at a time
    // (unconditional loads, clamped into the pool) so that their round trips
    // overlap.
    const uint32_t nch = (m + 63) / 64, c0 = min(nch, (a.S + 63) / 64);
    auto group = [&](uint32_t cfirst, uint32_t step, uint32_t cend) {
      for (uint32_t cb = cfirst; cb < cend; cb += 8 * step) {
        uint64_t kk[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
          const uint32_t i = 64 * (cb + u * step) + lane;
          kk[u] = a.pool[base + min(i, m - 1)];
        }
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
          const uint32_t c = cb + u * step;
          if (c < cend) chunk(64 * c + lane < m ? kk[u] : kNoKey);
        }
      }
    };
    group(0, 1, c0);
    for (uint32_t p = 0; p < 4; ++p) group(c0 + p, 4, nch);
    if (pc) {
      wave_fence();
      top_merge<SC>(sv, static_cast<uint32_t>(lane) < pc ? pend[lane] : kNoKey, a.S);
    }
    wave_fence();
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

  // the query's bytes and a zero block in LDS (the exact products' B operand)
  {
    const int8_t *qh = a.qhi + static_cast<size_t>(q) * a.dc;
    const int8_t *ql = a.qlo + static_cast<size_t>(q) * a.dc;
    for (uint32_t o = 16 * lane; o < a.dc; o += 1024) {
      *reinterpret_cast<v4i *>(&qb[w][o]) = load16(qh + o);
      *reinterpret_cast<v4i *>(&qb[w][a.dc + o]) = load16(ql + o);
    }
    if (lane == 0) *reinterpret_cast<v4i *>(&qb[w][2 * RF_MAXDC]) = v4i{0, 0, 0, 0};
  }
  // each survivor's row << 32 | id, in survivor order
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) {
    const uint32_t i = lane + 64 * r;
    if (i < ns) {
      const uint32_t id = static_cast<uint32_t>(sv.e[r]);
      const uint32_t row = id < a.npts ? a.id_row[id] : 0xffffffffu;
      ex[w][i] = static_cast<uint64_t>(row) << 32 | id;
    }
  }
  wave_fence();
  exact_keys<RB>(a, ex[w], ns, qb[w], &qb[w][2 * RF_MAXDC], a.qnorm[q]);
  wave_fence();

  // The arrival rule in closed form. Survivors arrive in descending rank order
  // (here: index ns - 1 first) and enter a K-entry max-heap of (dist, id) that
  // admits a new one while it has room, else iff its dist is strictly below
  // the heap's largest. The heap's dists are always the K smallest of those
  // arrived so far, so with M = the K-th smallest dist of all survivors the
  // final heap is: every survivor with dist < M, and of those with dist == M,
  // the ones admitted (fewer than K earlier arrivals with dist <= M) minus the
  // largest ids evicted once the heap's largest dist is M, i.e. the admitted
  // ones with the smallest ids. Sorting that set by (dist, id) and taking K
  // entries gives exactly the heap's output order.
  TopList<SC> e;
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) {
    const uint32_t i = lane + 64 * r;
    e.e[r] = i < ns ? ex[w][i] : kNoKey;
  }
  TopList<SC> srt = e;
  list_sort<SC>(srt);
  if (top_count<SC>(srt) > a.K) {
    const uint32_t mb = static_cast<uint32_t>(top_get<SC>(srt, a.K - 1) >> 32);  // bits of M (dists >= 0)
    uint32_t above = 0;  // survivors with dist <= M in higher registers (earlier arrivals)
    const uint64_t higher = lane == 63 ? 0ull : ~((2ull << lane) - 1ull);
#pragma unroll
    for (int r = SC / 64 - 1; r >= 0; --r) {
      const uint64_t k = e.e[r];
      const uint32_t fb = static_cast<uint32_t>(k >> 32);
      const bool valid = k != kNoKey;
      const uint64_t le = ballot(valid && fb <= mb);
      const uint32_t pre = above + static_cast<uint32_t>(__popcll(le & higher));
      const bool in = valid && (fb < mb || (fb == mb && pre < a.K));
      e.e[r] = in ? k : kNoKey;
      above += static_cast<uint32_t>(__popcll(le));
    }
    list_sort<SC>(e);
  } else {
    e = srt;
  }
#pragma unroll
  for (int r = 0; r < SC / 64; ++r) {
    const uint32_t g = lane + 64 * r;
    if (g < a.K) {
      const bool ok = e.e[r] != kNoKey;
      a.out_ids[static_cast<size_t>(q) * a.K + g] = ok ? static_cast<uint32_t>(e.e[r]) : 0xffffffffu;
      a.out_dists[static_cast<size_t>(q) * a.K + g] = ok ? key_float(e.e[r]) : FLT_MAX;
    }
  }
}

// ---- survivor counts above 1,024 (up to BIG_SB): one block per query, lists in LDS
constexpr uint32_t BIG_SB = 4096;
constexpr size_t BIG_LDS = 3 * BIG_SB * sizeof(uint64_t);  // one list and the sort's two buffers

// Ascending sort of a bitonic sequence a[0, M) in LDS (256 threads).
__device__ inline void block_bitonic_merge(uint64_t *a, uint32_t M) {
  for (uint32_t j = M >> 1; j > 0; j >>= 1) {
    for (uint32_t p = threadIdx.x; p < M / 2; p += 256) {
      const uint32_t i = (p / j) * 2 * j + p % j;
      const uint64_t x = a[i], y = a[i + j];
      a[i] = umin64(x, y);
      a[i + j] = umax64(x, y);
    }
    __syncthreads();
  }
}

// Keys below kNoKey in the ascending list a[0, n): n minus the first kNoKey's index.
__device__ inline uint32_t block_count_sorted(const uint64_t *a, uint32_t n, uint32_t *tmp) {
  uint32_t c = 0;
  for (uint32_t i = threadIdx.x; i < n; i += 256) c += a[i] != kNoKey ? 1u : 0u;
  const uint32_t off = block_exscan256(c, tmp);
  __shared__ uint32_t total;
  if (threadIdx.x == 255) total = off + c;
  __syncthreads();
  const uint32_t t = total;
  __syncthreads();
  return t;
}

// k_update for S > 1024: the previous sorted list, then the new keys in
// chunks of BIG_SB, each chunk sorted and merged in (the BIG_SB smallest of
// list and chunk by one bitonic merge, entries past S dropped); keys at or
// above the list's S-th are dropped before sorting.
__global__ __launch_bounds__(256) void k_update_big(UpdateArgs a) {
  extern __shared__ uint64_t lds[];
  __shared__ uint32_t stmp[4];
  uint64_t *cur = lds, *buf = lds + BIG_SB;
  const uint32_t q = blockIdx.x, tid = threadIdx.x;
  const uint32_t n = min(a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride], a.stride);
  const uint32_t sorted = a.pool_sorted[q] & 0x7fffffffu;
  if (n == sorted) return;
  uint64_t *pq = a.pool + static_cast<size_t>(q) * a.stride;
  for (uint32_t i = tid; i < BIG_SB; i += 256) cur[i] = i < sorted ? pq[i] : kNoKey;
  __syncthreads();
  uint64_t cap = cur[a.S - 1];  // kNoKey until the list holds S keys
  for (uint32_t b0 = sorted; b0 < n; b0 += BIG_SB) {
    for (uint32_t i = tid; i < BIG_SB; i += 256) {
      const uint64_t k = b0 + i < n ? pq[b0 + i] : kNoKey;
      buf[i] = k < cap ? k : kNoKey;
    }
    __syncthreads();
    const uint64_t *srt = block_sort_keys(buf, BIG_SB);
    // cur ascending, srt ascending: min(cur[i], srt[M - 1 - i]) is a bitonic
    // sequence holding the BIG_SB smallest of both
    for (uint32_t i = tid; i < BIG_SB; i += 256) cur[i] = umin64(cur[i], srt[BIG_SB - 1 - i]);
    __syncthreads();
    block_bitonic_merge(cur, BIG_SB);
    for (uint32_t i = a.S + tid; i < BIG_SB; i += 256) cur[i] = kNoKey;
    __syncthreads();
    cap = cur[a.S - 1];
    __syncthreads();
  }
  const uint32_t cnt = block_count_sorted(cur, a.S, stmp);
  for (uint32_t i = tid; i < cnt; i += 256) pq[i] = cur[i];
  if (tid == 0) {
    a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride] = cnt;
    a.pool_sorted[q] = cnt;
    if (cnt == a.S) {
      const uint64_t kth = cur[a.S - 1];
      if (kth < a.thr[q]) {
        a.thr[q] = kth;
        a.lbound[q] = static_cast<int32_t>(l2_bound(key_float(kth), a.step_sq, a.lmax));
      }
    }
  }
}

// k_update for S > 1,024 when no consumer needs the pool sorted: the S-th
// smallest key of the pool by a most-significant-digit radix select (8-bit
// digits of the u64 key; keys are unique), then the keys at or below it moved
// to the front, in no particular order. The pool is read from global memory
// for the histograms and moved through LDS UR_CHUNK keys at a time (a chunk
// is read whole before any of it is written, and keys only move down).
constexpr uint32_t UR_CHUNK = 8192;

__global__ __launch_bounds__(256) void k_update_radix(UpdateArgs a) {
  extern __shared__ uint64_t lds[];  // UR_CHUNK keys
  __shared__ uint32_t hist[256];
  __shared__ uint32_t stmp[4];
  __shared__ uint32_t s_digit, s_rank, s_cnt, s_out;
  __shared__ uint64_t s_pivot;
  const uint32_t q = blockIdx.x, tid = threadIdx.x;
  const uint32_t n = min(a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride], a.stride);
  const uint32_t sorted = a.pool_sorted[q] & 0x7fffffffu;
  if (n == sorted) return;
  if (n <= a.S) {  // every key stays; no threshold yet
    if (tid == 0) a.pool_sorted[q] = n;
    return;
  }
  uint64_t *pq = a.pool + static_cast<size_t>(q) * a.stride;
  uint64_t prefix = 0, pmask = 0;
  uint32_t rank = a.S;  // 1-based rank of the wanted key among those matching the prefix
  bool unique = false;
  for (int shift = 56; shift >= 0; shift -= 8) {
    hist[tid] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < n; i += 256) {
      const uint64_t k = pq[i];
      if ((k & pmask) == prefix) atomicAdd(&hist[(k >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    const uint32_t h = hist[tid];
    const uint32_t ex = block_exscan256(h, stmp);
    if (ex < rank && rank <= ex + h) {
      s_digit = tid;
      s_rank = rank - ex;
      s_cnt = h;
    }
    __syncthreads();
    prefix |= static_cast<uint64_t>(s_digit) << shift;
    pmask |= 0xffull << shift;
    rank = s_rank;
    unique = s_cnt == 1;
    __syncthreads();
    if (unique) break;
  }
  uint64_t pivot = prefix;
  if (unique) {  // the one key that matches the prefix
    for (uint32_t i = tid; i < n; i += 256) {
      const uint64_t k = pq[i];
      if ((k & pmask) == prefix) s_pivot = k;
    }
    __syncthreads();
    pivot = s_pivot;
  }
  // the keys at or below the pivot (exactly S) to the front
  if (tid == 0) s_out = 0;
  __syncthreads();
  for (uint32_t c0 = 0; c0 < n; c0 += UR_CHUNK) {
    const uint32_t len = min(UR_CHUNK, n - c0);
    for (uint32_t i = tid; i < len; i += 256) lds[i] = pq[c0 + i];
    __syncthreads();
    uint32_t cnt = 0;
    for (uint32_t i = tid; i < len; i += 256) cnt += lds[i] <= pivot ? 1u : 0u;
    uint32_t pos = s_out + block_exscan256(cnt, stmp);
    for (uint32_t i = tid; i < len; i += 256)
      if (lds[i] <= pivot) pq[pos++] = lds[i];
    __syncthreads();
    if (tid == 255) s_out = pos;
    __syncthreads();
  }
  if (tid == 0) {
    a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride] = a.S;
    a.pool_sorted[q] = a.S;
    if (pivot < a.thr[q]) {
      a.thr[q] = pivot;
      a.lbound[q] = static_cast<int32_t>(l2_bound(key_float(pivot), a.step_sq, a.lmax));
    }
  }
}

// k_refine for S > 1024: the pool after the last update holds the query's
// survivors sorted ascending; the four waves take groups of 16 for the exact
// distances; then the arrival rule in closed form (as in k_refine) over the
// block: survivor i arrives after every survivor above it (i descending).
__global__ __launch_bounds__(256) void k_refine_big(RefineArgs a) {
  extern __shared__ uint64_t lds[];
  __shared__ uint32_t stmp[4];
  __shared__ uint32_t s_mb;
  // lists of C keys: the smallest power of two holding S (the launch sizes
  // the LDS to three such lists)
  uint32_t C = 64;
  while (C < a.S) C <<= 1;
  uint64_t *ex = lds, *buf = lds + C;
  const int lane = lane_id(), w = threadIdx.x >> 6;
  const uint32_t q = blockIdx.x, tid = threadIdx.x;
  const uint64_t *pq = a.pool + static_cast<size_t>(q) * a.stride;
  const uint32_t m = min(min(a.pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride], a.stride), a.S);
  const uint64_t tq = a.thr[q];
  for (uint32_t i = tid; i < C; i += 256) {
    const uint64_t k = i < m ? pq[i] : kNoKey;
    buf[i] = k <= tq ? k : kNoKey;
  }
  __syncthreads();
  const uint32_t ns = block_count_sorted(buf, C, stmp);
  // the pool holds the survivors in no particular order (k_update_radix);
  // they arrive in descending key order, so sort them ascending first
  const uint64_t *sv = block_sort_keys(buf, C);
  if (a.surv)
    for (uint32_t i = tid; i < a.S; i += 256) a.surv[static_cast<size_t>(q) * a.S + i] = sv[i];
  if (a.nsurv && tid == 0) a.nsurv[q] = ns;
  const int8_t *qh = a.qhi + static_cast<size_t>(q) * a.dc;
  const int8_t *ql = a.qlo + static_cast<size_t>(q) * a.dc;
  const int64_t qn = a.qnorm[q];
  for (uint32_t t0 = 16 * w; t0 < ns; t0 += 64) {
    const uint32_t s = t0 + (lane & 15);
    const uint64_t key = sv[min(s, C - 1)];
    const uint64_t e = exact_key16(a, key, s < ns, qh, ql, qn);
    if (lane < 16 && t0 + lane < ns) ex[t0 + lane] = e;
  }
  __syncthreads();
  uint32_t M = 64;
  while (M < ns) M <<= 1;
  for (uint32_t i = tid; i < M; i += 256) buf[i] = i < ns ? ex[i] : kNoKey;
  __syncthreads();
  const uint64_t *srt = block_sort_keys(buf, M);
  const uint64_t *fin = srt;
  if (ns > a.K) {
    if (tid == 0) s_mb = static_cast<uint32_t>(srt[a.K - 1] >> 32);  // bits of M (dists >= 0)
    __syncthreads();
    const uint32_t mb = s_mb;
    // thread t owns survivors [R t, R t + R); pre = arrivals before it (higher
    // indices) with dist <= M
    constexpr uint32_t R = BIG_SB / 256;
    const uint32_t i0 = R * tid;
    uint32_t c = 0;
    for (uint32_t i = i0; i < i0 + R && i < ns; ++i) c += static_cast<uint32_t>(ex[i] >> 32) <= mb ? 1u : 0u;
    const uint32_t before = block_exscan256(c, stmp);  // in lower-index threads
    __shared__ uint32_t s_tot;
    if (tid == 255) s_tot = before + c;
    __syncthreads();
    uint32_t pre = s_tot - before - c;  // in higher-index threads
    __syncthreads();
    for (uint32_t i = i0 + R; i-- > i0;) {
      if (i >= ns) continue;
      const uint64_t k = ex[i];
      const uint32_t fb = static_cast<uint32_t>(k >> 32);
      const bool in = k != kNoKey && (fb < mb || (fb == mb && pre < a.K));
      if (k != kNoKey && fb <= mb) ++pre;
      ex[i] = in ? k : kNoKey;
    }
    __syncthreads();
    for (uint32_t i = tid; i < M; i += 256) buf[i] = i < ns ? ex[i] : kNoKey;
    __syncthreads();
    fin = block_sort_keys(buf, M);
  }
  for (uint32_t g = tid; g < a.K; g += 256) {
    const uint64_t k = g < M ? fin[g] : kNoKey;
    const bool ok = k != kNoKey;
    a.out_ids[static_cast<size_t>(q) * a.K + g] = ok ? static_cast<uint32_t>(k) : 0xffffffffu;
    a.out_dists[static_cast<size_t>(q) * a.K + g] = ok ? key_float(k) : FLT_MAX;
  }
}

}  // namespace

namespace katoann_hip {

uint32_t refine_cap(uint32_t S) {
  return S <= 64     ? 64
         : S <= 128  ? 128
         : S <= 256  ? 256
         : S <= 512  ? 512
         : S <= 1024 ? 1024
         : S <= BIG_SB ? BIG_SB
                       : 0;
}

template <class K>
static void big_lds_attr(K kernel) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(BIG_LDS));
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
                  const uint32_t *gcnt, const uint32_t *qoff, uint32_t qc, uint32_t rt, uint32_t qsq, uint32_t qsrt,
                  uint32_t cap, uint32_t seed, uint32_t join, const uint32_t *gbv, uint32_t *ihist, uint4 *items,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_items, dim3((cap + 255) / 256), dim3(256), 0, s, item_off, nslots, nlist, list_off, gcnt, qoff,
                     qc, rt, qsq, qsrt, cap, seed, join, gbv, ihist, items);
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

// Rows per wave of the wave-independent ranking kernel: 16 T, with T one of
// two register budgets per ranking width (B operands of 32 or 16 x 16 bytes).
// Wave-kernel tiles per prefix chunk count KC = P / 64 and variant: rows per
// wave (16 T) and query buffers in the ring (NB). Variants 0 and 1 take the
// larger and smaller T with a two-buffer ring; variant 2 a three-buffer ring
// (codes two fragments ahead) for KC <= 8, and variant 0's tile above;
// variant 3 for KC = 4 twice variant 0's rows at one wave per SIMD (the
// kernel's launch bounds), and variant 0's tile otherwise. The ring kernel
// takes T of variants 0 and 1.
struct WaveTile {
  uint8_t t, nb;
};
static const WaveTile kWaveTiles[17][4] = {
    {{0, 0}, {0, 0}, {0, 0}, {0, 0}}, {{8, 2}, {4, 2}, {8, 3}, {8, 2}}, {{8, 2}, {4, 2}, {8, 3}, {8, 2}},
    {{6, 2}, {4, 2}, {6, 3}, {6, 2}}, {{6, 2}, {4, 2}, {5, 3}, {12, 4}}, {{4, 2}, {2, 2}, {3, 3}, {4, 2}},
    {{4, 2}, {2, 2}, {3, 3}, {4, 2}}, {{3, 2}, {2, 2}, {2, 3}, {3, 2}}, {{2, 2}, {1, 2}, {2, 3}, {2, 2}},
    {{2, 2}, {1, 2}, {2, 2}, {2, 2}}, {{2, 2}, {1, 2}, {2, 2}, {2, 2}}, {{1, 2}, {1, 2}, {1, 2}, {1, 2}},
    {{1, 2}, {1, 2}, {1, 2}, {1, 2}}, {{1, 2}, {1, 2}, {1, 2}, {1, 2}}, {{1, 2}, {1, 2}, {1, 2}, {1, 2}},
    {{1, 2}, {1, 2}, {1, 2}, {1, 2}}, {{1, 2}, {1, 2}, {1, 2}, {1, 2}}};

uint32_t rank_wave_rows(uint32_t P, uint32_t variant) {
  const uint32_t kc = P / 64;
  if (kc == 0 || kc > 16 || P % 64 || variant > 3) return 0;
  return 16u * kWaveTiles[kc][variant].t;
}

uint32_t rank_wave_blocks_per_cu(uint32_t P, uint32_t variant) {
  const uint32_t t = rank_wave_rows(P, variant) / 16;
  return t * (P / 64) > 32 ? 1u : 2u;  // the kernel's launch bounds
}

template <int KC, int T, int NB>
static void wave_launch(const WaveArgs &a, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL((k_rank_wave<KC, T, NB>), dim3(grid), dim3(256), 0, s, a);
}

template <int KC, int T>
static void ring_launch(const WaveArgs &a, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL((k_rank_ring<KC, T>), dim3(grid), dim3(256), 0, s, a);
}

bool launch_rank_wave(uint32_t band, uint32_t variant, bool ring, const int8_t *plane, const int32_t *xnorm, const uint32_t *ids,
                      uint32_t nlist, uint32_t P, const int8_t *qhi, uint32_t dc, const int32_t *qnorm,
                      const int32_t *lbound, const uint32_t *qlist, const uint32_t *item_off, const uint4 *items,
                      uint32_t stride, double step_sq, uint64_t *pool, uint32_t *pool_cnt, uint32_t *state,
                      uint32_t grid, hipStream_t s) {
  if (rank_wave_rows(P, variant) == 0 || (ring && variant > 1)) return false;
  const WaveTile w = kWaveTiles[P / 64][variant];
  const WaveArgs a{plane, xnorm, ids, qhi, dc, qnorm, lbound, qlist, item_off, items, nlist, stride, pool, pool_cnt,
                   state, band, step_sq};
  if (ring) {
    switch (P / 64 * 16 + w.t) {
#define R_CASE(kc, t)              \
  case kc * 16 + t:                \
    ring_launch<kc, t>(a, grid, s); \
    return true;
      R_CASE(1, 8) R_CASE(1, 4) R_CASE(2, 8) R_CASE(2, 4) R_CASE(3, 6) R_CASE(3, 4) R_CASE(4, 6) R_CASE(4, 4)
      R_CASE(5, 4) R_CASE(5, 2) R_CASE(6, 4) R_CASE(6, 2) R_CASE(7, 3) R_CASE(7, 2) R_CASE(8, 2) R_CASE(8, 1)
      R_CASE(9, 2) R_CASE(9, 1) R_CASE(10, 2) R_CASE(10, 1) R_CASE(11, 1) R_CASE(12, 1) R_CASE(13, 1)
      R_CASE(14, 1) R_CASE(15, 1) R_CASE(16, 1)
#undef R_CASE
      default:
        return false;
    }
  }
  switch ((P / 64 * 16 + w.t) * 4 + w.nb) {
#define W_CASE(kc, t, nb)                \
  case (kc * 16 + t) * 4 + nb:           \
    wave_launch<kc, t, nb>(a, grid, s);  \
    return true;
    W_CASE(1, 8, 2) W_CASE(1, 4, 2) W_CASE(1, 8, 3) W_CASE(2, 8, 2) W_CASE(2, 4, 2) W_CASE(2, 8, 3)
    W_CASE(3, 6, 2) W_CASE(3, 4, 2) W_CASE(3, 6, 3) W_CASE(4, 6, 2) W_CASE(4, 4, 2) W_CASE(4, 5, 3) W_CASE(4, 12, 4)
    W_CASE(5, 4, 2) W_CASE(5, 2, 2) W_CASE(5, 3, 3) W_CASE(6, 4, 2) W_CASE(6, 2, 2) W_CASE(6, 3, 3)
    W_CASE(7, 3, 2) W_CASE(7, 2, 2) W_CASE(7, 2, 3) W_CASE(8, 2, 2) W_CASE(8, 1, 2) W_CASE(8, 2, 3)
    W_CASE(9, 2, 2) W_CASE(9, 1, 2) W_CASE(10, 2, 2) W_CASE(10, 1, 2) W_CASE(11, 1, 2) W_CASE(12, 1, 2)
    W_CASE(13, 1, 2) W_CASE(14, 1, 2) W_CASE(15, 1, 2) W_CASE(16, 1, 2)
#undef W_CASE
    default:
      return false;
  }
}

// Query-stationary tiles per prefix chunk count KC = P / 64 and variant:
// query fragments held (QF) and row tiles in flight (NR); the larger variant
// first. Zero where the registers do not fit.
struct QsTile {
  uint8_t qf, nr;
};
static const QsTile kQsTiles[17][2] = {
    {{0, 0}, {0, 0}}, {{4, 4}, {2, 4}}, {{4, 4}, {2, 4}}, {{4, 3}, {2, 3}}, {{4, 4}, {2, 3}}, {{2, 3}, {1, 3}},
    {{2, 3}, {1, 3}}, {{2, 2}, {1, 3}}, {{2, 2}, {1, 2}}, {{1, 3}, {1, 2}}, {{1, 3}, {1, 2}}, {{1, 2}, {1, 2}},
    {{1, 2}, {1, 2}}, {{1, 2}, {1, 2}}, {{1, 2}, {1, 2}}, {{1, 2}, {1, 2}}, {{1, 2}, {1, 2}}};

uint32_t rank_qs_queries(uint32_t P, uint32_t variant) {
  const uint32_t kc = P / 64;
  if (kc == 0 || kc > 16 || P % 64 || variant > 1) return 0;
  return 16u * kQsTiles[kc][variant].qf;
}

uint32_t rank_qs_blocks_per_cu(uint32_t, uint32_t) { return 2u; }  // the kernel's launch bounds

template <int KC, int QF, int NR>
static void qs_launch(const WaveArgs &a, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL((k_rank_qs<KC, QF, NR>), dim3(grid), dim3(256), 0, s, a);
}

bool launch_rank_qs(uint32_t band, uint32_t variant, const int8_t *plane, const int32_t *xnorm, const uint32_t *ids,
                    uint32_t nlist, uint32_t P, const int8_t *qhi, uint32_t dc, const int32_t *qnorm,
                    const int32_t *lbound, const uint32_t *qlist, const uint32_t *item_off, const uint4 *items,
                    uint32_t stride, double step_sq, uint64_t *pool, uint32_t *pool_cnt, uint32_t *state,
                    uint32_t grid, hipStream_t s) {
  if (rank_qs_queries(P, variant) == 0) return false;
  const QsTile q = kQsTiles[P / 64][variant];
  const WaveArgs a{plane, xnorm, ids, qhi, dc, qnorm, lbound, qlist, item_off, items, nlist, stride, pool, pool_cnt,
                   state, band, step_sq};
  switch ((P / 64 * 8 + q.qf) * 8 + q.nr) {
#define Q_CASE(kc, qf, nr)           \
  case (kc * 8 + qf) * 8 + nr:       \
    qs_launch<kc, qf, nr>(a, grid, s); \
    return true;
    Q_CASE(1, 4, 4) Q_CASE(1, 2, 4) Q_CASE(2, 4, 4) Q_CASE(2, 2, 4) Q_CASE(3, 4, 3) Q_CASE(3, 2, 3)
    Q_CASE(4, 4, 4) Q_CASE(4, 2, 3) Q_CASE(5, 2, 3) Q_CASE(5, 1, 3) Q_CASE(6, 2, 3) Q_CASE(6, 1, 3)
    Q_CASE(7, 2, 2) Q_CASE(7, 1, 3) Q_CASE(8, 2, 2) Q_CASE(8, 1, 2) Q_CASE(9, 1, 3) Q_CASE(9, 1, 2)
    Q_CASE(10, 1, 3) Q_CASE(10, 1, 2) Q_CASE(11, 1, 2) Q_CASE(12, 1, 2) Q_CASE(13, 1, 2) Q_CASE(14, 1, 2)
    Q_CASE(15, 1, 2) Q_CASE(16, 1, 2)
#undef Q_CASE
    default:
      return false;
  }
}

template <int KC, int T, int NB, int QF, int NR>
static void hyb_launch(const WaveArgs &a, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL((k_rank_hyb<KC, T, NB, QF, NR>), dim3(grid), dim3(256), 0, s, a);
}

bool launch_rank_hyb(uint32_t band, uint32_t variant, const int8_t *plane, const int32_t *xnorm, const uint32_t *ids,
                     uint32_t nlist, uint32_t P, const int8_t *qhi, uint32_t dc, const int32_t *qnorm,
                     const int32_t *lbound, const uint32_t *qlist, const uint32_t *item_off, const uint4 *items,
                     uint32_t stride, double step_sq, uint64_t *pool, uint32_t *pool_cnt, uint32_t *state,
                     uint32_t grid, hipStream_t s) {
  if (variant != 0 || rank_wave_rows(P, 0) == 0 || rank_qs_queries(P, 0) == 0) return false;
  const WaveArgs a{plane, xnorm, ids, qhi, dc, qnorm, lbound, qlist, item_off, items, nlist, stride, pool, pool_cnt,
                   state, band, step_sq};
  switch (P / 64) {
#define H_CASE(kc, t, nb, qf, nr)            \
  case kc:                                   \
    hyb_launch<kc, t, nb, qf, nr>(a, grid, s); \
    return true;
    H_CASE(1, 8, 2, 4, 4) H_CASE(2, 8, 2, 4, 4) H_CASE(3, 6, 2, 4, 3) H_CASE(4, 6, 2, 4, 4) H_CASE(5, 4, 2, 2, 3)
    H_CASE(6, 4, 2, 2, 3) H_CASE(7, 3, 2, 2, 2) H_CASE(8, 2, 2, 2, 2) H_CASE(9, 2, 2, 1, 3) H_CASE(10, 2, 2, 1, 3)
    H_CASE(11, 1, 2, 1, 2) H_CASE(12, 1, 2, 1, 2) H_CASE(13, 1, 2, 1, 2) H_CASE(14, 1, 2, 1, 2)
    H_CASE(15, 1, 2, 1, 2) H_CASE(16, 1, 2, 1, 2)
#undef H_CASE
    default:
      return false;
  }
}

uint32_t rank_full_rows() { return 16 * FULL_T; }
uint32_t rank_bf16_rows() { return BF16_ROWS; }
uint32_t rank_bf16_blocks_per_cu() { return 5; }  // 82 VGPRs: five waves per SIMD

void launch_rank_bf16(uint32_t band, const uint16_t *rows, const float *xnorm, const uint32_t *ids, uint32_t nlist,
                      const uint16_t *qbf, const float *qnorm, uint32_t dp, const uint64_t *thr,
                      const uint32_t *qlist, const uint32_t *item_off, const uint4 *items, uint32_t stride,
                      uint64_t *pool, uint32_t *pool_cnt, uint32_t *state, uint32_t grid, hipStream_t s) {
  const Bf16Args a{reinterpret_cast<const uint32_t *>(rows), xnorm, ids, reinterpret_cast<const uint32_t *>(qbf),
                   qnorm, thr, qlist, item_off, items, nlist, band, dp / 2, stride, pool, pool_cnt, state};
  hipLaunchKernelGGL(k_rank_bf16, dim3(grid), dim3(256), 0, s, a);
}

void launch_emit_bf16(const uint64_t *pool, uint32_t stride, const uint32_t *pool_cnt, uint32_t B, uint32_t K,
                      uint32_t *out_ids, float *out_dists, uint64_t *surv, uint32_t *nsurv, hipStream_t s) {
  if (!B) return;
  hipLaunchKernelGGL(k_emit_bf16, dim3(B), dim3(64), 0, s, pool, stride, pool_cnt, K, out_ids, out_dists, surv, nsurv);
}

uint32_t rank_full_blocks_per_cu() { return 3; }  // 142 VGPRs: three waves per SIMD

void launch_rank_full(uint32_t band, const int8_t *hi, const int8_t *lo, const int64_t *xnorm, const uint32_t *ids,
                      uint32_t nlist, const int8_t *qhi, const int8_t *qlo, uint32_t dc, const int64_t *qnorm,
                      const uint64_t *thr, const uint32_t *qlist, const uint32_t *item_off, const uint4 *items,
                      uint32_t stride, double step_sq, uint64_t *pool, uint32_t *pool_cnt, uint32_t *state,
                      uint32_t grid, hipStream_t s) {
  const FullArgs a{hi, lo, xnorm, ids, qhi, qlo, qnorm, thr, qlist, item_off, items, nlist, band, dc, stride,
                   pool, pool_cnt, state, step_sq};
  hipLaunchKernelGGL(k_rank_full, dim3(grid), dim3(256), 0, s, a);
}

bool launch_update(uint64_t *pool, uint32_t *pool_cnt, uint32_t *pool_sorted, uint64_t *thr, int32_t *lbound,
                   uint32_t B, uint32_t S, uint32_t stride, uint32_t P, double step_sq, bool keep_sorted,
                   hipStream_t s) {
  UpdateArgs a{pool, pool_cnt, pool_sorted, thr, lbound, B, S, stride, static_cast<int64_t>(P) * 65025, step_sq};
  const dim3 grid(B);
  if (refine_cap(S) == BIG_SB && !keep_sorted) {
    static const bool attr = (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_update_radix),
                                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  static_cast<int>(UR_CHUNK * sizeof(uint64_t))),
                              true);
    (void)attr;
    hipLaunchKernelGGL(k_update_radix, grid, dim3(256), UR_CHUNK * sizeof(uint64_t), s, a);
    return true;
  }
  // large pools by selection, then the rest (and any the selection left) by merging
  if (S <= 128) hipLaunchKernelGGL(k_update_sel<false>, grid, dim3(256), 0, s, a);
  else if (S <= 256) hipLaunchKernelGGL(k_update_sel<true>, grid, dim3(256), 0, s, a);
  switch (refine_cap(S)) {
    case 64: hipLaunchKernelGGL(k_update<64>, grid, dim3(256), 0, s, a); return true;
    case 128: hipLaunchKernelGGL(k_update<128>, grid, dim3(256), 0, s, a); return true;
    case 256: hipLaunchKernelGGL(k_update<256>, grid, dim3(256), 0, s, a); return true;
    case 512: hipLaunchKernelGGL(k_update<512>, grid, dim3(256), 0, s, a); return true;
    case 1024: hipLaunchKernelGGL(k_update<1024>, grid, dim3(256), 0, s, a); return true;
    case BIG_SB: {
      static const bool attr = (big_lds_attr(&k_update_big), true);
      (void)attr;
      hipLaunchKernelGGL(k_update_big, grid, dim3(256), BIG_LDS, s, a);
      return true;
    }
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
  if (dc == 0 || dc % 64 || dc > RF_MAXDC) return false;
  // chunks of 64 bytes per gather step: the largest of 8, 2, 1 dividing dc / 64
  const uint32_t nk = dc / 64, rb = nk % 8 == 0 ? 8u : nk % 2 == 0 ? 2u : 1u;
#define RF_CASE(sc)                                                                       \
  case sc:                                                                                \
    if (rb == 8) hipLaunchKernelGGL((k_refine<sc, 8>), grid, dim3(256), 0, s, a);         \
    else if (rb == 2) hipLaunchKernelGGL((k_refine<sc, 2>), grid, dim3(256), 0, s, a);    \
    else hipLaunchKernelGGL((k_refine<sc, 1>), grid, dim3(256), 0, s, a);                 \
    return true;
  switch (refine_cap(S)) {
    RF_CASE(64)
    RF_CASE(128)
    RF_CASE(256)
    RF_CASE(512)
    RF_CASE(1024)
#undef RF_CASE
    case BIG_SB: {
      static const bool attr = (big_lds_attr(&k_refine_big), true);
      (void)attr;
      uint32_t C = 64;
      while (C < S) C <<= 1;
      hipLaunchKernelGGL(k_refine_big, dim3(B), dim3(256), 3 * C * sizeof(uint64_t), s, a);
      return true;
    }
    default: return false;
  }
}

}  // namespace katoann_hip
