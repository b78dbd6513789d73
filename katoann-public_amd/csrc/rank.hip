// Ranking pass, its threshold seed, and the survivors / refinement / top-k
// kernel of the gfx950 search (reference src/index.cpp:998-1127 with
// kernels.cpp:296-316).
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
// 1. k_seed (query-major, one wave per query): the S-th smallest key over the
//    first candidates of the query's nearest lists sets the threshold.
// 2. Probe ranks are split into bands [2^b - 1, 2^(b+1) - 1). For each band:
//    k_rank_band (list-major, persistent) takes items (list, up to 64 queries,
//    row tile) of the band from a device counter; rows are read from HBM once
//    into LDS for all queries of the item (prefetched through registers while
//    the previous 64 rows are in use), each wave owns 16 queries (int8 MFMA),
//    and a candidate whose integer l2 is within its query's threshold bound is
//    appended to the query's pool. k_update then keeps each pool's S smallest
//    keys and lowers the threshold to the S-th: nearer lists come first, so the
//    bound is tight when the far lists are ranked and few candidates pass.
// 3. k_refine (one wave per query): the S smallest keys of the pool, exact
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
struct SeedArgs {
  const int8_t *plane;
  const int32_t *xnorm;
  const uint32_t *ids;
  const uint64_t *list_off;
  uint32_t P;
  const int8_t *qhi;
  uint32_t dc;
  const int32_t *qnorm;
  const uint32_t *probes;
  const uint32_t *npq;
  uint32_t np, B, S;
  uint32_t target;    // candidates to gather before stopping
  uint32_t maxlists;  // probes to look at, at most
  double step_sq;
  uint64_t *thr;
};

constexpr int SEED_CH = 16;  // up to 1024 candidates per query

__global__ __launch_bounds__(256) void k_seed(SeedArgs a) {
  __shared__ uint64_t keys[4][SEED_CH * 64];
  const int lane = lane_id(), w = threadIdx.x >> 6;
  const uint32_t q = blockIdx.x * 4 + w;
  if (q >= a.B) return;
  const uint32_t KC = a.P / 64;
  const uint32_t cap = SEED_CH * 64;
  const uint32_t npq = a.npq ? min(a.npq[q], a.np) : a.np;
  const uint32_t nl = min(npq, a.maxlists);
  const int32_t qn = a.qnorm[q];
  const int8_t *qp = a.qhi + static_cast<size_t>(q) * a.dc + 16 * (lane >> 4);
  uint32_t m = 0;
  for (uint32_t p = 0; p < nl && m < a.target && m < cap; ++p) {
    const uint32_t c = a.probes[static_cast<size_t>(q) * a.np + p];
    const uint64_t r0 = a.list_off[c];
    const uint32_t n = static_cast<uint32_t>(a.list_off[c + 1] - r0);
    for (uint32_t t0 = 0; t0 < n && m < cap; t0 += 16) {
      const uint32_t ri = min(t0 + (lane & 15), n - 1);
      const int8_t *rp = a.plane + (r0 + ri) * a.P + 16 * (lane >> 4);
      v4i acc = {0, 0, 0, 0};
      for (uint32_t kc = 0; kc < KC; ++kc) acc = mfma(load16(rp + 64 * kc), load16(qp + 64 * kc), acc);
      // every column holds the same dot; column 0 sits in lanes 0, 16, 32, 48
      if ((lane & 15) == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t i = t0 + 4 * (lane >> 4) + r;
          const uint32_t slot = m + (i - t0);
          if (i < n && slot < cap) {
            const uint64_t row = r0 + i;
            const int32_t l2 = qn + a.xnorm[row] - 2 * acc[r];
            keys[w][slot] = make_key(fscale(l2, a.step_sq), a.ids[row]);
          }
        }
      }
      m = min(cap, m + min(16u, n - t0));
    }
  }
  wave_fence();
  if (m < a.S) return;
  uint64_t e[SEED_CH];
#pragma unroll
  for (int i = 0; i < SEED_CH; ++i) {
    const uint32_t idx = 64 * i + lane;
    e[i] = idx < m ? keys[w][idx] : kNoKey;
  }
  const uint64_t kth = kth_key<SEED_CH>(e, a.S);
  if (lane == 0)
    atomicMin(reinterpret_cast<unsigned long long *>(&a.thr[q]), static_cast<unsigned long long>(kth));
}

// --------------------------------------------------------------------------
// Ranking pass of one probe-rank band. Items of the band (list, up to 64
// queries, up to rt rows) are taken from the band's queue; each query's
// threshold is fixed for the band (set by the seed and the previous bands'
// updates), so the filter is one compare per candidate and kept keys are rare
// enough to be appended to the query's pool directly.
struct BandArgs {
  const int8_t *plane;       // N x P, row-major, shifted high bytes
  const int32_t *xnorm;      // N: shifted norm of those bytes
  const uint32_t *ids;       // N
  uint32_t nlist;
  uint32_t band;
  const int8_t *qhi;         // B x dc
  uint32_t dc;
  const int32_t *qnorm;      // B
  const uint32_t *qlist;
  const uint32_t *item_off;  // per (band, list) slot, nslots + 1
  const uint4 *items;        // item descriptors: first query slot, queries, first row, end row
  uint32_t stride;           // pool slots per query
  int32_t lmax;              // P * 255^2
  double step_sq;
  uint64_t *pool;            // B x stride
  uint32_t *pool_cnt;        // B
  const uint64_t *thr;       // B
  uint32_t *state;           // [0] overflow, [1] kept, [3 + band] queue head
};

constexpr int RK_ROWS = 64;
constexpr int RK_QC = 64;

__device__ __forceinline__ int32_t bound_of(uint64_t key, double step_sq, int32_t lmax) {
  return key == kNoKey ? 0x7fffffff : static_cast<int32_t>(l2_bound(key_float(key), step_sq, lmax));
}

// Descriptor of every ranking item: thread per item, binary search of its
// (band, list) slot in item_off.
__global__ void k_items(const uint32_t *__restrict__ item_off, uint32_t nslots, uint32_t nlist,
                        const uint64_t *__restrict__ list_off, const uint32_t *__restrict__ gcnt,
                        const uint32_t *__restrict__ qoff, uint32_t rt, uint32_t cap, uint4 *items) {
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
  d.x = qoff[slot] + qch * RK_QC;
  d.y = min(static_cast<uint32_t>(RK_QC), gcnt[slot] - qch * RK_QC);
  d.z = static_cast<uint32_t>(list_off[c] + static_cast<uint64_t>(rtile) * rt);
  d.w = static_cast<uint32_t>(list_off[c] + min(n, static_cast<uint64_t>(rtile + 1) * rt));
  items[it] = d;
}

template <int KC>
__global__ __launch_bounds__(256) void k_rank_band(BandArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  constexpr uint32_t P = KC * 64, PIECES = P / 16;
  int8_t *rows = reinterpret_cast<int8_t *>(smem);                           // 64 x P, MFMA fragment order
  int32_t *xns = reinterpret_cast<int32_t *>(smem + RK_ROWS * P);            // 64 row norms
  uint64_t *stage = reinterpret_cast<uint64_t *>(xns + RK_ROWS);            // 4 waves x 16 queries x 64 kept keys
  __shared__ uint32_t s_next;
  __shared__ uint4 s_desc;
  __shared__ uint32_t sqid[RK_QC];
  __shared__ int32_t sqn[RK_QC];
  __shared__ uint32_t skept[4][16];
  const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
  const uint32_t ib = a.item_off[a.band * a.nlist], ie = a.item_off[(a.band + 1) * a.nlist];
  uint64_t *stw = stage + w * 16 * 64;
  uint32_t n_kept = 0;

  if (tid == 0) s_next = ib + atomicAdd(&a.state[3 + a.band], 1u);
  __syncthreads();
  uint32_t it = s_next;
  if (it >= ie) return;
  uint4 desc = a.items[it];
  for (;;) {
    const uint32_t qbeg = desc.x, nq = desc.y, rbeg = desc.z, rend = desc.w;
    // next item's index and descriptor in flight during this item
    uint32_t nxt = 0;
    if (tid == 0) nxt = ib + atomicAdd(&a.state[3 + a.band], 1u);
    v4i pf[KC];
    int32_t pxn = 0;
    auto prefetch = [&](uint32_t r0) {
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const uint32_t e = tid + 256 * k;
        const uint32_t rl = e / PIECES, pc = e % PIECES;
        pf[k] = load16(a.plane + static_cast<size_t>(min(r0 + rl, rend - 1)) * P + 16 * pc);
      }
      if (tid < RK_ROWS) pxn = a.xnorm[min(r0 + tid, rend - 1)];
    };
    prefetch(rbeg);
    if (tid < RK_QC) {
      const bool ok = tid < static_cast<int>(nq);
      const uint32_t qid = a.qlist[qbeg + (ok ? tid : 0)];
      sqid[tid] = qid;
      sqn[tid] = ok ? a.qnorm[qid] : 0;
    }
    __syncthreads();
    const int wq = max(0, min(16, static_cast<int>(nq) - 16 * w));
    v4i areg[KC];
    int32_t qn4[4], L4[4];
    if (wq > 0) {
      const bool ok = (lane & 15) < wq;
      const int8_t *qp = a.qhi + static_cast<size_t>(sqid[16 * w + (ok ? (lane & 15) : 0)]) * a.dc +
                         16 * (lane >> 4);
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) areg[kc] = ok ? load16(qp + 64 * kc) : v4i{0, 0, 0, 0};
      int32_t Lv = -1;
      if (lane < wq) Lv = bound_of(a.thr[sqid[16 * w + lane]], a.step_sq, a.lmax);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int sq = 4 * (lane >> 4) + r;
        qn4[r] = sqn[16 * w + sq];
        L4[r] = __shfl(Lv, sq);
      }
    }
    for (uint32_t r0 = rbeg; r0 < rend; r0 += RK_ROWS) {
      __syncthreads();
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const uint32_t e = tid + 256 * k;
        const uint32_t rl = e / PIECES, pc = e % PIECES;
        const uint32_t off = (((rl >> 4) * KC + (pc >> 2)) * 64 + (rl & 15) + 16 * (pc & 3)) * 16;
        *reinterpret_cast<v4i *>(rows + off) = pf[k];
      }
      if (tid < RK_ROWS) xns[tid] = pxn;
      if (tid < 64) skept[tid >> 4][tid & 15] = 0;
      __syncthreads();
      if (r0 + RK_ROWS < rend) prefetch(r0 + RK_ROWS);
      if (wq == 0) continue;
      v4i acc[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = v4i{0, 0, 0, 0};
#pragma unroll
      for (int kc = 0; kc < KC; ++kc)
#pragma unroll
        for (int t = 0; t < 4; ++t)
          acc[t] = mfma(areg[kc], *reinterpret_cast<const v4i *>(rows + ((t * KC + kc) * 64 + lane) * 16), acc[t]);
      bool any = false;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const uint32_t rl = 16 * t + (lane & 15);
        const bool rok = r0 + rl < rend;
        const int32_t xn = xns[rl];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int sq = 4 * (lane >> 4) + r;
          const int32_t l2 = qn4[r] + xn - 2 * acc[t][r];
          const bool keep = rok && sq < wq && l2 <= L4[r];
          if (keep) {
            const uint32_t pos = atomicAdd(&skept[w][sq], 1u);
            stw[sq * 64 + pos] = make_key(fscale(l2, a.step_sq), a.ids[r0 + rl]);
          }
          any |= keep;
        }
      }
      if (ballot(any)) {
        wave_fence();
        // one global reservation per query for its kept keys of these 64 rows
        const uint32_t nk = lane < 16 ? skept[w][lane] : 0u;
        uint32_t base = 0;
        if (nk) base = atomicAdd(&a.pool_cnt[sqid[16 * w + lane]], nk);
        for (int sq = 0; sq < wq; ++sq) {
          const uint32_t n = __shfl(nk, sq);
          if (!n) continue;
          n_kept += n;
          const uint32_t b0 = __shfl(base, sq);
          const uint32_t qid = sqid[16 * w + sq];
          if (lane < static_cast<int>(n)) {
            const uint32_t pos = b0 + lane;
            if (pos < a.stride) a.pool[static_cast<size_t>(qid) * a.stride + pos] = stw[sq * 64 + lane];
            else atomicOr(&a.state[0], 1u);
          }
        }
      }
    }
    if (tid == 0) s_next = nxt;
    __syncthreads();
    it = s_next;
    if (it >= ie) break;
    desc = a.items[it];
  }
  if (lane == 0 && n_kept) atomicAdd(&a.state[1], n_kept);
}

// After a band: each query's pool back to its S smallest keys, ascending, and
// the query's threshold to the S-th of them (an upper bound of its final S-th
// key). Entries [0, sorted) are the previous update's sorted list.
struct UpdateArgs {
  uint64_t *pool;
  uint32_t *pool_cnt, *pool_sorted;
  uint64_t *thr;
  uint32_t B, S, stride;
};

template <int SC>
__global__ __launch_bounds__(256) void k_update(UpdateArgs a) {
  const int lane = lane_id();
  const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= a.B) return;
  const uint32_t n = min(a.pool_cnt[q], a.stride), sorted = a.pool_sorted[q];
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
    a.pool_cnt[q] = cnt;
    a.pool_sorted[q] = cnt;
  }
  if (cnt == a.S) {
    const uint64_t kth = top_get<SC>(t, a.S - 1);
    if (lane == 0 && kth < a.thr[q]) a.thr[q] = kth;
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
    const uint32_t m = min(a.pool_cnt[q], a.stride);
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

void launch_seed(const int8_t *plane, const int32_t *xnorm, const uint32_t *ids, const uint64_t *list_off,
                 uint32_t P, const int8_t *qhi, uint32_t dc, const int32_t *qnorm, const uint32_t *probes,
                 const uint32_t *npq, uint32_t np, uint32_t B, uint32_t S, double step_sq, uint64_t *thr,
                 hipStream_t s) {
  if (S > SEED_CH * 64) return;
  SeedArgs a{plane, xnorm, ids, list_off, P, qhi, dc, qnorm, probes, npq, np, B, S,
             std::min(static_cast<uint32_t>(SEED_CH * 64), std::max(256u, 4 * S)), 8u, step_sq, thr};
  hipLaunchKernelGGL(k_seed, dim3((B + 3) / 4), dim3(256), 0, s, a);
}

size_t band_smem(uint32_t P) {
  return static_cast<size_t>(RK_ROWS) * P + RK_ROWS * 4 + 4 * 16 * 64 * 8;
}

void launch_items(const uint32_t *item_off, uint32_t nslots, uint32_t nlist, const uint64_t *list_off,
                  const uint32_t *gcnt, const uint32_t *qoff, uint32_t rt, uint32_t cap, uint4 *items,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_items, dim3((cap + 255) / 256), dim3(256), 0, s, item_off, nslots, nlist, list_off, gcnt, qoff,
                     rt, cap, items);
}

template <int KC>
static void band_launch(const BandArgs &a, uint32_t grid, hipStream_t s) {
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_rank_band<KC>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(band_smem(KC * 64)));
    attr = true;
  }
  hipLaunchKernelGGL(k_rank_band<KC>, dim3(grid), dim3(256), band_smem(KC * 64), s, a);
}

bool launch_rank_band(uint32_t band, const int8_t *plane, const int32_t *xnorm, const uint32_t *ids, uint32_t nlist,
                      uint32_t P, const int8_t *qhi, uint32_t dc, const int32_t *qnorm, const uint32_t *qlist,
                      const uint32_t *item_off, const uint4 *items, uint32_t stride, double step_sq, uint64_t *pool,
                      uint32_t *pool_cnt, const uint64_t *thr, uint32_t *state, uint32_t grid, hipStream_t s) {
  BandArgs a{plane, xnorm, ids, nlist, band, qhi, dc, qnorm, qlist, item_off, items, stride,
             static_cast<int32_t>(P * 65025u), step_sq, pool, pool_cnt, thr, state};
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

bool launch_update(uint64_t *pool, uint32_t *pool_cnt, uint32_t *pool_sorted, uint64_t *thr, uint32_t B, uint32_t S,
                   uint32_t stride, hipStream_t s) {
  UpdateArgs a{pool, pool_cnt, pool_sorted, thr, B, S, stride};
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
