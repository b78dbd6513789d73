// Ranking pass of one probe-rank band, block-tiled form (rank kernel 5): the
// same keys, filter and pools as the other ranking kernels (rank.hip), laid
// out as a tiled int8 product of an item's queries by its rows.
//
// An item is (list, up to 64 QTW of the band's queries of the list, a row
// range). A block of 4 waves takes items dealt round-robin over the blocks;
// wave w holds queries [16 QTW w, 16 QTW (w + 1)) of the item as MFMA A
// operands (registers), and the item's rows stream through an LDS ring of NST
// stages of 64 rows, copied straight from HBM into LDS (global_load_lds)
// NST - 1 stages ahead, across item boundaries, so the copies of the next
// stages are in flight while a stage is ranked. Every wave ranks its queries
// against the 4 row tiles of 16 rows of each stage: the rows come from HBM once
// per item and feed 4 QTW query tiles, so a list with many queries runs on the
// matrix cores at tile sizes a GEMM uses, and a list with few streams at the
// copy rate.
//
// Waits: each wave waits for its own copies of a stage with an explicit
// vmcnt (the copies of the stages issued after it stay in flight; vector
// loads complete in order, so any other memory operation issued later only
// lengthens the wait), then a bare s_barrier makes all four waves' copies of
// the stage visible. Every LDS access of the ranking loop is written as asm
// that waits for its own results, so the compiler inserts no wait for the
// copies in flight.
//
// Kept keys are staged per wave in LDS with their query slot (lane s of the
// wave holds the query of slot s) and flushed to the pools at the end of an
// item, or when a fragment would not fit: one returning atomic per slot
// reserves its run in the query's pool.
#include "device.hpp"
#include "kernels.hpp"

using namespace katoann_dev;

namespace {

struct GemmArgs {
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

template <int KC>
struct GemmShape {
  static constexpr int QTW = KC <= 4 ? 4 : 2;              // 16-query tiles per wave
  static constexpr int NST = KC <= 4 ? 3 : 2;              // LDS stages
  static constexpr uint32_t ROWS = 64;                     // rows per stage
  static constexpr uint32_t STAGE = ROWS * 64u * KC;       // code bytes per stage
  static constexpr uint32_t DMA_OPS = KC + 2;              // copies per wave per stage
};
constexpr uint32_t G_CAP = 512;  // staged keys per wave

// s_waitcnt immediate (gfx9 encoding) that waits for the vector memory
// counter only: at most n of the wave's vector memory operations in flight.
constexpr int vm_imm(int n) { return (n & 0xf) | (((n >> 4) & 3) << 14) | (0x7 << 4) | (0xf << 8); }

typedef __attribute__((address_space(3))) void lds_void_t;

__device__ __forceinline__ v4i g_load16(const int8_t *p) { return *reinterpret_cast<const v4i *>(p); }

__device__ __forceinline__ v4i g_mfma(v4i a, v4i b, v4i c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// Item descriptors through the scalar cache: they were written by an earlier
// launch and are read-only here, and a scalar load is not waited for with the
// vector memory counter (that wait would also wait for the copies in flight).
__device__ __forceinline__ uint4 g_item(const uint4 *items, uint32_t i) {
  typedef const __attribute__((address_space(4))) uint32_t *cptr;
  const cptr q = (cptr)(reinterpret_cast<uintptr_t>(items + i));
  return uint4{q[0], q[1], q[2], q[3]};
}

__device__ __forceinline__ uint32_t g_lds_addr(const void *p) {
  return static_cast<uint32_t>(reinterpret_cast<size_t>((const __attribute__((address_space(3))) char *)p));
}

// Copies of 16 / 4 bytes per lane from global memory into LDS: lane l's bytes
// land at l * size past the wave-uniform LDS address.
__device__ __forceinline__ void g_dma16(const void *g, void *l) {
  __builtin_amdgcn_global_load_lds(g, (lds_void_t *)l, 16, 0, 0);
}
__device__ __forceinline__ void g_dma4(const void *g, void *l) {
  __builtin_amdgcn_global_load_lds(g, (lds_void_t *)l, 4, 0, 0);
}

// LDS accesses, each asm block waiting for its own results.
__device__ __forceinline__ void g_rd4(uint32_t a, v4i &x0, v4i &x1, v4i &x2, v4i &x3) {
  asm volatile(
      "ds_read_b128 %0, %4\n\t"
      "ds_read_b128 %1, %4 offset:1024\n\t"
      "ds_read_b128 %2, %4 offset:2048\n\t"
      "ds_read_b128 %3, %4 offset:3072\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(x0), "=&v"(x1), "=&v"(x2), "=&v"(x3)
      : "v"(a));
}
__device__ __forceinline__ void g_rd1(uint32_t a, v4i &x0) {
  asm volatile(
      "ds_read_b128 %0, %1\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(x0)
      : "v"(a));
}
__device__ __forceinline__ void g_rd_pair(uint32_t ax, uint32_t ai, int32_t &x, uint32_t &i) {
  asm volatile(
      "ds_read_b32 %0, %2\n\t"
      "ds_read_b32 %1, %3\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(x), "=&v"(i)
      : "v"(ax), "v"(ai));
}

// Waits for the copies of the stage to be ranked: `younger` stages were
// issued after it (their copies may stay in flight). Written with the builtin
// so that the compiler's own wait insertion knows the counter's state.
template <int KC>
__device__ __forceinline__ void wait_stage(uint32_t younger) {
  constexpr int D = static_cast<int>(GemmShape<KC>::DMA_OPS);
  if (younger == 0) __builtin_amdgcn_s_waitcnt(vm_imm(0));
  else if (younger == 1) __builtin_amdgcn_s_waitcnt(vm_imm(D));
  else __builtin_amdgcn_s_waitcnt(vm_imm(2 * D));
}

// Inclusive prefix sums within each 16-lane row of the wave (DPP row shifts).
__device__ __forceinline__ uint32_t g_row_scan16(uint32_t x) {
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x111, 0xf, 0xf, true));
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x112, 0xf, 0xf, true));
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x114, 0xf, 0xf, true));
  x += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x118, 0xf, 0xf, true));
  return x;
}

__device__ __forceinline__ uint32_t g_byte_sum(uint32_t x) {
  return (x & 0xffu) + ((x >> 8) & 0xffu) + ((x >> 16) & 0xffu) + (x >> 24);
}

// v[t] for a lane-varying t < T (a select chain, not an indexed stack array).
template <int T, class V>
__device__ __forceinline__ V g_pick(const V (&v)[T], int t) {
  V r = v[0];
#pragma unroll
  for (int i = 1; i < T; ++i) {
    r = t == i ? v[i] : r;
    asm volatile("" : "+v"(r));
  }
  return r;
}

// The wave's key stage in LDS: keys, metas (slot | index << 6: the key's
// query slot and its index among the slot's staged keys) and per-slot counts.
// Its accesses are plain: the copies in flight land in other LDS arrays, and
// the compiler's wait insertion sees that they cannot alias.
struct GStage {
  uint64_t *key;   // G_CAP
  uint32_t *meta;  // G_CAP
  uint32_t *cnt;   // 64
  uint32_t n;      // keys staged (wave-uniform)
};

// Staged keys -> pools; lane s holds the query of slot s (slotq) and whether
// the slot is in the item (slotok).
__device__ __forceinline__ void g_flush(GStage &st, const GemmArgs &a, uint32_t slotq, bool slotok) {
  const int lane = lane_id();
  const uint32_t c = st.cnt[lane];
  uint32_t base = 0;
  if (slotok && c) base = atomicAdd(&a.pool_cnt[static_cast<size_t>(slotq) * katoann_hip::kCntStride], c);
  bool ovf = false;
  for (uint32_t e0 = 0; e0 < st.n; e0 += 64) {
    const uint32_t e = min(e0 + lane, st.n - 1);
    const uint64_t key = st.key[e];
    const uint32_t m = st.meta[e];
    const int slot = static_cast<int>(m & 63u);
    const uint32_t pos = __shfl(base, slot) + (m >> 6);
    const uint32_t q = __shfl(slotq, slot);
    if (e0 + lane < st.n) {
      if (pos < a.stride) a.pool[static_cast<size_t>(q) * a.stride + pos] = key;
      else ovf = true;
    }
  }
  if (ovf) atomicOr(&a.state[0], 1u);
  st.cnt[lane] = 0;
  st.n = 0;
  // the flush's memory operations complete here (it waited for its atomics,
  // and with them for every copy issued before): nothing of it is left for a
  // later wait of the compiler to cover
  __builtin_amdgcn_s_waitcnt(vm_imm(0));
}

// The kept keys of query tile f (slots 16 f .. 16 f + 15 of the wave) against
// T row tiles: acc[t][r] = dot(slot 16 f + 4 g16 + r, row 16 t + c16) -
// ceil(xn / 2); a key is kept iff acc >= H of its query (H4[r]) and its row is
// in the item (vmask bit t); its l2 is qn - 2 acc - odd (oddm bit t). Keys are
// staged (a fragment that does not fit flushes the stage first), or, when more
// than the stage holds, written straight to the pools. Returns the keys kept.
template <int T>
__device__ __forceinline__ uint32_t g_keep(GStage &st, const GemmArgs &a, const v4i (&acc)[T], int f,
                                           uint32_t vmask, uint32_t oddm, const uint32_t (&xid)[T],
                                           const int32_t (&H4)[4], int32_t qn16, uint32_t slotq, bool slotok) {
  const int lane = lane_id(), g16 = lane >> 4, c16 = lane & 15;
  int32_t qnr[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) qnr[r] = __shfl(qn16, 4 * g16 + r);
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
  // per 16-lane group g, byte r: keys of slot 4 g + r up to this lane (at most 16 T < 256)
  const uint32_t inc = g_row_scan16(c), ex = inc - c;
  const uint32_t G0 = __builtin_amdgcn_readlane(inc, 15), G1 = __builtin_amdgcn_readlane(inc, 31);
  const uint32_t G2 = __builtin_amdgcn_readlane(inc, 47), G3 = __builtin_amdgcn_readlane(inc, 63);
  const uint32_t S0 = g_byte_sum(G0), S1 = g_byte_sum(G1), S2 = g_byte_sum(G2), S3 = g_byte_sum(G3);
  const uint32_t total = S0 + S1 + S2 + S3;
  if (total == 0) return 0;
  const uint32_t Gm = g16 == 0 ? G0 : g16 == 1 ? G1 : g16 == 2 ? G2 : G3;
  const bool direct = total > G_CAP;
  if ((direct || st.n + total > G_CAP) && st.n) g_flush(st, a, slotq, slotok);
  uint32_t so = st.n + (g16 > 0 ? S0 : 0u) + (g16 > 1 ? S1 : 0u) + (g16 > 2 ? S2 : 0u);
  if (direct) {
    // one returning atomic per slot of the fragment (lanes c16 < 4 of each
    // group), then every lane writes its keys to the pools
    uint32_t dbase = 0;
    if (c16 < 4) {
      const uint32_t qj = __shfl(slotq, 16 * f + 4 * g16 + c16);
      const uint32_t n = (Gm >> (8 * c16)) & 0xffu;
      if (n) dbase = atomicAdd(&a.pool_cnt[static_cast<size_t>(qj) * katoann_hip::kCntStride], n);
    }
    bool ovf = false;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint32_t q = __shfl(slotq, static_cast<int>(16u * f + 4 * g16 + r));
      const uint32_t b0 = __shfl(dbase, 16 * g16 + r);
      uint32_t idx = (ex >> (8 * r)) & 0xffu;
      int32_t av[T];
#pragma unroll
      for (int t = 0; t < T; ++t) av[t] = acc[t][r];
      uint32_t m = bits[r];
      while (m) {
        const int t = __builtin_ctz(m);
        m &= m - 1u;
        const int32_t l2 = qnr[r] - 2 * g_pick(av, t) - static_cast<int32_t>((oddm >> t) & 1u);
        const uint32_t pos = b0 + idx;
        if (pos < a.stride) a.pool[static_cast<size_t>(q) * a.stride + pos] = make_key(fscale(l2, a.step_sq), g_pick(xid, t));
        else ovf = true;
        ++idx;
      }
    }
    if (ovf) atomicOr(&a.state[0], 1u);
    __builtin_amdgcn_s_waitcnt(vm_imm(0));
    return total;
  }
  uint32_t prev[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) prev[r] = st.cnt[16 * f + 4 * g16 + r];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t slot = 16u * f + 4 * g16 + r;
    uint32_t idx = (ex >> (8 * r)) & 0xffu;
    int32_t av[T];
#pragma unroll
    for (int t = 0; t < T; ++t) av[t] = acc[t][r];
    uint32_t m = bits[r];
    while (m) {
      const int t = __builtin_ctz(m);
      m &= m - 1u;
      const int32_t l2 = qnr[r] - 2 * g_pick(av, t) - static_cast<int32_t>((oddm >> t) & 1u);
      st.key[so + idx] = make_key(fscale(l2, a.step_sq), g_pick(xid, t));
      st.meta[so + idx] = slot | ((prev[r] + idx) << 6);
      ++idx;
    }
    so += (Gm >> (8 * r)) & 0xffu;
  }
  // slot counts: lane c16 < 4 of group g16 adds slot 4 g16 + c16's keys
  if (c16 < 4) {
    uint32_t pv = prev[0];
#pragma unroll
    for (int r = 1; r < 4; ++r) pv = c16 == r ? prev[r] : pv;
    st.cnt[16 * f + 4 * g16 + c16] = pv + ((Gm >> (8 * c16)) & 0xffu);
  }
  st.n += total;
  return total;
}

template <int KC>
__global__ __launch_bounds__(256, (KC <= 6 ? 2 : 1)) void k_rank_gemm(GemmArgs a) {
  using G = GemmShape<KC>;
  constexpr int QTW = G::QTW, NST = G::NST;
  constexpr uint32_t P = 64 * KC, STAGE = G::STAGE, ROWS = G::ROWS;
  __shared__ __attribute__((aligned(16))) int8_t s_rows[NST][STAGE];
  __shared__ int32_t s_xn[NST][ROWS];
  __shared__ uint32_t s_id[NST][ROWS];
  __shared__ uint64_t s_key[4][G_CAP];
  __shared__ uint32_t s_meta[4][G_CAP];
  __shared__ uint32_t s_cnt[4][64];
  const int lane = lane_id(), w = threadIdx.x >> 6, g16 = lane >> 4, c16 = lane & 15;
  GStage st{s_key[w], s_meta[w], s_cnt[w], 0u};
  st.cnt[lane] = 0;
  const uint32_t ib = a.item_off[a.band * a.nlist], ie = a.item_off[(a.band + 1) * a.nlist];
  uint32_t n_kept = 0;

  // copy cursor: the next stage to copy (item, its descriptor, stage within)
  uint32_t cit = ib + blockIdx.x, cst = 0, nissued = 0;
  uint4 cd = cit < ie ? g_item(a.items, cit) : uint4{0, 0, 0, 0};
  auto issue = [&]() {
    while (cit < ie && cd.w <= cd.z) {  // an item without rows has no stage
      cit += gridDim.x;
      if (cit < ie) cd = g_item(a.items, cit);
    }
    if (cit >= ie) return false;
    const int slot = static_cast<int>(nissued % NST);
    const uint32_t r0 = cd.z + ROWS * cst;
    const uint32_t row = min(r0 + 16u * w + c16, cd.w - 1);
    const int8_t *rp = a.plane + static_cast<size_t>(row) * P + 16 * g16;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) g_dma16(rp + 64 * kc, &s_rows[slot][(w * KC + kc) * 1024]);
    const uint32_t rl = min(r0 + 16u * w + static_cast<uint32_t>(lane & 15), cd.w - 1);
    if (lane < 16) {
      g_dma4(a.xnorm + rl, &s_xn[slot][16 * w]);
      g_dma4(a.ids + rl, &s_id[slot][16 * w]);
    }
    ++nissued;
    if (ROWS * (cst + 1) < cd.w - cd.z) {
      ++cst;
    } else {
      cst = 0;
      cit += gridDim.x;
      if (cit < ie) cd = g_item(a.items, cit);
    }
    return true;
  };
#pragma unroll
  for (int i = 0; i + 1 < NST; ++i) issue();

  uint32_t ncomp = 0;  // stages ranked so far
  // query ids of an item's wave share, loaded one item ahead so that the loads
  // of its codes need no wait (a wait for a load issued after copies in flight
  // would wait for those copies too)
  auto share_ids = [&](uint32_t item, uint32_t (&qid)[QTW], uint32_t &sq) {
    const uint4 e = g_item(a.items, item);
    const uint32_t q0 = 16u * QTW * w, nqw = e.y > q0 ? min(e.y - q0, 16u * QTW) : 0u;
    const uint32_t *ql = a.qlist + e.x + q0;
#pragma unroll
    for (int f = 0; f < QTW; ++f) qid[f] = nqw ? ql[min(16u * f + c16, nqw - 1)] : 0u;
    sq = nqw ? ql[min(static_cast<uint32_t>(lane), nqw - 1)] : 0u;
  };
  uint32_t nqid[QTW], nslotq = 0;
  if (ib + blockIdx.x < ie) share_ids(ib + blockIdx.x, nqid, nslotq);
  for (uint32_t it = ib + blockIdx.x; it < ie; it += gridDim.x) {
    const uint4 d = g_item(a.items, it);
    const uint32_t nq = d.y, rows = d.w - d.z, nstage = (rows + ROWS - 1) / ROWS;
    const uint32_t q0 = 16u * QTW * w;
    const uint32_t nqw = nq > q0 ? min(nq - q0, 16u * QTW) : 0u;
    const uint32_t nfw = (nqw + 15) / 16;
    // the wave's queries: lane c16 of tile f holds query 16 f + c16 (the last
    // one again past the wave's share); lane s holds the query of slot s
    uint32_t qid[QTW];
#pragma unroll
    for (int f = 0; f < QTW; ++f) qid[f] = nqid[f];
    const uint32_t slotq = nslotq;
    const bool slotok = static_cast<uint32_t>(lane) < nqw;
    if (it + gridDim.x < ie) share_ids(it + gridDim.x, nqid, nslotq);
    v4i A[QTW][KC];
    int32_t H4[QTW][4], qn16[QTW], lb16[QTW];
#pragma unroll
    for (int f = 0; f < QTW; ++f) {
      const int8_t *src = a.qhi + static_cast<size_t>(qid[f]) * a.dc + 16 * g16;
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) A[f][kc] = g_load16(src + 64 * kc);
      qn16[f] = a.qnorm[qid[f]];
      lb16[f] = a.lbound[qid[f]];
    }
    // every operand of the item in registers before its stages: the stage loop
    // then waits only for its copies (a wait for a register load inside it
    // would be a wait for every copy in flight)
    __builtin_amdgcn_s_waitcnt(vm_imm(0));
#pragma unroll
    for (int f = 0; f < QTW; ++f) {
      // H = floor((qn - L) / 2), or 2^30 (above every accumulator) past the share
      const int32_t H = 16u * f + c16 < nqw ? (qn16[f] - lb16[f]) >> 1 : (1 << 30);
#pragma unroll
      for (int r = 0; r < 4; ++r) H4[f][r] = __shfl(H, 4 * g16 + r);
    }
    for (uint32_t s = 0; s < nstage; ++s, ++ncomp) {
      wait_stage<KC>(nissued - ncomp - 1);
      __builtin_amdgcn_s_barrier();
      issue();
      if (!nfw) continue;
      const int slot = static_cast<int>(ncomp % NST);
      const uint32_t r0 = d.z + ROWS * s;
      const uint32_t aR = g_lds_addr(&s_rows[slot][0]) + 16u * lane;
      // the stage's row tiles in two halves of 2 (rows past the item are copies
      // of its last row, outside vmask): B operands, norms, ids, then every
      // query tile of the wave against them
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (r0 + 32u * h >= d.w) break;
        v4i b[2][KC];
        int32_t xn[2], sd[2];
        uint32_t xid[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int rt = 2 * h + u;
#pragma unroll
          for (int kc = 0; kc + 4 <= KC; kc += 4)
            g_rd4(aR + (rt * KC + kc) * 1024, b[u][kc], b[u][kc + 1], b[u][kc + 2], b[u][kc + 3]);
#pragma unroll
          for (int kc = KC & ~3; kc < KC; ++kc) g_rd1(aR + (rt * KC + kc) * 1024, b[u][kc]);
          g_rd_pair(g_lds_addr(&s_xn[slot][16 * rt + c16]), g_lds_addr(&s_id[slot][16 * rt + c16]), xn[u], xid[u]);
        }
        uint32_t vmask = 0, oddm = 0;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          sd[u] = -((xn[u] + 1) >> 1);
          vmask |= static_cast<uint32_t>(r0 + 32u * h + 16u * u + c16 < d.w) << u;
          oddm |= static_cast<uint32_t>(xn[u] & 1) << u;
        }
#pragma unroll
        for (int f = 0; f < QTW; ++f) {
          if (static_cast<uint32_t>(f) >= nfw) break;
          v4i acc[2];
#pragma unroll
          for (int u = 0; u < 2; ++u) acc[u] = g_mfma(A[f][0], b[u][0], v4i{sd[u], sd[u], sd[u], sd[u]});
#pragma unroll
          for (int kc = 1; kc < KC; ++kc)
#pragma unroll
            for (int u = 0; u < 2; ++u) acc[u] = g_mfma(A[f][kc], b[u][kc], acc[u]);
          // whether any row of the fragment passes: max over the lane's rows per slot
          bool any = false;
#pragma unroll
          for (int r = 0; r < 4; ++r) any |= max(acc[0][r], acc[1][r]) >= H4[f][r];
          if (!ballot(any)) continue;
          n_kept += g_keep<2>(st, a, acc, f, vmask, oddm, xid, H4[f], qn16[f], slotq, slotok);
        }
      }
    }
    if (st.n) g_flush(st, a, slotq, slotok);
  }
  if (lane == 0 && n_kept) atomicAdd(&a.state[1], n_kept);
}

template <int KC>
void gemm_launch(const GemmArgs &a, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL((k_rank_gemm<KC>), dim3(grid), dim3(256), 0, s, a);
}

}  // namespace

namespace katoann_hip {

uint32_t rank_gemm_queries(uint32_t P) {
  const uint32_t kc = P / 64;
  if (kc == 0 || kc > 8 || P % 64) return 0;
  return 64u * (kc <= 4 ? GemmShape<4>::QTW : GemmShape<8>::QTW);
}

// two blocks per CU fit the LDS up to 384 prefix bytes, one above
uint32_t rank_gemm_blocks_per_cu(uint32_t P) { return P <= 384 ? 2u : 1u; }

bool launch_rank_gemm(uint32_t band, const int8_t *plane, const int32_t *xnorm, const uint32_t *ids, uint32_t nlist,
                      uint32_t P, const int8_t *qhi, uint32_t dc, const int32_t *qnorm, const int32_t *lbound,
                      const uint32_t *qlist, const uint32_t *item_off, const uint4 *items, uint32_t stride,
                      double step_sq, uint64_t *pool, uint32_t *pool_cnt, uint32_t *state, uint32_t grid,
                      hipStream_t s) {
  const GemmArgs a{plane, xnorm, ids, qhi, dc, qnorm, lbound, qlist, item_off, items, nlist, stride, pool, pool_cnt,
                   state, band, step_sq};
  switch (P / 64) {
#define G_CASE(n)               \
  case n:                       \
    gemm_launch<n>(a, grid, s); \
    return true;
    G_CASE(1) G_CASE(2) G_CASE(3) G_CASE(4) G_CASE(5) G_CASE(6) G_CASE(7) G_CASE(8)
#undef G_CASE
    default:
      return false;
  }
}

}  // namespace katoann_hip
