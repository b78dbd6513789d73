// Host launchers of the gfx950 kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace katoann_hip {

// Each query's pool counter sits alone in a 128-byte line (the ranking pass
// updates it with atomics from every CU).
constexpr uint32_t kCntStride = 32;

// Hybrid ranking items: a (band, list) slot of at most qsq queries (qsq > 0)
// becomes query-stationary items of qsrt rows that hold all its queries, with
// kQsItem set in the item's query count; the other slots are split into items
// of qc queries and rt rows as usual.
constexpr uint32_t kQsItem = 0x80000000u;
__host__ __device__ inline bool slot_qs(uint32_t g, uint32_t qsq) { return qsq && g <= qsq; }
// Items of rows x g queries in chunks of qc queries and tiles of rt rows.
__host__ __device__ inline uint32_t chunk_items(uint32_t g, uint64_t rows, uint32_t qc, uint32_t rt) {
  return (g && rows) ? ((g + qc - 1) / qc) * static_cast<uint32_t>((rows + rt - 1) / rt) : 0u;
}
// Joined seed remainder (join = 1, with a seed band and the probe-list
// inversion): the queries whose seed band ranked the first R0 rows of a list
// longer than R0 rank the rest of it in the first probe-rank band (band 2)
// instead of a band of its own, so those rows are read once with the band's
// other queries of the list, and band 1 stays empty (no launch, no pool
// update after it). Such a slot of band 2 lists its gb own queries first,
// then its ga joined ones: part 1 is rows [0, min(R0, n)) x the gb queries,
// part 2 rows [min(R0, n), n) x all gb + ga.
constexpr uint32_t kJoinBand = 2;

// Cost classes of the ranking items. A ranking kernel deals a band's items to
// its waves in item order, and an item's time grows with the query fragments
// it streams (16 queries each): dealt in list order, some waves drew far more
// fragments than others and the band waited for them. Each band's items are
// therefore laid out by class, most fragments first (> 64, 49-64, 33-48,
// 25-32, 17-24, 13-16, 9-12, 7-8, 5-6, 3-4, 2, 1), and dealt in rounds.
constexpr uint32_t kItemClasses = 12;
constexpr uint32_t kPlanBands = 16;  // bands of one plan at most
// Words after the plan's state: per (band, class) item counts, then cursors.
constexpr uint32_t kItemHistWords = 2 * kPlanBands * kItemClasses;
__host__ __device__ inline uint32_t item_class(uint32_t nq) {
  const uint32_t nf = (nq + 15) / 16;
  return nf > 64 ? 0u : nf > 48 ? 1u : nf > 32 ? 2u : nf > 24 ? 3u : nf > 16 ? 4u : nf > 12 ? 5u
       : nf > 8 ? 6u : nf > 6 ? 7u : nf > 4 ? 8u : nf > 2 ? 9u : nf > 1 ? 10u : 11u;
}

void launch_normalize(float *x, uint32_t n, uint32_t d, hipStream_t s);
void launch_rotate(const float *X, uint32_t n, uint32_t d, const float *R, float *Y, hipStream_t s);
void launch_quantize(const float *Y, uint32_t n, uint32_t d, uint32_t dc, uint32_t P, float lo,
                     double inv_step, int8_t *hi, int8_t *lop, int64_t *norm, int32_t *hinorm,
                     int32_t *headnorm, int8_t *pre, uint16_t *codes, hipStream_t s);
void launch_minmax(const float *y, uint64_t n, float *part, uint32_t nblocks, hipStream_t s);
void launch_coarse(const int8_t *qhi, const int8_t *qlo, const int64_t *qnorm, uint32_t B, const int8_t *chi,
                   const int8_t *clo, const int64_t *cnorm, uint32_t nlist, uint32_t dc, double step_sq,
                   float *dist, int cus, hipStream_t s);
// Per-query fp32 coarse distances (LoadOptions::batched_coarse = false):
// dist[q][c] = l2_fp32(unrotated query q, float centroid c).
void launch_coarse_flat(const float *Q, uint32_t B, uint32_t d, const float *cen, uint32_t nlist, float *dist,
                        hipStream_t s);
size_t select_smem(uint32_t nlist, uint32_t np);
void launch_select(const float *dist, uint32_t B, uint32_t nlist, uint32_t np, uint32_t *probes, float *probe_d,
                   hipStream_t s);
// Probe-rank bands for band factor F = 2^shift; the work plan has one
// (band, list) slot per band and list.
// Without a seed (seed = 0), band b holds ranks [(F^b - 1) / (F - 1),
// (F^(b+1) - 1) / (F - 1)).
// With a seed of R0 rows, band 0 holds a query's nearest lists up to R0 rows
// (the first r0 ranks: the fewest whose lists, each counted up to R0 rows,
// reach R0), each list's first R0 rows only; band 1 the remaining rows of
// those lists; band 2 + j ranks [r0 + F (F^j - 1) / (F - 1),
// r0 + F (F^(j+1) - 1) / (F - 1)). Every query with R0 >= S then has a
// threshold before its wider bands, and no pool takes more than 2 R0 keys
// without one.
inline uint32_t plan_bands(uint32_t np, uint32_t shift, uint32_t seed = 0) {
  if (!seed) {
    const uint32_t v = ((1u << shift) - 1u) * (np - 1u) + 1u;
    return (31u - static_cast<uint32_t>(__builtin_clz(v))) / shift + 1u;
  }
  uint32_t nb = 2;  // the regular bands for the smallest r0 (1)
  uint64_t s = 1, step = 1ull << shift;
  while (s < np) {
    ++nb;
    s += step;
    step = step << shift > (1ull << 32) ? (1ull << 32) : step << shift;
  }
  return nb;
}
uint32_t plan_slots(uint32_t nlist, uint32_t np, uint32_t shift, uint32_t seed = 0);
// The inversion reads the probe lists (B x nprobe pairs, atomic slot counters)
// when nprobe is well below nlist, else the coarse matrix in query tiles
// (ascending queries per slot, no atomics: lists probed by most of the batch
// would serialise on their counters).
inline bool inv_from_probes(uint32_t nlist, uint32_t np) { return 16ull * np <= nlist; }
// 32-bit words of plan scratch for a batch (tile partial counts, item counts,
// block sums, seed ranks).
size_t plan_scratch_words(uint32_t B, uint32_t nlist, uint32_t np, uint32_t shift, uint32_t seed = 0);
// Inversion and work plan from the coarse matrix and the sorted probes. With
// check_npq, per-query counts outside [1, np] are clamped in place and flag
// state[2].
void launch_plan(const float *dist, const uint32_t *probes, const float *probe_d, const uint32_t *npq, bool check_npq,
                 uint32_t B, uint32_t np, const uint64_t *list_off, uint32_t nlist, uint32_t qc, uint32_t rt,
                 uint32_t qsq, uint32_t qsrt, uint32_t *scratch, uint32_t *gcnt, uint32_t *qoff, uint32_t *item_off,
                 uint32_t *qlist, uint32_t *totals, uint32_t *pool_cnt, uint32_t *pool_sorted, uint64_t *thr,
                 int32_t *lbound, uint32_t *state, uint32_t nstate, uint32_t shift, uint32_t seed, uint32_t join,
                 hipStream_t s);
// The joined seed remainder's per-list counts of a slot's own queries (first
// probe-rank band), in the plan scratch after the inversion's words.
inline size_t plan_join_gb_offset(uint32_t B, uint32_t nlist, uint32_t np, uint32_t shift, uint32_t seed) {
  return plan_scratch_words(B, nlist, np, shift, seed) - nlist;
}
void launch_fill32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t s);
// Queries per ranking item for ranking width P (0 if P is unsupported).
uint32_t rank_queries_per_item(uint32_t P);
// Item descriptors, each band's in cost-class order; ihist: the per (band,
// class) item counts the plan left after its state (nstate words), and their
// cursors.
void launch_items(const uint32_t *item_off, uint32_t nslots, uint32_t nlist, const uint64_t *list_off,
                  const uint32_t *gcnt, const uint32_t *qoff, uint32_t qc, uint32_t rt, uint32_t qsq, uint32_t qsrt,
                  uint32_t cap, uint32_t seed, uint32_t join, const uint32_t *gbv, uint32_t *ihist, uint4 *items,
                  hipStream_t s);
// Rows [first, end) of list c's n rows that slot band `band` ranks.
__host__ __device__ inline void slot_rows(uint32_t band, uint64_t n, uint32_t seed, uint64_t &first, uint64_t &end) {
  first = 0;
  end = n;
  if (seed && band == 0) end = n < seed ? n : seed;
  if (seed && band == 1) first = n < seed ? n : seed;
}
bool launch_rank_band(uint32_t band, const int8_t *plane, const int32_t *xnorm, const uint32_t *ids, uint32_t nlist,
                      uint32_t P, const int8_t *qhi, uint32_t dc, const int32_t *qnorm, const int32_t *lbound,
                      const uint32_t *qlist, const uint32_t *item_off, const uint4 *items, uint32_t stride,
                      double step_sq, uint64_t *pool, uint32_t *pool_cnt, uint32_t *state, uint32_t grid,
                      hipStream_t s);
// Wave-item ranking kernels (wave-independent, or a block of 4 waves sharing a
// query ring when `ring`): rows per wave for ranking width P and register
// variant 0 (large) or 1 (half); 0 if unsupported.
// Blocks of the wave kernel resident per CU (1 for the tiles that take a whole
// SIMD's registers per wave).
uint32_t rank_wave_blocks_per_cu(uint32_t P, uint32_t variant);
uint32_t rank_wave_rows(uint32_t P, uint32_t variant);
bool launch_rank_wave(uint32_t band, uint32_t variant, bool ring, const int8_t *plane, const int32_t *xnorm, const uint32_t *ids,
                      uint32_t nlist, uint32_t P, const int8_t *qhi, uint32_t dc, const int32_t *qnorm,
                      const int32_t *lbound, const uint32_t *qlist, const uint32_t *item_off, const uint4 *items,
                      uint32_t stride, double step_sq, uint64_t *pool, uint32_t *pool_cnt, uint32_t *state,
                      uint32_t grid, hipStream_t s);
// Query-stationary ranking kernel: a wave holds up to rank_qs_queries(P,
// variant) queries of a list as MFMA A operands and streams the item's rows
// (16 at a time, several tiles in flight) as B operands; 0 if unsupported.
uint32_t rank_qs_queries(uint32_t P, uint32_t variant);
// Hybrid launch (both forms per item, kQsItem): the wave form's tile of
// `variant`, the query-stationary form's variant 0.
bool launch_rank_hyb(uint32_t band, uint32_t variant, const int8_t *plane, const int32_t *xnorm, const uint32_t *ids,
                     uint32_t nlist, uint32_t P, const int8_t *qhi, uint32_t dc, const int32_t *qnorm,
                     const int32_t *lbound, const uint32_t *qlist, const uint32_t *item_off, const uint4 *items,
                     uint32_t stride, double step_sq, uint64_t *pool, uint32_t *pool_cnt, uint32_t *state,
                     uint32_t grid, hipStream_t s);
uint32_t rank_qs_blocks_per_cu(uint32_t P, uint32_t variant);
// Block-tiled ranking kernel (rank_gemm.hip): queries per item (4 waves of
// query tiles; 0 if unsupported, P above 512), blocks per CU, launch of one band.
uint32_t rank_gemm_queries(uint32_t P);
uint32_t rank_gemm_blocks_per_cu(uint32_t P);
bool launch_rank_gemm(uint32_t band, const int8_t *plane, const int32_t *xnorm, const uint32_t *ids, uint32_t nlist,
                      uint32_t P, const int8_t *qhi, uint32_t dc, const int32_t *qnorm, const int32_t *lbound,
                      const uint32_t *qlist, const uint32_t *item_off, const uint4 *items, uint32_t stride,
                      double step_sq, uint64_t *pool, uint32_t *pool_cnt, uint32_t *state, uint32_t grid,
                      hipStream_t s);
bool launch_rank_qs(uint32_t band, uint32_t variant, const int8_t *plane, const int32_t *xnorm, const uint32_t *ids,
                    uint32_t nlist, uint32_t P, const int8_t *qhi, uint32_t dc, const int32_t *qnorm,
                    const int32_t *lbound, const uint32_t *qlist, const uint32_t *item_off, const uint4 *items,
                    uint32_t stride, double step_sq, uint64_t *pool, uint32_t *pool_cnt, uint32_t *state,
                    uint32_t grid, hipStream_t s);
// Ranking over the full codes (exact distances, pools of the k smallest
// keys): rows per item, and the launch of one band.
uint32_t rank_full_rows();
uint32_t rank_full_blocks_per_cu();
void launch_rank_full(uint32_t band, const int8_t *hi, const int8_t *lo, const int64_t *xnorm, const uint32_t *ids,
                      uint32_t nlist, const int8_t *qhi, const int8_t *qlo, uint32_t dc, const int64_t *qnorm,
                      const uint64_t *thr, const uint32_t *qlist, const uint32_t *item_off, const uint4 *items,
                      uint32_t stride, double step_sq, uint64_t *pool, uint32_t *pool_cnt, uint32_t *state,
                      uint32_t grid, hipStream_t s);
// bf16 precision: rows to bf16 (dp apart, zero padded) and their norms;
// ranking of one band in the reference's bf16 arithmetic (pools of the k
// nearest, keys of order-preserving float bits); the k nearest out of the
// pools.
void launch_to_bf16(const float *X, uint32_t n, uint32_t d, uint32_t dp, uint16_t *out, float *norms,
                    hipStream_t s);
void launch_bf16_pairs(const uint16_t *q, const uint16_t *x, const float *qn, const float *xn, uint32_t m,
                       uint32_t n, uint32_t dim, float *out, hipStream_t s);
uint32_t rank_bf16_rows();
uint32_t rank_bf16_blocks_per_cu();
void launch_rank_bf16(uint32_t band, const uint16_t *rows, const float *xnorm, const uint32_t *ids, uint32_t nlist,
                      const uint16_t *qbf, const float *qnorm, uint32_t dp, const uint64_t *thr,
                      const uint32_t *qlist, const uint32_t *item_off, const uint4 *items, uint32_t stride,
                      uint64_t *pool, uint32_t *pool_cnt, uint32_t *state, uint32_t grid, hipStream_t s);
void launch_emit_bf16(const uint64_t *pool, uint32_t stride, const uint32_t *pool_cnt, uint32_t B, uint32_t K,
                      uint32_t *out_ids, float *out_dists, uint64_t *surv, uint32_t *nsurv, hipStream_t s);
uint32_t refine_cap(uint32_t S);
// Pools back to their S smallest keys after a band, and the thresholds; with
// keep_sorted false and S above 1,024 the kept keys stay in no particular
// order (the refinement sorts them).
bool launch_update(uint64_t *pool, uint32_t *pool_cnt, uint32_t *pool_sorted, uint64_t *thr, int32_t *lbound,
                   uint32_t B, uint32_t S, uint32_t stride, uint32_t P, double step_sq, bool keep_sorted,
                   hipStream_t s);
bool launch_refine(const uint64_t *pool, uint32_t stride, const uint32_t *pool_cnt, const uint64_t *thr,
                   uint32_t B, uint32_t S, uint32_t K, const int8_t *hi, const int8_t *lo, const int64_t *norm,
                   const uint32_t *id_row, uint32_t npts, uint32_t dc, const int8_t *qhi, const int8_t *qlo,
                   const int64_t *qnorm, double step_sq, uint32_t *out_ids, float *out_dists, uint64_t *surv,
                   uint32_t *nsurv, hipStream_t s);
void launch_sq16_pairs(const int8_t *qh, const int8_t *ql, const int64_t *qn, const int8_t *xh, const int8_t *xl,
                       const int64_t *xn, uint32_t m, uint32_t n, uint32_t dc, uint32_t qstride, uint32_t xstride,
                       int full, double step_sq, float *out, hipStream_t s);

// Learned nprobe (mlp.hip).
void launch_gemm_bf16(uint32_t M, uint32_t N, uint32_t K, const void *A, size_t sam, size_t sak, const void *B,
                      size_t sbn, size_t sbk, const float *bias, const float *gate, size_t ldg, int relu, float *C32,
                      void *C16, size_t ldc, hipStream_t s);
void launch_to_bf16(const float *x, void *y, uint64_t n, hipStream_t s);
void launch_mse(const float *p, const float *y, uint32_t n, float *grad, float *loss, hipStream_t s);
void launch_pinball(const float *p, const float *y, uint32_t n, float tau, float *grad, float *loss,
                    hipStream_t s);
void launch_colsum(const float *g, uint32_t R, uint32_t C, float *out, hipStream_t s);
void launch_adam(float *w, const float *g, float *m, float *v, uint32_t n, float lr, float b1, float b2, float eps,
                 uint32_t step, void *w16, hipStream_t s);
void launch_features(const float *pd, uint32_t B, uint32_t np, uint32_t F, const float *mean, const float *inv_std,
                     float *x32, void *x16, hipStream_t s);
void launch_nprobe_of(const float *z, uint32_t B, float y_mean, float y_std, float margin, uint32_t nmin,
                      uint32_t nmax, uint32_t *npq, hipStream_t s);

}  // namespace katoann_hip
