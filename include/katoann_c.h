/* C ABI of the gfx950 KatoANN library (libkatoann.so).
 *
 * This is the drop-in boundary for the batched IVF search: the entry points are
 * what the reference's front ends bind (python/katoann_module.cpp:39-128, the
 * CLI tools/katoann.cpp), expressed with plain pointers and sizes. The C++ API
 * of the reference (include/katoann/index.h) is rebuilt on top of this ABI in
 * include/katoann/index.h.
 *
 * Errors: every int-returning function returns KATOANN_OK or the class of the
 * exception the reference throws (invalid_argument, runtime_error, logic_error);
 * katoann_last_error() gives the message (thread-local).
 * Queries, ids and dists are host memory unless stated otherwise. A loaded index
 * owns all its device memory; concurrent searches on one index are safe.
 */
#ifndef KATOANN_C_H_
#define KATOANN_C_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum katoann_status {
  KATOANN_OK = 0,
  KATOANN_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
  KATOANN_RUNTIME_ERROR = 2,    /* std::runtime_error */
  KATOANN_LOGIC_ERROR = 3       /* std::logic_error */
};

enum katoann_precision { KATOANN_SQ16 = 0, KATOANN_BF16 = 1 };          /* index.h:32-35 */
enum katoann_ranking { KATOANN_RANK_FULL = 0, KATOANN_RANK_HIGH = 1, KATOANN_RANK_PREFIX = 2 }; /* index.h:45-49 */
enum katoann_selection { KATOANN_SEL_THRESHOLD = 0, KATOANN_SEL_LOCAL = 1 };                   /* index.h:51-54 */
enum katoann_engine { KATOANN_ENGINE_AMX = 0, KATOANN_ENGINE_AVX512 = 1, KATOANN_ENGINE_GFX950 = 2 }; /* kernels.h:19 */

/* BuildOptions (index.h:22-30) */
typedef struct katoann_build_options {
  uint32_t nlist;
  uint32_t train_size;
  uint32_t kmeans_iters;
  uint64_t seed;
  int32_t normalized;
  uint32_t rotation_sample;
  uint32_t jacobi_sweeps;
} katoann_build_options;

/* LoadOptions (index.h:37-43); use_amx is accepted and ignored.
 * precision KATOANN_BF16 keeps the unrotated vectors as bf16 (no codes, no
 * .dimc read) and serves Ranking::kFull only (src/index.cpp:777-786, 885-889);
 * batched_coarse = 0 selects the per-query fp32 coarse search
 * (src/index.cpp:832-844), which bf16 precision always uses. */
typedef struct katoann_load_options {
  int32_t precision;
  int32_t use_amx;
  uint32_t prefix_dims;
  int32_t batched_coarse;
} katoann_load_options;

/* SearchParams (index.h:56-62) */
typedef struct katoann_search_params {
  uint32_t k;
  uint32_t nprobe;
  int32_t ranking;
  uint32_t survivors;
  int32_t selection;
} katoann_search_params;

typedef struct katoann_index katoann_index;

const char *katoann_last_error(void);
void katoann_default_build_options(katoann_build_options *o);
void katoann_default_load_options(katoann_load_options *o);
void katoann_default_search_params(katoann_search_params *p);

/* Index::build (index.h:74-75, src/index.cpp:419-545): k-means partition and
 * PCA rotation on the GPU (the vectors stream from host memory through pinned
 * buffers), index files under prefix. The method is this library's own (see
 * csrc/build.hip); the files, options and errors are the reference's. */
int katoann_build(const float *data, uint32_t n, uint32_t dim, const katoann_build_options *opt,
                  const char *prefix);

/* No counterpart in the reference (its build takes host vectors and writes
 * files): the same build for vectors already in device memory, for callers
 * whose data lives in HBM; katoann_load_arrays can load the result with no
 * files. Every pointer is device memory except `rotated`: centroids (nlist x
 * dim), order (n: the ids grouped by list, ascending inside each list), sizes
 * (nlist), rotation (dim x dim, rows by decreasing variance; NULL skips the
 * PCA); *rotated (host, may be NULL) = 1 when a rotation was computed. */
int katoann_build_device(const float *data, uint32_t n, uint32_t dim, const katoann_build_options *opt,
                         float *centroids, uint32_t *order, uint32_t *sizes, float *rotation, int32_t *rotated);

/* The build's nearest-centroid assignment (the reference assigns every vector
 * to its nearest centroid, src/index.cpp:457-463): n vectors (device, n x dim)
 * against nlist centroids (device, nlist x dim), squared L2 in fp32, ties to
 * the lowest index; out (device, n). */
int katoann_assign_lists(const float *x, uint32_t n, uint32_t dim, const float *centroids, uint32_t nlist,
                         uint32_t *out);

/* The build's Lloyd iterations (the reference's k-means: src/kmeans.cpp:191-278)
 * from the given centroids (device, nlist x dim, replaced by the result) over
 * n training vectors (device): assignment as katoann_assign_lists, each centroid
 * the mean of its vectors (fp64, ascending vector order), a centroid left
 * without vectors unchanged. */
int katoann_kmeans_device(const float *train, uint32_t n, uint32_t dim, float *centroids, uint32_t nlist,
                          uint32_t iters);

/* The build's rotation (the reference's PCA: src/index.cpp:276-380, 516-533)
 * from m sample vectors (device, m x dim): rotation (device, dim x dim), row j
 * the eigenvector of the j-th largest eigenvalue of the fp64 sample covariance,
 * signed so its first largest-magnitude component is positive; eigenvalues
 * (host, dim, decreasing; may be NULL); *rotated (host, may be NULL) = 0 and
 * nothing written when m < 2 or dim < 2. */
int katoann_pca_rotation_device(const float *sample, uint32_t m, uint32_t dim, float *rotation, double *eigenvalues,
                                int32_t *rotated);

int katoann_index_create(katoann_index **out);
void katoann_index_destroy(katoann_index *h);

/* Index::load (index.h:77, src/index.cpp:777-807): reads the files and builds
 * the device-resident index (rotation, 16-bit codes, norms) on the GPU. */
int katoann_load(katoann_index *h, const char *prefix, const katoann_load_options *opt);

/* Index::search (index.h:80-81, src/index.cpp:879-904): nq x dim queries, not
 * normalised here; ids and dists are nq x k, nearest first. */
int katoann_search(const katoann_index *h, const float *queries, uint32_t nq, const katoann_search_params *p,
                   uint32_t *ids, float *dists);

/* Index loads (no reference counterpart: the reference reads the whole
 * .invlist into host memory, src/index.cpp:568-591) stream the vectors to the
 * GPU through two pinned staging buffers of this many bytes each (default
 * 64 MiB, 4 KiB to 4 GiB), after every file check has passed. Process-wide. */
int katoann_set_load_chunk(uint64_t bytes);

/* Accessors (index.h:83-93) */
uint32_t katoann_dim(const katoann_index *h);
uint32_t katoann_nlist(const katoann_index *h);
uint32_t katoann_size(const katoann_index *h);
int32_t katoann_normalized(const katoann_index *h);
int32_t katoann_engine(const katoann_index *h);
int32_t katoann_precision(const katoann_index *h);
uint32_t katoann_prefix_dims(const katoann_index *h);
int32_t katoann_rotated(const katoann_index *h);
int32_t katoann_batched_coarse(const katoann_index *h);
int katoann_list_sizes(const katoann_index *h, uint32_t *out); /* nlist entries */
uint64_t katoann_memory_bytes(const katoann_index *h);

/* ---- extensions (no reference counterpart) ----------------------------- */

/* A trained learned-nprobe model (no reference counterpart): an MLP over each
 * query's nfeat smallest sorted coarse distances (x_0 = d_0, x_i = d_i - d_0,
 * standardised), two ReLU layers of `hidden`, one output z; the query's
 * nprobe is ceil(2^(z * y_std + y_mean + margin)) clamped to [nmin, nprobe of
 * the search]. Device pointers, owned by the caller; weights row-major
 * (out x in), bf16. */
typedef struct katoann_nprobe_model {
  uint32_t nfeat, hidden;
  const float *feat_mean, *feat_inv_std; /* nfeat */
  const void *w1;                        /* hidden x nfeat */
  const float *b1;                       /* hidden */
  const void *w2;                        /* hidden x hidden */
  const float *b2;                       /* hidden */
  const void *w3;                        /* 1 x hidden */
  const float *b3;                       /* 1 */
  float y_mean, y_std, margin;
  uint32_t nmin;
} katoann_nprobe_model;

/* Optional inputs and outputs of one search. */
typedef struct katoann_search_extras {
  const uint32_t *nprobe_per_query; /* nq entries in [1, nprobe]: per-query probe count, or NULL */
  int32_t normalize;                /* normalise the queries first, as the Python module and CLI do */
  int32_t device_io;                /* queries, ids, dists (and the outputs below) are device pointers */
  uint32_t *probes;                 /* nq x nprobe: probed lists, nearest first (UINT32_MAX past a query's count) */
  float *probe_dists;               /* nq x nprobe: their coarse distances */
  uint64_t *survivors;              /* nq x max(survivors, k) (k for full ranking): ranking keys (float bits << 32 | id), ascending */
  uint32_t *n_survivors;            /* nq */
  float *rotated;                   /* nq x dim: rotated (normalised) queries */
  uint16_t *codes;                  /* nq x Dc: 16-bit query codes */
  void *stream;                     /* hipStream_t to run on (device_io only), or NULL for the index's own */
  const katoann_nprobe_model *nprobe_model; /* per-query nprobe from the model (exclusive with nprobe_per_query) */
  uint32_t *nprobe_used;            /* nq: the per-query nprobe the search used */
  int32_t coarse_only;              /* stop after the probe selection (probes / probe_dists only) */
  uint64_t max_bytes;               /* nonzero: an overflow rerun whose buffers would exceed this many bytes
                                       stops with KATOANN_RUNTIME_ERROR (and *needed_bytes set) instead */
  uint64_t *needed_bytes;           /* out: device bytes of the search's buffers (or of the refused rerun) */
} katoann_search_extras;

int katoann_search_ex(const katoann_index *h, const float *queries, uint32_t nq, const katoann_search_params *p,
                      const katoann_search_extras *x, uint32_t *ids, float *dists);

/* Allocation-free search. No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). katoann_search_ex with device
 * pointers (x->device_io = 1, x required), run in a static device array of
 * KATOANN_NOALLOC_SCRATCH_BYTES on each GPU instead of buffers of its own,
 * on x->stream (NULL: the null stream); it calls no hipMalloc, hipFree or
 * other allocating routine. A search whose buffers exceed the array (e.g.
 * B x nlist x 4 bytes of coarse distances at large batches) is refused with
 * KATOANN_INVALID_ARGUMENT; a survivor-pool overflow is rerun within the
 * array, or refused with KATOANN_RUNTIME_ERROR. No autotuning, graph replay
 * or stage timing. Calls on one GPU share its array: the library serialises
 * them, and the array keeps the last call's contents. Results equal
 * katoann_search_ex's. */
#define KATOANN_NOALLOC_SCRATCH_BYTES (256ull << 20)
int katoann_search_noalloc(const katoann_index *h, const float *queries, uint32_t nq, const katoann_search_params *p,
                           const katoann_search_extras *x, uint32_t *ids, float *dists);

/* Load from arrays instead of files: vectors (n x dim, grouped by list in list
 * order; host memory, or device memory when vectors_on_device), ids in the same
 * order, list_sizes (nlist), centroids (nlist x dim) and the rotation (dim x dim,
 * row j = basis j, or NULL) in host memory. prefix_dims as stored in .dimc. */
int katoann_load_arrays(katoann_index *h, uint32_t dim, uint32_t nlist, uint32_t n, int32_t normalized,
                        const float *centroids, const uint32_t *list_sizes, const uint32_t *ids,
                        const float *vectors, int32_t vectors_on_device, const float *rotation,
                        uint32_t prefix_dims, const katoann_load_options *opt);

/* Scale of the codes: lo, 65535 / (hi - lo), step^2 (src/index.cpp:648-653). */
int katoann_index_scale(const katoann_index *h, float *lo, double *inv_step, double *step_sq);

/* The unsigned 16-bit codes of the loaded index (n x Dc in list order, nlist x Dc)
 * and the ids in list order; any pointer may be NULL. */
int katoann_export_codes(const katoann_index *h, uint16_t *codes, uint16_t *centroid_codes, uint32_t *ids);
/* The same for the rows of the given lists only, list after list in the
 * order given (codes: rows x Dc, ids: rows); either pointer may be NULL. */
int katoann_export_lists(const katoann_index *h, const uint32_t *lists, uint32_t nl, uint16_t *codes,
                         uint32_t *ids);

/* Per-stage GPU times of the last search on this index, in ms:
 * [0] normalise + rotate + quantise, [1] coarse, [2] probe selection,
 * [3] inversion + plan, [4] ranking kernels (all bands), [5] survivor pool
 * updates between bands, [6] survivors + refinement + top-k.
 * Timing is recorded only while enabled. */
#define KATOANN_NSTAGES 7
int katoann_set_timing(katoann_index *h, int32_t on);
int katoann_stage_times(const katoann_index *h, float *ms, int32_t n);
/* Ranking-launch durations (ms) of each probe-rank band of the last timed
 * search, band 0 first; 0 past its bands. */
int katoann_band_times(const katoann_index *h, float *ms, int32_t n);
/* Work of the last search: [0] (query, list) pairs, [1] ranking items,
 * [2] candidates kept by the ranking filter, [3] probe-rank bands. */
int katoann_last_work(const katoann_index *h, uint64_t *out, int32_t n);

/* Tile parameters of the ranking pass: rows per work item (multiple of 64). */
int katoann_set_rank_rows(katoann_index *h, uint32_t rows);
/* Probe-rank bands of the ranking pass: band b holds ranks [(F^b - 1) / (F - 1),
 * (F^(b+1) - 1) / (F - 1)) with F = 2^log2_factor (1..12); the queries' thresholds
 * tighten between bands. Default 5 (F = 32). */
int katoann_set_band_factor(katoann_index *h, uint32_t log2_factor);
/* Seed band (no reference counterpart): with rows = R0 > 0 the ranking pass
 * first ranks each query's nearest lists up to R0 rows in all (at most the
 * first R0 rows of each; R0 is raised to max(survivors, k)), then the rest of
 * those lists, then probe-rank bands of factor F from the next rank. Every
 * query thus has a threshold before its wider bands. 0 = bands from rank 0.
 * Default 1024. Results do not depend on it. */
int katoann_set_seed_rows(katoann_index *h, uint32_t rows);
/* Joined seed remainder (no reference counterpart): with on != 0, a seed band
 * and nprobe <= nlist / 16, the rest of each query's seed lists (rows past R0)
 * is ranked in the first probe-rank band with that band's queries of the same
 * list, on the thresholds the seed band produced, instead of in a band of its
 * own: those rows are read once and one pool update goes. Default on.
 * Results do not depend on it. */
int katoann_set_seed_join(katoann_index *h, int32_t on);
/* Survivor pool room per query beyond S for one band's kept keys (no
 * reference counterpart): 0 (default) sizes it per search. A band that keeps
 * more runs the batch again with 4x the room; later searches of the shape
 * start from the room that sufficed. Results do not depend on it. */
int katoann_set_pool_allowance(katoann_index *h, uint32_t slots);
/* Ranking kernel: 0 = block items (a list's queries staged into LDS, rows
 * streamed), 1 = wave items (a wave keeps up to 16 T rows of a list in
 * registers and streams up to queries_per_item of its queries), 2 = block items
 * of 4 x 16 T rows whose 4 waves share the queries through an LDS ring,
 * 3 = query-stationary wave items (a wave keeps up to 64 of a list's queries in
 * registers, 32 with variant 1, and streams the list's rows 16 at a time with
 * several tiles in flight; queries_per_item gives its rows per item),
 * 4 = both forms in one launch: lists with at most 64 of a band's queries
 * query-stationary (256 rows per item), the others as kernel 1 (variant 0),
 * 5 = block-tiled items (4 waves hold up to 256 of a list's queries, 128 above
 * 256 prefix dims, in registers; the list's rows stream through an LDS ring of
 * 64-row stages copied from HBM ahead of use; queries_per_item gives its rows
 * per item; ranked widths up to 512 bytes, wider ones run kernel 1 with 1024
 * queries per item);
 * variant 0 or 1 picks the larger or smaller T; variant 2 (kernel 1 only) a
 * smaller T with query codes loaded two fragments ahead instead of one;
 * variant 3 (kernel 1 only) a larger T at one wave per SIMD.
 * Default: 1, variant 0, 1024. */
int katoann_set_rank_kernel(katoann_index *h, uint32_t kernel, uint32_t variant, uint32_t queries_per_item);
/* Graph replay (no reference counterpart): each search context captures its
 * kernel sequence (normalisation through top-k) into a hipGraph once and
 * replays it while shapes, parameters and buffers stay the same (default on;
 * searches with stage timing on run eagerly). */
int katoann_set_graphs(katoann_index *h, int32_t on);
/* Tile autotuning (no reference counterpart): with it on, the first search of
 * each shape (ranking width, nlist, batch rounded up to a power of two,
 * nprobe, max(survivors, k)) times the candidate ranking tiles (kernel, row
 * tile, queries per item, band factor, seed rows, joined seed remainder) on its own inputs and keeps the fastest
 * for that shape; results are the same whatever the tile. Default off. */
int katoann_set_autotune(katoann_index *h, int32_t on);
/* The tile a search of this shape runs with: kernel, variant, queries per
 * item, log2 band factor, seed rows, joined seed remainder (out: 6 entries). */
int katoann_tile_config(const katoann_index *h, uint32_t nq, const katoann_search_params *p, uint32_t *out6);

/* Kernel-level distances (reference kernels.h:34-52) on the GPU, host memory in
 * and out, row-major planes (m x padded_dim and n x padded_dim bytes), norms of
 * the unsigned codes; out is m x n. High-byte form: the leading padded_dim bytes
 * of rows that are plane_dim wide (0 means padded_dim). */
int katoann_sq16_l2(const uint8_t *qhi, const uint8_t *qlo, const uint8_t *xhi, const uint8_t *xlo,
                    const int64_t *qnorm, const int64_t *xnorm, double step_sq, uint32_t m, uint32_t n,
                    uint32_t padded_dim, float *out);
int katoann_sq16_high_l2(const uint8_t *qhi, const uint8_t *xhi, const int64_t *qnorm, const int64_t *xnorm,
                         double step_sq, uint32_t m, uint32_t n, uint32_t padded_dim, float *out,
                         uint32_t plane_dim);
/* bf16 distances (reference kernels.h:64-67, bf16_l2; its AVX-512 arithmetic)
 * on the GPU: m queries and n rows of dim bf16 values each, row-major (dim a
 * multiple of 2), their norms; out[i][j] = (qnorm + xnorm) - 2 dot. */
int katoann_bf16_l2(const uint16_t *q, const uint16_t *x, const float *qnorm, const float *xnorm, uint32_t m,
                    uint32_t n, uint32_t dim, float *out);

/* Kernels of the learned-nprobe MLP on device pointers (no reference
 * counterpart), on `stream` (hipStream_t, NULL for the null stream).
 * katoann_mlp_gemm: C[m][n] = act(sum_k A(m,k) B(n,k) + bias[n]), zero where
 * gate[m][n] <= 0 when gate is given; A(m,k) = A[m sam + k sak] and
 * B(n,k) = B[n sbn + k sbk] in bf16; C32 (fp32) and/or C16 (bf16), ldc apart. */
int katoann_mlp_gemm(uint32_t M, uint32_t N, uint32_t K, const void *A, size_t sam, size_t sak, const void *B,
                     size_t sbn, size_t sbk, const float *bias, const float *gate, size_t ldg, int32_t relu,
                     float *C32, void *C16, size_t ldc, void *stream);
int katoann_mlp_to_bf16(const float *x, void *y, uint64_t n, void *stream);
/* grad = 2 (pred - y) / n; *loss += mean squared error (loss may be NULL) */
int katoann_mlp_mse(const float *pred, const float *y, uint32_t n, float *grad, float *loss, void *stream);
/* pinball loss at quantile tau: grad = -(tau - [y < pred]) / n; *loss += mean loss */
int katoann_mlp_pinball(const float *pred, const float *y, uint32_t n, float tau, float *grad, float *loss,
                        void *stream);
int katoann_mlp_colsum(const float *g, uint32_t rows, uint32_t cols, float *out, void *stream);
/* Adam step `step` (1-based) on fp32 weights; w16 (bf16 copy) may be NULL */
int katoann_mlp_adam(float *w, const float *g, float *m, float *v, uint32_t n, float lr, float b1, float b2,
                     float eps, uint32_t step, void *w16, void *stream);
/* Model inputs from sorted coarse distances (nq x nprobe): x32 and/or x16 (nq x nfeat) */
int katoann_mlp_features(const float *probe_dists, uint32_t nq, uint32_t nprobe, uint32_t nfeat, const float *mean,
                         const float *inv_std, float *x32, void *x16, void *stream);

/* Device bytes one search of nq queries with these parameters and extras
 * allocates (its per-search buffers at the first pool allowance), after the
 * same argument checks as katoann_search_ex. */
int katoann_search_footprint(const katoann_index *h, uint32_t nq, const katoann_search_params *p,
                             const katoann_search_extras *x, uint64_t *bytes);

/* ---- continuous-batching server (no reference counterpart) --------------
 * Replaces the synchronous idle batching of the CLI's `load` command
 * (tools/katoann.cpp:267-337): requests (host queries, any size, each with its
 * own k / nprobe / survivors / model) queue as they arrive; up to
 * max_inflight batches run at once, each on its own HIP stream, so that their
 * kernels overlap on the GPU. A batch is the queue's oldest request plus the
 * following requests with the same parameters, up to max_batch queries. It is
 * admitted when a slot is free and its search footprint fits the HBM budget
 * beside the batches in flight; a request that can never fit is refused at
 * submission. Each query's result does not depend on the batch it ran in. */
typedef struct katoann_server katoann_server;

typedef struct katoann_server_options {
  uint32_t max_inflight; /* batches on the GPU at once (default 4) */
  uint32_t max_batch;    /* queries per batch (default 4096) */
  uint64_t hbm_budget;   /* bytes for search buffers; 0: 90% of the free HBM at creation */
} katoann_server_options;

typedef struct katoann_server_stats {
  uint64_t requests, queries, batches; /* completed */
  double p50_ms, p95_ms, p99_ms, max_ms; /* arrival (submission) to results ready; sorted[min(n-1, ceil(q n)-1)] */
  double mean_batch;                   /* queries per batch */
  double achieved_qps;                 /* queries / (last completion - first arrival) */
  uint64_t budget, peak_reserved;      /* bytes */
  uint32_t peak_inflight;
  uint64_t peak_used;                  /* bytes: the largest search buffers a batch ran with */
} katoann_server_stats;

void katoann_default_server_options(katoann_server_options *o);
/* The index must outlive the server. */
int katoann_server_create(katoann_index *h, const katoann_server_options *o, katoann_server **out);
/* Drains the queue, then stops. */
void katoann_server_destroy(katoann_server *s);
/* Queues nq x dim host queries (copied; normalised if normalize) for a search
 * with p and an optional model (copied; its device weights must stay valid
 * until the request completes); *ticket names the request. */
int katoann_server_submit(katoann_server *s, const float *queries, uint32_t nq, const katoann_search_params *p,
                          const katoann_nprobe_model *model, int32_t normalize, uint64_t *ticket);
/* *done = 1 once the request's results are ready (or it failed). */
int katoann_server_poll(katoann_server *s, uint64_t ticket, int32_t *done);
/* Waits for the request, writes nq x k ids and dists (host) and its latency in
 * ms (may be NULL), and forgets the ticket; returns the search's error, if any. */
int katoann_server_wait(katoann_server *s, uint64_t ticket, uint32_t *ids, float *dists, double *latency_ms);
/* Waits until no request is queued or running. */
int katoann_server_drain(katoann_server *s);
int katoann_server_stats_get(katoann_server *s, katoann_server_stats *out);

/* Row normalisation of the front ends (python/katoann_module.cpp:97-110) on the GPU. */
int katoann_normalize_rows(float *rows, uint64_t n, uint32_t dim);

/* The rotation of the load and search paths (src/index.cpp:139-270) on the GPU. */
int katoann_rotate_rows(const float *in, float *out, uint64_t n, uint32_t dim, const float *rotation);

#ifdef __cplusplus
}
#endif

#endif /* KATOANN_C_H_ */
