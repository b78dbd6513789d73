// Host side of the gfx950 KatoANN library: index files, the device-resident
// index, the search pipeline and the C ABI (include/katoann_c.h).
//
// The search runs as one stream of kernels per batch (kernels.hip):
//   normalise? -> rotate -> quantise -> coarse MFMA -> probe select ->
//   inversion + plan -> ranking (persistent) -> survivors + refine + top-k
// Each concurrent search takes its own context (stream, workspace, events)
// from a pool, so Index::search stays const and thread-safe as in the
// reference (src/index.cpp:879-904, README.md:228).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <map>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/katoann_c.h"
#include "kernels.hpp"

namespace {

constexpr uint32_t kMaxDim = 4096;          // kernels.h:24
constexpr uint32_t kDimcMagic = 0x434D4944u;  // index.cpp:27
constexpr uint32_t kDimcVersion = 2u;       // index.cpp:28
constexpr uint32_t kMaxSurvivors = 4096;
constexpr uint32_t kPoolAllowance = 1024;  // pool slots per query beyond S for one band's kept keys
constexpr uint64_t kPoolAllowanceCap = 65536;  // ... at most, unless one list is larger (overflow reruns)
constexpr uint32_t kMaxProbes = 4096;
constexpr uint32_t kMaxBands = 16;  // probe-rank bands at most (nprobe 4096, factor 2: 13)

thread_local std::string g_last_error;

struct logic_err : std::logic_error {
  using std::logic_error::logic_error;
};

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}
#define HIPCHECK(x) hip_check((x), #x)

// Device allocation that grows on demand.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t n) {
    if (n <= bytes && p) return;
    release();
    const size_t want = std::max<size_t>(n, 256);
    HIPCHECK(hipMalloc(&p, want));
    bytes = want;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

template <class T>
void upload(DevBuf &b, const T *src, size_t count) {
  b.ensure(count * sizeof(T));
  if (count) HIPCHECK(hipMemcpy(b.p, src, count * sizeof(T), hipMemcpyHostToDevice));
}

struct SearchCtx {
  hipStream_t stream = nullptr;
  hipEvent_t ev[KATOANN_NSTAGES] = {};  // stage boundaries (the pool updates are timed as the rest of the ranking interval)
  hipEvent_t bev[2 * kMaxBands] = {};   // around each band's ranking kernel
  DevBuf ws;                            // the search's buffers, carved from one allocation
  hipGraphExec_t exec = nullptr;  // the last captured pass and its launch arguments
  std::vector<uint64_t> gkey;
  SearchCtx() {
    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (auto &e : ev) HIPCHECK(hipEventCreate(&e));
    for (auto &e : bev) HIPCHECK(hipEventCreate(&e));
  }
  ~SearchCtx() {
    if (exec) (void)hipGraphExecDestroy(exec);
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto &e : bev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

}  // namespace

// Tile shape of the ranking pass: kernel (0 block items, 1 wave items, 2 block
// items with a query ring), row-tile variant, queries per item, log2 band factor.
struct TileConfig {
  uint32_t kernel, variant, qc, shift;
  uint32_t seed;      // rows of the seed band (0: geometric bands from rank 0)
  uint32_t join = 0;  // the seed remainder ranked in the first probe-rank band (kernels.hpp)
};
using TileKey = std::array<uint32_t, 6>;  // P, nlist, batch rounded up to 2^k, nprobe, S, ranking

struct katoann_index {
  uint32_t dim = 0, nlist = 0, npts = 0, dc = 0, head_dims = 0;
  bool normalized = false, rotated = false, batched = true;
  int32_t precision = KATOANN_SQ16;
  float lo = 0.0f;
  double inv_step = 0.0, step = 0.0, step_sq = 0.0;
  std::vector<uint32_t> sizes;
  std::vector<uint64_t> list_off;
  uint64_t nrows = 0;
  uint32_t rank_rows = 1024;
  uint32_t band_shift = 5;     // probe-rank bands grow by 2^band_shift
  uint32_t seed_rows = 1024;  // rows of the seed band (0: none)
  uint32_t seed_join = 1;     // seed remainder joined to the first probe-rank band (where it applies)
  uint32_t pool_allowance = 0;  // pool slots per query beyond S (0: sized per search)
  uint32_t rank_kernel = 1, wave_variant = 0, wave_qc = 1024;
  std::atomic<bool> graphs{true};  // replay each search's kernels from a captured graph
  std::atomic<bool> autotune{false};
  std::mutex tune_mu;
  std::map<TileKey, TileConfig> tuned;  // per shape, the fastest tile measured
  // per shape and tile, the pool allowance the last overflow rerun needed
  std::map<std::array<uint64_t, 8>, uint64_t> allowance_seen;
  uint64_t gen = 0;                // bumped by every load: captured graphs name the old buffers
  std::vector<uint64_t> tile_prefix;  // sum of the p largest tile counts, p = 0..nlist
  uint32_t tile_prefix_rows = 0;

  DevBuf rot, list_off_d, ids, id_row, hi, lo_plane, pre, norm, hinorm, headnorm, cen_hi, cen_lo, cen_norm;
  DevBuf cen_f;  // float centroids, unrotated: the per-query coarse search (batched = false)
  // bf16 precision: unrotated rows as bf16, dp = roundup(dim, 32) apart, and their norms
  DevBuf bf, bfn;
  uint32_t dp = 0;

  std::mutex mu;
  std::vector<std::unique_ptr<SearchCtx>> free_ctx;
  std::atomic<bool> timing{false};
  // last search's stage times and work, per calling thread
  std::mutex stat_mu;
  float stage_ms[KATOANN_NSTAGES] = {};
  float band_ms[kMaxBands] = {};
  uint64_t work[4] = {};

  void reset() {
    ++gen;
    {
      std::lock_guard<std::mutex> g(tune_mu);
      tuned.clear();
      allowance_seen.clear();
    }
    dim = nlist = npts = dc = head_dims = 0;
    normalized = rotated = false;
    batched = true;
    precision = KATOANN_SQ16;
    lo = 0.0f;
    inv_step = step = step_sq = 0.0;
    sizes.clear();
    list_off.clear();
    nrows = 0;
    tile_prefix.clear();
    tile_prefix_rows = 0;
    for (DevBuf *b : {&rot, &list_off_d, &ids, &id_row, &hi, &lo_plane, &pre, &norm, &hinorm, &headnorm, &cen_hi,
                      &cen_lo, &cen_norm, &cen_f, &bf, &bfn})
      b->release();
    dp = 0;
  }

  std::unique_ptr<SearchCtx> acquire() const {
    auto *self = const_cast<katoann_index *>(this);
    std::lock_guard<std::mutex> g(self->mu);
    if (!self->free_ctx.empty()) {
      auto c = std::move(self->free_ctx.back());
      self->free_ctx.pop_back();
      return c;
    }
    return std::make_unique<SearchCtx>();
  }
  void release(std::unique_ptr<SearchCtx> c) const {
    auto *self = const_cast<katoann_index *>(this);
    std::lock_guard<std::mutex> g(self->mu);
    self->free_ctx.push_back(std::move(c));
  }

  // Upper bound of the candidates of any one query at nprobe np: the sizes of
  // the np largest lists.
  uint64_t max_rows_per_query(uint32_t np) {
    std::lock_guard<std::mutex> g(mu);
    if (tile_prefix.size() != nlist + 1u) {
      std::vector<uint64_t> t(sizes.begin(), sizes.end());
      std::sort(t.begin(), t.end(), std::greater<uint64_t>());
      tile_prefix.assign(nlist + 1, 0);
      for (uint32_t c = 0; c < nlist; ++c) tile_prefix[c + 1] = tile_prefix[c] + t[c];
    }
    return tile_prefix[std::min(np, nlist)];
  }
};

namespace {

uint32_t round_up(uint32_t x, uint32_t m) { return (x + m - 1) / m * m; }

int device_cus() {
  static int cus = [] {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
    return prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  }();
  return cus;
}

// ---------------------------------------------------------------- files

struct RawIndex {
  uint32_t dim = 0, nlist = 0, npts = 0;
  bool normalized = false;
  std::vector<float> centroids, vecs;
  std::vector<uint32_t> sizes, ids;
};

// Pinned staging of the .invlist vectors on their way to device memory: two
// buffers, the file read into one while the other's copy runs.
std::atomic<uint64_t> g_load_chunk{64ull << 20};

class Stager {
 public:
  explicit Stager(size_t bytes) : cap_(bytes) {
    HIPCHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
      HIPCHECK(hipHostMalloc(&buf_[i], cap_, hipHostMallocDefault));
      HIPCHECK(hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming));
      HIPCHECK(hipEventRecord(ev_[i], s_));
    }
  }
  ~Stager() {
    (void)hipStreamSynchronize(s_);
    for (int i = 0; i < 2; ++i) {
      (void)hipHostFree(buf_[i]);
      (void)hipEventDestroy(ev_[i]);
    }
    (void)hipStreamDestroy(s_);
  }
  size_t cap() const { return cap_; }
  // the other buffer, once its last copy has finished
  char *next() {
    cur_ ^= 1;
    HIPCHECK(hipEventSynchronize(ev_[cur_]));
    return static_cast<char *>(buf_[cur_]);
  }
  void copy(void *dst, size_t n) {
    HIPCHECK(hipMemcpyAsync(dst, buf_[cur_], n, hipMemcpyHostToDevice, s_));
    HIPCHECK(hipEventRecord(ev_[cur_], s_));
  }
  void finish() { HIPCHECK(hipStreamSynchronize(s_)); }

 private:
  size_t cap_;
  hipStream_t s_ = nullptr;
  void *buf_[2] = {nullptr, nullptr};
  hipEvent_t ev_[2] = {};
  int cur_ = 0;
};

// Reads .meta, .centroids, .invlist with the reference's checks and messages
// (src/index.cpp:550-591). With `vec_off`, the vectors are not read: their
// file offsets (one per list) are recorded after the same checks (the file
// must hold every list's ids and vectors), for stream_vectors.
void read_files(const std::string &prefix, RawIndex &r, std::vector<uint64_t> *vec_off = nullptr) {
  {
    std::ifstream f(prefix + ".meta", std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + prefix + ".meta");
    uint32_t meta[3] = {0, 0, 0};
    f.read(reinterpret_cast<char *>(meta), sizeof(meta));
    if (!f) throw std::runtime_error("truncated " + prefix + ".meta");
    r.dim = meta[0];
    r.nlist = meta[1];
    r.npts = meta[2];
    uint32_t extra[2] = {0, 0};
    f.read(reinterpret_cast<char *>(extra), sizeof(extra));
    if (f.gcount() == static_cast<std::streamsize>(sizeof(extra))) {
      if (extra[0] != 0) throw std::runtime_error("only the L2 metric is supported");
      r.normalized = extra[1] != 0;
    }
  }
  if (r.dim == 0 || r.dim > kMaxDim || r.nlist == 0)
    throw std::runtime_error("bad header in " + prefix + ".meta");
  r.centroids.assign(static_cast<size_t>(r.nlist) * r.dim, 0.0f);
  {
    std::ifstream f(prefix + ".centroids", std::ios::binary);
    f.read(reinterpret_cast<char *>(r.centroids.data()),
           static_cast<std::streamsize>(r.centroids.size() * sizeof(float)));
    if (!f) throw std::runtime_error("cannot read " + prefix + ".centroids");
  }
  std::ifstream f(prefix + ".invlist", std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + prefix + ".invlist");
  r.sizes.assign(r.nlist, 0);
  uint64_t total = 0;
  r.ids.clear();
  r.vecs.clear();
  const uint64_t row_bytes = static_cast<uint64_t>(r.dim) * 4;
  uint64_t fsize = 0;
  if (vec_off) {
    vec_off->assign(r.nlist, 0);
    f.seekg(0, std::ios::end);
    fsize = static_cast<uint64_t>(f.tellg());
    f.seekg(0, std::ios::beg);
  }
  for (uint32_t c = 0; c < r.nlist; ++c) {
    uint32_t sz = 0;
    f.read(reinterpret_cast<char *>(&sz), sizeof(sz));
    if (!f) break;
    r.sizes[c] = sz;
    const size_t i0 = r.ids.size(), v0 = r.vecs.size();
    r.ids.resize(i0 + sz);
    f.read(reinterpret_cast<char *>(r.ids.data() + i0), static_cast<std::streamsize>(sz) * 4);
    if (vec_off) {
      if (!f) break;
      const uint64_t at = static_cast<uint64_t>(f.tellg());
      (*vec_off)[c] = at;
      if (at + sz * row_bytes > fsize) {
        f.setstate(std::ios::failbit);
        break;
      }
      f.seekg(static_cast<std::streamoff>(at + sz * row_bytes), std::ios::beg);
    } else {
      r.vecs.resize(v0 + static_cast<size_t>(sz) * r.dim);
      f.read(reinterpret_cast<char *>(r.vecs.data() + v0), static_cast<std::streamsize>(sz) * r.dim * 4);
    }
    total += sz;
    if (!f) break;
  }
  if (!f || total != r.npts) throw std::runtime_error("truncated " + prefix + ".invlist");
}

// The vectors of a checked .invlist (offsets from read_files) into device
// memory, npts x dim floats in list order, through the pinned stager: the host
// never holds the whole file.
void stream_vectors(const std::string &prefix, const RawIndex &r, const std::vector<uint64_t> &vec_off, DevBuf &dev) {
  const uint64_t row_bytes = static_cast<uint64_t>(r.dim) * 4;
  dev.ensure(std::max<uint64_t>(r.npts, 1) * row_bytes);
  std::ifstream f(prefix + ".invlist", std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + prefix + ".invlist");
  Stager st(static_cast<size_t>(std::max<uint64_t>(row_bytes, g_load_chunk.load())));
  char *dst = dev.as<char>();
  for (uint32_t c = 0; c < r.nlist; ++c) {
    f.seekg(static_cast<std::streamoff>(vec_off[c]), std::ios::beg);
    for (uint64_t left = r.sizes[c] * row_bytes; left > 0;) {
      const size_t n = static_cast<size_t>(std::min<uint64_t>(left, st.cap()));
      char *b = st.next();
      f.read(b, static_cast<std::streamsize>(n));
      if (!f) throw std::runtime_error("truncated " + prefix + ".invlist");
      st.copy(dst, n);
      dst += n;
      left -= n;
    }
  }
  st.finish();
}

// .dimc header and rotation (src/index.cpp:593-611); false when absent.
bool read_rotation(const std::string &prefix, const RawIndex &r, uint32_t *stored, bool *rotated,
                   std::vector<float> &rotation) {
  std::ifstream f(prefix + ".dimc", std::ios::binary);
  if (!f) return false;
  uint32_t hdr[8] = {0};
  f.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
  if (!f || hdr[0] != kDimcMagic || (hdr[1] != 1u && hdr[1] != kDimcVersion))
    throw std::runtime_error("bad header in " + prefix + ".dimc");
  if (hdr[2] != r.dim || hdr[3] != r.nlist || hdr[4] != r.npts)
    throw std::runtime_error(prefix + ".dimc does not match the index");
  *stored = hdr[5];
  *rotated = hdr[6] == 1u;
  if (*rotated) {
    rotation.resize(static_cast<size_t>(r.dim) * r.dim);
    f.read(reinterpret_cast<char *>(rotation.data()), static_cast<std::streamsize>(rotation.size() * 4));
    if (!f) throw std::runtime_error("truncated rotation in " + prefix + ".dimc");
  }
  return true;
}

// ---------------------------------------------------------------- load

void check_load_options(const katoann_load_options &o) {
  if (o.precision != KATOANN_SQ16 && o.precision != KATOANN_BF16) throw std::invalid_argument("precision is 'sq16' or 'bf16'");
}

// Everything of Index::load after reading the files (src/index.cpp:777-807):
// rotation of every list, the global scale, the 16-bit codes and norms, the
// prefix plane, the centroid codes and the id map, all on the GPU.
void build_device_index(katoann_index &ix, uint32_t dim, uint32_t nlist, uint32_t npts, bool normalized,
                        const float *centroids, const uint32_t *sizes, const uint32_t *ids, const float *vecs,
                        bool vecs_on_device, const float *rotation, bool has_dimc, uint32_t dims, bool batched,
                        int32_t precision) {
  ix.reset();
  ix.batched = batched;
  ix.precision = precision;
  ix.dim = dim;
  ix.nlist = nlist;
  ix.npts = npts;
  ix.normalized = normalized;
  ix.dc = round_up(dim, 64u);
  ix.sizes.assign(sizes, sizes + nlist);
  ix.list_off.assign(nlist + 1, 0);
  for (uint32_t c = 0; c < nlist; ++c) ix.list_off[c + 1] = ix.list_off[c] + sizes[c];
  const uint64_t N = ix.list_off[nlist];
  ix.nrows = N;
  if (precision == KATOANN_BF16) {
    // src/index.cpp:777-786: the unrotated vectors as bf16 and their norms;
    // no codes, no rotation, no prefix plane, no id map; the coarse search is
    // the per-query one
    ix.batched = false;
    ix.dp = round_up(dim, 32u);
    hipStream_t s = nullptr;
    DevBuf raw;
    const float *dv = vecs;
    if (!vecs_on_device && N) {
      upload(raw, vecs, static_cast<size_t>(N) * dim);
      dv = raw.as<float>();
    }
    ix.bf.ensure(static_cast<size_t>(N) * ix.dp * 2);
    ix.bfn.ensure(N * 4);
    const uint64_t chunk = 1u << 22;
    for (uint64_t r0 = 0; r0 < N; r0 += chunk) {
      const uint32_t nr = static_cast<uint32_t>(std::min<uint64_t>(chunk, N - r0));
      katoann_hip::launch_to_bf16(dv + r0 * dim, nr, dim, ix.dp, ix.bf.as<uint16_t>() + r0 * ix.dp,
                                  ix.bfn.as<float>() + r0, s);
    }
    HIPCHECK(hipGetLastError());
    upload(ix.cen_f, centroids, static_cast<size_t>(nlist) * dim);
    upload(ix.ids, ids, N);
    upload(ix.list_off_d, ix.list_off.data(), ix.list_off.size());
    (void)ix.max_rows_per_query(1);
    HIPCHECK(hipDeviceSynchronize());
    return;
  }
  if (has_dimc && dims) {
    if (dims % 64 || dims > ix.dc)
      throw std::invalid_argument("prefix_dims must be a multiple of 64 within the dimension");
    ix.head_dims = dims;
  }
  ix.rotated = rotation != nullptr;
  hipStream_t s = nullptr;
  const uint32_t d = dim, dc = ix.dc;

  DevBuf raw, y;
  const float *dv = vecs;
  if (!vecs_on_device && N) {
    upload(raw, vecs, static_cast<size_t>(N) * d);
    dv = raw.as<float>();
  }
  const float *yv = dv;
  if (ix.rotated) {
    upload(ix.rot, rotation, static_cast<size_t>(d) * d);
    if (N) {
      y.ensure(static_cast<size_t>(N) * d * 4);
      // rows in chunks so each launch stays within 32-bit grid sizes
      const uint64_t chunk = 1u << 22;
      for (uint64_t r0 = 0; r0 < N; r0 += chunk) {
        const uint32_t nr = static_cast<uint32_t>(std::min<uint64_t>(chunk, N - r0));
        katoann_hip::launch_rotate(dv + r0 * d, nr, d, ix.rot.as<float>(), y.as<float>() + r0 * d, s);
      }
      HIPCHECK(hipGetLastError());
      yv = y.as<float>();
    }
    raw.release();
  }
  // global scale (src/index.cpp:634-653)
  float glo = std::numeric_limits<float>::max(), ghi = -std::numeric_limits<float>::max();
  if (N) {
    const uint32_t nb = 2048;
    DevBuf part;
    part.ensure(nb * 2 * sizeof(float));
    katoann_hip::launch_minmax(yv, N * d, part.as<float>(), nb, s);
    std::vector<float> hp(nb * 2);
    HIPCHECK(hipMemcpy(hp.data(), part.p, hp.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t b = 0; b < nb; ++b) {
      glo = std::min(glo, hp[2 * b]);
      ghi = std::max(ghi, hp[2 * b + 1]);
    }
  }
  ix.lo = glo;
  const double rng = static_cast<double>(ghi) - glo;
  ix.inv_step = 65535.0 / rng;
  ix.step = rng / 65535.0;
  ix.step_sq = ix.step * ix.step;

  const uint32_t P = ix.head_dims;
  ix.hi.ensure(static_cast<size_t>(N) * dc);
  ix.lo_plane.ensure(static_cast<size_t>(N) * dc);
  ix.norm.ensure(N * 8);
  ix.hinorm.ensure(N * 4);
  ix.headnorm.ensure(N * 4);
  if (P) ix.pre.ensure(static_cast<size_t>(N) * P);
  if (N) {
    const uint64_t chunk = 1u << 22;
    for (uint64_t r0 = 0; r0 < N; r0 += chunk) {
      const uint32_t nr = static_cast<uint32_t>(std::min<uint64_t>(chunk, N - r0));
      katoann_hip::launch_quantize(yv + r0 * d, nr, d, dc, P, ix.lo, ix.inv_step, ix.hi.as<int8_t>() + r0 * dc,
                                   ix.lo_plane.as<int8_t>() + r0 * dc, ix.norm.as<int64_t>() + r0,
                                   ix.hinorm.as<int32_t>() + r0, ix.headnorm.as<int32_t>() + r0,
                                   P ? ix.pre.as<int8_t>() + r0 * P : nullptr, nullptr, s);
    }
    HIPCHECK(hipGetLastError());
  }
  y.release();
  raw.release();

  // centroids with the database scale (src/index.cpp:718-750)
  {
    DevBuf c, cy, tmp32;
    upload(c, centroids, static_cast<size_t>(nlist) * d);
    const float *cv = c.as<float>();
    if (ix.rotated) {
      cy.ensure(static_cast<size_t>(nlist) * d * 4);
      katoann_hip::launch_rotate(cv, nlist, d, ix.rot.as<float>(), cy.as<float>(), s);
      cv = cy.as<float>();
    }
    ix.cen_hi.ensure(static_cast<size_t>(nlist) * dc);
    ix.cen_lo.ensure(static_cast<size_t>(nlist) * dc);
    ix.cen_norm.ensure(static_cast<size_t>(nlist) * 8);
    katoann_hip::launch_quantize(cv, nlist, d, dc, 0, ix.lo, ix.inv_step, ix.cen_hi.as<int8_t>(),
                                 ix.cen_lo.as<int8_t>(), ix.cen_norm.as<int64_t>(), nullptr, nullptr, nullptr,
                                 nullptr, s);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipDeviceSynchronize());
  }
  // the per-query coarse search compares unrotated queries with the float
  // centroids (src/index.cpp:832-844)
  if (!batched) upload(ix.cen_f, centroids, static_cast<size_t>(nlist) * d);
  // ids and id -> row (src/index.cpp:795-806; a later duplicate wins)
  upload(ix.ids, ids, N);
  {
    std::vector<uint32_t> idr(npts, UINT32_MAX);
    for (uint64_t r = 0; r < N; ++r)
      if (ids[r] < npts) idr[ids[r]] = static_cast<uint32_t>(r);
    upload(ix.id_row, idr.data(), idr.size());
  }
  upload(ix.list_off_d, ix.list_off.data(), ix.list_off.size());
  (void)ix.max_rows_per_query(1);  // the largest lists' sums, ahead of any search
  HIPCHECK(hipDeviceSynchronize());
}

// ---------------------------------------------------------------- search

// Validated shape of one search and the bytes of every per-search buffer, so
// that the search and the admission estimate of a serving loop agree.
struct SearchPlan {
  uint32_t B, d, dc, np, K, S, nlist, P, shift, seed, join, qc, nslots, nbands, nstate, rt, item_cap, variant;
  bool prefix, full, bf16, wave, ring, qs, hyb, gemm;
  uint32_t qsq, qsrt;  // hybrid items: slots of at most qsq queries as query-stationary items of qsrt rows
  uint64_t max_rows, allowance;
  size_t q, qcodes, qpre, qnorm, q4, dist, probes, qlist, gcnt, pscratch, item_off, pool_cnt, thr, state, items, out, surv,
      codes, npq, mx16, mh, qbf, nsurv;
  std::array<uint64_t, 8> akey;  // shape and tile: the key of the pool room an overflow rerun needed
};

// Device buffers of one search, carved in this order from one allocation
// (every buffer 256-byte aligned; the pool, whose room varies, last).
struct Workspace {
  float *q, *y;
  int8_t *qhi, *qlo, *qpre;
  uint16_t *qbf;
  float *qbn;
  int64_t *qnorm;
  int32_t *qhn, *qh0;
  float *dist;
  uint32_t *probes;
  float *probe_d;
  uint32_t *npq, *gcnt, *qoff, *pscratch, *item_off, *qlist, *totals, *pool_cnt, *pool_sorted;
  uint64_t *thr;
  int32_t *lbound;
  uint32_t *state;
  uint4 *items;
  uint32_t *ids;
  float *dists;
  uint64_t *surv;
  uint32_t *nsurv;
  uint16_t *codes;
  void *mx16, *mh1, *mh2;
  float *mz;
  uint64_t *pool;
};

// Bytes of the workspace of plan z with `stride` pool slots per query; with
// base, also the buffers' addresses.
size_t carve(const SearchPlan &z, uint64_t stride, char *base, Workspace *w) {
  size_t off = 0;
  auto take = [&](auto &ptr, size_t bytes) {
    using Ptr = std::remove_reference_t<decltype(ptr)>;
    ptr = bytes && base ? reinterpret_cast<Ptr>(base + off) : nullptr;
    off += (bytes + 255) / 256 * 256;
  };
  Workspace tmp{};
  Workspace &v = w ? *w : tmp;
  const size_t B = z.B;
  take(v.q, z.q);
  take(v.y, z.q);
  take(v.qhi, z.qcodes);
  take(v.qlo, z.qcodes);
  take(v.qpre, z.qpre);
  take(v.qbf, z.qbf);
  take(v.qbn, z.qbf ? B * 4 : 0);
  take(v.qnorm, z.qnorm);
  take(v.qhn, z.q4);
  take(v.qh0, z.q4);
  take(v.dist, z.dist);
  take(v.probes, z.probes);
  take(v.probe_d, z.probes);
  take(v.npq, z.npq);
  take(v.gcnt, z.gcnt);
  take(v.qoff, z.gcnt);
  take(v.pscratch, z.pscratch);
  take(v.item_off, z.item_off);
  take(v.qlist, z.qlist);
  take(v.totals, 16);
  take(v.pool_cnt, z.pool_cnt);
  take(v.pool_sorted, z.q4);
  take(v.thr, z.thr);
  take(v.lbound, z.q4);
  take(v.state, z.state);
  take(v.items, z.items);
  take(v.ids, z.out);
  take(v.dists, z.out);
  take(v.surv, z.surv);
  take(v.nsurv, z.nsurv);
  take(v.codes, z.codes);
  take(v.mx16, z.mx16);
  take(v.mh1, z.mh);
  take(v.mh2, z.mh);
  take(v.mz, z.mx16 ? z.q4 : 0);
  take(v.pool, std::max<uint64_t>(stride * B, 1) * 8);
  return off;
}

// Pool slots per query of plan z with room `allowance` beyond S.
uint64_t pool_stride(const SearchPlan &z, uint64_t allowance) {
  return std::min<uint64_t>(z.S + allowance, 0xffffffffu);
}

TileKey tile_key(const katoann_index &ix, uint32_t nq, const katoann_search_params &p) {
  const bool prefix = p.ranking == KATOANN_RANK_PREFIX;
  uint32_t b = 1;
  while (b < nq) b <<= 1;
  return TileKey{prefix ? ix.head_dims : ix.dc, ix.nlist, b, p.nprobe, std::max(p.survivors, p.k),
                 static_cast<uint32_t>(p.ranking)};
}

// The tile a search runs with: the forced one, else the one measured fastest
// for its shape, else the index's settings.
TileConfig tile_for(const katoann_index &ix, uint32_t nq, const katoann_search_params &p, const TileConfig *force) {
  if (force) return *force;
  auto &mix = const_cast<katoann_index &>(ix);
  {
    std::lock_guard<std::mutex> g(mix.tune_mu);
    auto it = mix.tuned.find(tile_key(ix, nq, p));
    if (it != mix.tuned.end()) return it->second;
  }
  return TileConfig{ix.rank_kernel, ix.wave_variant, ix.wave_qc, ix.band_shift, ix.seed_rows, ix.seed_join};
}

SearchPlan plan_search(const katoann_index &ix, uint32_t nq, const katoann_search_params &p,
                       const katoann_search_extras *x, const TileConfig *force = nullptr) {
  if (ix.nlist == 0) throw logic_err("search before load");
  if (p.k == 0 || p.nprobe == 0 || p.nprobe > ix.nlist)
    throw std::invalid_argument("need k >= 1 and 1 <= nprobe <= nlist");
  if (ix.precision == KATOANN_BF16 && p.ranking != KATOANN_RANK_FULL)
    throw std::invalid_argument("bf16 precision supports only Ranking::kFull");
  if (p.ranking == KATOANN_RANK_PREFIX && ix.head_dims == 0)
    throw std::invalid_argument("prefix ranking needs the rotation file (.dimc)");
  if (p.ranking != KATOANN_RANK_PREFIX && p.ranking != KATOANN_RANK_HIGH && p.ranking != KATOANN_RANK_FULL)
    throw std::invalid_argument("ranking is 'prefix', 'high' or 'full'");
  if (p.selection != KATOANN_SEL_THRESHOLD && p.selection != KATOANN_SEL_LOCAL)
    throw std::invalid_argument("selection is 'threshold' or 'local'");
  SearchPlan z{};
  z.full = p.ranking == KATOANN_RANK_FULL;
  z.bf16 = ix.precision == KATOANN_BF16;
  if (z.bf16 && x && x->codes) throw std::invalid_argument("a bf16 index holds no 16-bit codes");
  // full ranking keeps the k nearest of all candidates (src/index.cpp:1130-1219)
  z.S = z.full ? p.k : std::max(p.survivors, p.k);
  if (z.S > kMaxSurvivors)
    throw std::invalid_argument("max(survivors, k) above 4096 is not supported by the gfx950 engine");
  if (p.nprobe > kMaxProbes) throw std::invalid_argument("nprobe above 4096 is not supported by the gfx950 engine");
  z.prefix = p.ranking == KATOANN_RANK_PREFIX;
  z.P = z.prefix ? ix.head_dims : ix.dc;
  if (!z.full && z.P / 64 > 16)
    throw std::invalid_argument("high-byte ranking above 1024 dimensions is not supported yet");
  z.B = nq, z.d = ix.dim, z.dc = ix.dc, z.np = p.nprobe, z.K = p.k, z.nlist = ix.nlist;
  const TileConfig tc = tile_for(ix, nq, p, force);
  z.shift = tc.shift;
  // a seed band below S rows could leave a query without a threshold
  z.seed = tc.seed ? std::max(tc.seed, z.S) : 0u;
  z.wave = !z.full && (tc.kernel == 1 || tc.kernel == 2 || tc.kernel == 4);
  z.ring = !z.full && tc.kernel == 2;
  z.qs = !z.full && tc.kernel == 3;
  z.hyb = !z.full && tc.kernel == 4;
  // the block-tiled kernel ranks widths up to 512 bytes; wider ones run kernel 1
  z.gemm = !z.full && tc.kernel == 5 && katoann_hip::rank_gemm_queries(z.P) != 0;
  if (!z.full && tc.kernel == 5 && !z.gemm) z.wave = true;
  z.variant = z.full || z.hyb || z.gemm ? 0u : tc.variant;
  z.qsq = z.hyb ? katoann_hip::rank_qs_queries(z.P, 0) : 0u;
  z.qsrt = z.hyb ? 256u : 0u;
  // the query-stationary kernel: its queries per item are fixed by the
  // variant, and the tile's queries_per_item field gives its rows per item
  // (and the block-tiled kernel's likewise, in stages of 64 rows)
  z.qc = z.qs                ? katoann_hip::rank_qs_queries(z.P, z.variant)
         : z.gemm           ? katoann_hip::rank_gemm_queries(z.P)
         : tc.kernel == 5 && !z.full ? 1024u
         : z.wave || z.full ? tc.qc
                            : katoann_hip::rank_queries_per_item(z.P);
  z.rt = z.bf16   ? katoann_hip::rank_bf16_rows()
         : z.full ? katoann_hip::rank_full_rows()
         : z.qs   ? std::max(16u, tc.qc / 16 * 16)
         : z.gemm ? std::max(64u, tc.qc / 64 * 64)
         : z.wave ? katoann_hip::rank_wave_rows(z.P, z.variant) * (z.ring ? 4 : 1)
                  : ix.rank_rows;
  if (z.qc == 0 || z.rt == 0) throw std::invalid_argument("unsupported ranking shape (prefix dims)");
  z.nslots = katoann_hip::plan_slots(z.nlist, z.np, z.shift, z.seed);
  z.nbands = katoann_hip::plan_bands(z.np, z.shift, z.seed);
  if (z.nbands > kMaxBands) throw std::invalid_argument("too many probe-rank bands");
  // the joined seed remainder needs the probe-list inversion and a first
  // probe-rank band (bands 0, 1, 2)
  z.join = tc.join && z.seed && katoann_hip::inv_from_probes(z.nlist, z.np) && z.nbands > katoann_hip::kJoinBand &&
                   !z.bf16 && !z.full
               ? 1u
               : 0u;
  z.nstate = 3 + z.nbands;
  z.max_rows = const_cast<katoann_index &>(ix).max_rows_per_query(z.np);
  // items <= sum over lists of row tiles x (ceil(B / qc) + bands): a list's query
  // groups over all bands hold at most B queries
  uint64_t tiles_total = 0;
  for (uint32_t cc = 0; cc < z.nlist; ++cc) tiles_total += (ix.sizes[cc] + z.rt - 1) / z.rt;
  // (a seed band splits each list's rows in two: one more tile per list)
  const uint64_t item_cap64 = tiles_total * ((z.B + z.qc - 1) / z.qc + z.nbands) +
                              (z.seed ? static_cast<uint64_t>(z.nlist) * ((z.B + z.qc - 1) / z.qc + 1) : 0);
  if (item_cap64 > 0x7fffffffull) throw std::invalid_argument("batch too large for the ranking plan");
  z.item_cap = static_cast<uint32_t>(std::max<uint64_t>(item_cap64, 1));
  // the first bands run without a threshold (or with one from few candidates):
  // room for every row of a query's 1 + F nearest lists
  // (with a seed band: at most 2 R0 keys before the first threshold)
  z.allowance =
      z.seed ? std::min<uint64_t>(z.max_rows, std::max<uint64_t>(kPoolAllowance, 2ull * z.seed))
             : std::min(z.max_rows,
                        std::max<uint64_t>({kPoolAllowance, 2 * const_cast<katoann_index &>(ix).max_rows_per_query(1),
                                            std::min(kPoolAllowanceCap,
                                                     const_cast<katoann_index &>(ix).max_rows_per_query(
                                                         std::min(z.np, 1u + (1u << z.shift))))}));
  if (ix.pool_allowance) z.allowance = std::min<uint64_t>(z.max_rows, ix.pool_allowance);
  z.akey = {z.B, z.np, z.S, z.P, z.shift | (uint64_t(z.seed) << 8) | (uint64_t(z.join) << 40), z.rt, z.qc,
            uint64_t(z.variant) | (uint64_t(z.wave) << 8) | (uint64_t(z.ring) << 9) | (uint64_t(z.full) << 10) |
                (uint64_t(z.bf16) << 11) | (uint64_t(z.qs) << 12) | (uint64_t(z.hyb) << 13) |
                (uint64_t(z.gemm) << 14)};
  {
    // later searches of a shape start from the room its last overflow rerun needed
    auto &mix = const_cast<katoann_index &>(ix);
    std::lock_guard<std::mutex> g(mix.tune_mu);
    auto it = mix.allowance_seen.find(z.akey);
    if (it != mix.allowance_seen.end()) z.allowance = std::max(z.allowance, it->second);
  }
  const size_t B = z.B;
  z.q = B * z.d * 4;
  z.qcodes = B * z.dc;
  // the queries' prefix bytes, compact (rows P bytes apart): the ranking
  // kernels gather them, and rows dc bytes apart would fall on a few L2
  // channels only
  z.qpre = z.prefix ? B * z.P : 0;
  z.qbf = z.bf16 ? B * ix.dp * 2 : 0;
  z.qnorm = B * 8;
  z.q4 = B * 4;
  z.dist = B * z.nlist * 4;
  z.probes = B * z.np * 4;
  z.gcnt = z.nslots * 4ull;
  z.pscratch = katoann_hip::plan_scratch_words(z.B, z.nlist, z.np, z.shift, z.seed) * 4;
  // query lists: every probed pair once, seed pairs with rows past R0 twice
  z.qlist = z.probes * (z.seed ? 2 : 1);
  z.item_off = (z.nslots + 1ull) * 4;
  z.pool_cnt = B * 4 * katoann_hip::kCntStride;
  z.thr = B * 8;
  // the state, then the ranking items' counts and cursors per (band, cost class)
  z.state = (z.nstate + katoann_hip::kItemHistWords) * 4ull;
  z.items = static_cast<size_t>(z.item_cap) * 16;
  z.out = B * z.K * 4;
  z.surv = (x && x->survivors) ? B * z.S * 8 : 0;
  z.nsurv = (x && x->n_survivors) ? B * 4 : 0;
  z.codes = (x && x->codes) ? B * z.dc * 2 : 0;
  const katoann_nprobe_model *model = x ? x->nprobe_model : nullptr;
  if (model) {
    if (x->nprobe_per_query) throw std::invalid_argument("nprobe_model and nprobe_per_query are exclusive");
    if (model->nfeat == 0 || model->nfeat > z.np || model->hidden == 0)
      throw std::invalid_argument("nprobe model needs 1 <= nfeat <= nprobe and hidden >= 1");
    z.mx16 = B * model->nfeat * 2;
    z.mh = B * model->hidden * 2;
  }
  z.npq = (model || (x && x->nprobe_per_query)) ? B * 4 : 0;
  return z;
}

// Device memory a search runs in instead of buffers of its own (the
// allocation-free entry point): base, bytes.
struct FixedScratch {
  char *base;
  uint64_t bytes;
};

void search_impl(const katoann_index &ix, const float *queries, uint32_t nq, const katoann_search_params &p,
                 const katoann_search_extras *x, uint32_t *out_ids, float *out_dists,
                 const TileConfig *force = nullptr, const FixedScratch *fixed = nullptr) {
  if (nq == 0) return;
  const SearchPlan z = plan_search(ix, nq, p, x, force);
  const uint32_t B = z.B, d = z.d, dc = z.dc, np = z.np, K = z.K, nlist = z.nlist, S = z.S, P = z.P;
  const bool prefix = z.prefix;
  const bool dev_io = x && x->device_io;
  if (fixed && !dev_io) throw std::invalid_argument("the allocation-free search takes device pointers (device_io)");
  auto &mix = const_cast<katoann_index &>(ix);
  // the allocation-free search runs on the caller's stream with no context
  // (a context owns a stream, events and buffers of its own)
  std::unique_ptr<SearchCtx> ctx = fixed ? nullptr : ix.acquire();
  struct Release {  // the context goes back to the pool on every exit
    const katoann_index &ix;
    std::unique_ptr<SearchCtx> &c;
    ~Release() {
      if (c) ix.release(std::move(c));
    }
  } release_ctx{ix, ctx};
  hipStream_t s = (dev_io && x->stream) ? static_cast<hipStream_t>(x->stream) : (ctx ? ctx->stream : nullptr);
  const bool timing = ctx && ix.timing.load();
  const katoann_nprobe_model *model = x ? x->nprobe_model : nullptr;
  const bool coarse_only = x && x->coarse_only;
  const bool caller_npq = x && x->nprobe_per_query;
  const uint32_t shift = z.shift, qc = z.qc, nslots = z.nslots, nbands = z.nbands, nstate = z.nstate, rt = z.rt;
  const uint64_t max_rows = z.max_rows;
  const uint32_t item_cap = z.item_cap;
  // ranking blocks: as many as are resident at once
  const bool full = z.full;
  const uint32_t grid = static_cast<uint32_t>(device_cus()) *
                        (z.bf16              ? katoann_hip::rank_bf16_blocks_per_cu()
                         : full              ? katoann_hip::rank_full_blocks_per_cu()
                         : z.qs || z.hyb     ? katoann_hip::rank_qs_blocks_per_cu(P, z.variant)
                         : z.gemm            ? katoann_hip::rank_gemm_blocks_per_cu(P)
                         : z.wave && !z.ring ? katoann_hip::rank_wave_blocks_per_cu(P, z.variant)
                                             : 2u);
  Workspace w{};
  const uint32_t *npq = nullptr;
  uint32_t *dids = nullptr;
  float *ddst = nullptr;
  const float *yq = nullptr;
  uint32_t st[6] = {0, 0, 0, 0, 0, 0};
  auto mark = [&](int i) {
    if (timing) HIPCHECK(hipEventRecord(ctx->ev[i], s));
  };

  // the kernels of one pass (normalisation through top-k), launched eagerly
  // or captured once into a graph and replayed while every launch argument
  // stays the same
  auto launches = [&](uint32_t stride) {
    mark(0);
    if (x && x->normalize) katoann_hip::launch_normalize(w.q, B, d, s);
    yq = w.q;
    if (ix.rotated) {
      katoann_hip::launch_rotate(w.q, B, d, ix.rot.as<float>(), w.y, s);
      yq = w.y;
    }
    if (z.bf16)  // src/index.cpp:1230-1237
      katoann_hip::launch_to_bf16(w.q, B, d, ix.dp, w.qbf, w.qbn, s);
    else
      katoann_hip::launch_quantize(yq, B, d, dc, ix.head_dims, ix.lo, ix.inv_step, w.qhi, w.qlo, w.qnorm, w.qhn,
                                   w.qh0, prefix ? w.qpre : nullptr, w.codes, s);
    mark(1);
    if (ix.batched)
      katoann_hip::launch_coarse(w.qhi, w.qlo, w.qnorm, B, ix.cen_hi.as<int8_t>(), ix.cen_lo.as<int8_t>(),
                                 ix.cen_norm.as<int64_t>(), nlist, dc, ix.step_sq, w.dist, device_cus(), s);
    else
      katoann_hip::launch_coarse_flat(w.q, B, d, ix.cen_f.as<float>(), nlist, w.dist, s);
    mark(2);
    katoann_hip::launch_select(w.dist, B, nlist, np, w.probes, w.probe_d, s);
    if (model) {  // per-query nprobe from the sorted coarse distances
      const uint32_t F = model->nfeat, H = model->hidden;
      katoann_hip::launch_features(w.probe_d, B, np, F, model->feat_mean, model->feat_inv_std, nullptr, w.mx16, s);
      katoann_hip::launch_gemm_bf16(B, H, F, w.mx16, F, 1, model->w1, F, 1, model->b1, nullptr, 0, 1, nullptr, w.mh1,
                                    H, s);
      katoann_hip::launch_gemm_bf16(B, H, H, w.mh1, H, 1, model->w2, H, 1, model->b2, nullptr, 0, 1, nullptr, w.mh2,
                                    H, s);
      katoann_hip::launch_gemm_bf16(B, 1, H, w.mh2, H, 1, model->w3, H, 1, model->b3, nullptr, 0, 0, w.mz, nullptr, 1,
                                    s);
      katoann_hip::launch_nprobe_of(w.mz, B, model->y_mean, model->y_std, model->margin,
                                    std::max(1u, std::min(model->nmin, np)), np, w.npq, s);
    }
    mark(3);
    if (coarse_only) {
      mark(4);
      mark(5);
      mark(6);
      return;
    }
    katoann_hip::launch_plan(w.dist, w.probes, w.probe_d, npq, caller_npq, B, np, ix.list_off_d.as<uint64_t>(), nlist,
                             qc, rt, z.qsq, z.qsrt, w.pscratch, w.gcnt, w.qoff, w.item_off, w.qlist, w.totals,
                             w.pool_cnt, w.pool_sorted, w.thr, w.lbound, w.state, nstate, shift, z.seed, z.join, s);
    katoann_hip::launch_items(w.item_off, nslots, nlist, ix.list_off_d.as<uint64_t>(), w.gcnt, w.qoff, qc, rt, z.qsq,
                              z.qsrt, item_cap, z.seed, z.join,
                              w.pscratch + katoann_hip::plan_join_gb_offset(B, nlist, np, shift, z.seed),
                              w.state + nstate, w.items, s);
    mark(4);
    const int8_t *plane = prefix ? ix.pre.as<int8_t>() : ix.hi.as<int8_t>();
    const int32_t *xnorm = prefix ? ix.headnorm.as<int32_t>() : ix.hinorm.as<int32_t>();
    const int32_t *qn = prefix ? w.qh0 : w.qhn;
    const int8_t *qrank = prefix ? w.qpre : w.qhi;
    const uint32_t qstride = prefix ? P : dc;
    for (uint32_t band = 0; band < nbands; ++band) {
      if (timing) HIPCHECK(hipEventRecord(ctx->bev[2 * band], s));
      if (z.join && band == 1) {  // empty: its rows are ranked in band 2, on the seed band's thresholds
        if (timing) HIPCHECK(hipEventRecord(ctx->bev[2 * band + 1], s));
        continue;
      }
      if (z.bf16)
        katoann_hip::launch_rank_bf16(band, ix.bf.as<uint16_t>(), ix.bfn.as<float>(), ix.ids.as<uint32_t>(), nlist,
                                      w.qbf, w.qbn, ix.dp, w.thr, w.qlist, w.item_off, w.items, stride, w.pool,
                                      w.pool_cnt, w.state, grid, s);
      else if (full)
        katoann_hip::launch_rank_full(band, ix.hi.as<int8_t>(), ix.lo_plane.as<int8_t>(), ix.norm.as<int64_t>(),
                                      ix.ids.as<uint32_t>(), nlist, w.qhi, w.qlo, dc, w.qnorm, w.thr, w.qlist,
                                      w.item_off, w.items, stride, ix.step_sq, w.pool, w.pool_cnt, w.state, grid, s);
      const bool launched =
          full   ? true
          : z.hyb ? katoann_hip::launch_rank_hyb(band, 0, plane, xnorm, ix.ids.as<uint32_t>(), nlist, P, qrank,
                                                 qstride, qn, w.lbound, w.qlist, w.item_off, w.items, stride,
                                                 ix.step_sq, w.pool, w.pool_cnt, w.state, grid, s)
          : z.gemm ? katoann_hip::launch_rank_gemm(band, plane, xnorm, ix.ids.as<uint32_t>(), nlist, P, qrank,
                                                   qstride, qn, w.lbound, w.qlist, w.item_off, w.items, stride,
                                                   ix.step_sq, w.pool, w.pool_cnt, w.state, grid, s)
          : z.qs ? katoann_hip::launch_rank_qs(band, z.variant, plane, xnorm, ix.ids.as<uint32_t>(), nlist, P, qrank,
                                               qstride, qn, w.lbound, w.qlist, w.item_off, w.items, stride,
                                               ix.step_sq, w.pool, w.pool_cnt, w.state, grid, s)
          : z.wave ? katoann_hip::launch_rank_wave(band, z.variant, z.ring, plane, xnorm, ix.ids.as<uint32_t>(), nlist,
                                                   P, qrank, qstride, qn, w.lbound, w.qlist, w.item_off, w.items,
                                                   stride, ix.step_sq, w.pool, w.pool_cnt, w.state, grid, s)
                   : katoann_hip::launch_rank_band(band, plane, xnorm, ix.ids.as<uint32_t>(), nlist, P, qrank, qstride,
                                                   qn, w.lbound, w.qlist, w.item_off, w.items, stride, ix.step_sq,
                                                   w.pool, w.pool_cnt, w.state, grid, s);
      if (!launched) throw std::invalid_argument("unsupported ranking shape (prefix dims)");
      if (timing) HIPCHECK(hipEventRecord(ctx->bev[2 * band + 1], s));
      // after the last band, k_refine (S <= 1,024) takes the S smallest keys
      // of the whole pool itself: no update of its own
      if (band + 1 == nbands && !z.bf16 && katoann_hip::refine_cap(S) <= 1024) break;
      if (!katoann_hip::launch_update(w.pool, w.pool_cnt, w.pool_sorted, w.thr, w.lbound, B, S, stride, P, ix.step_sq,
                                      z.bf16, s))
        throw std::invalid_argument("unsupported survivor count");
    }
    mark(5);
    if (z.bf16)
      katoann_hip::launch_emit_bf16(w.pool, stride, w.pool_cnt, B, K, dids, ddst, w.surv, w.nsurv, s);
    else if (!katoann_hip::launch_refine(w.pool, stride, w.pool_cnt, w.thr, B, S, K, ix.hi.as<int8_t>(),
                                         ix.lo_plane.as<int8_t>(), ix.norm.as<int64_t>(), ix.id_row.as<uint32_t>(),
                                         ix.npts, dc, w.qhi, w.qlo, w.qnorm, ix.step_sq, dids, ddst, w.surv, w.nsurv,
                                         s))
      throw std::invalid_argument("unsupported refinement shape");
    mark(6);
  };
  const bool graphs = ctx && mix.graphs.load() && !timing;  // stage timing runs eagerly, events on the stream
  size_t ws_bytes = 0;
  auto run = [&](uint64_t allowance) {
    const uint32_t stride = static_cast<uint32_t>(pool_stride(z, allowance));
    ws_bytes = carve(z, stride, nullptr, nullptr);
    char *base = nullptr;
    if (fixed) {
      if (ws_bytes > fixed->bytes) throw std::logic_error("unreachable: workspace above the fixed scratch");
      base = fixed->base;
    } else {
      ctx->ws.ensure(ws_bytes);
      base = ctx->ws.as<char>();
    }
    carve(z, stride, base, &w);
    npq = (model || caller_npq) ? w.npq : nullptr;
    dids = dev_io ? out_ids : w.ids;
    ddst = dev_io ? out_dists : w.dists;
    if (caller_npq)
      HIPCHECK(hipMemcpyAsync(w.npq, x->nprobe_per_query, B * 4ull,
                              dev_io ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(w.q, queries, static_cast<size_t>(B) * d * 4,
                            dev_io ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    if (graphs) {
      SearchCtx &c = *ctx;
      // everything a launch reads: shapes, scalars, the workspace, the stream
      std::vector<uint64_t> key = {reinterpret_cast<uint64_t>(s), stride, B, d, dc, np, K, S, P, nlist, shift, z.seed, qc,
                                   rt, nbands, z.item_cap, z.variant, z.wave, z.ring, z.qs, z.hyb, z.gemm, prefix, full, z.bf16, grid, ix.gen,
                                   ix.rotated, x && x->normalize, coarse_only, caller_npq, reinterpret_cast<uint64_t>(npq),
                                   reinterpret_cast<uint64_t>(dids), reinterpret_cast<uint64_t>(ddst),
                                   reinterpret_cast<uint64_t>(base), ws_bytes, z.surv, z.nsurv, z.codes, z.qpre,
                                   z.qbf, z.npq, z.mx16, z.mh};
      if (model) {
        uint64_t words[(sizeof(katoann_nprobe_model) + 7) / 8] = {};
        std::memcpy(words, model, sizeof(katoann_nprobe_model));
        key.insert(key.end(), std::begin(words), std::end(words));
      }
      if (!c.exec || c.gkey != key) {
        if (c.exec) (void)hipGraphExecDestroy(c.exec);
        c.exec = nullptr;
        hipGraph_t g = nullptr;
        HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        try {
          launches(stride);
        } catch (...) {
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        HIPCHECK(hipStreamEndCapture(s, &g));
        const hipError_t ie = hipGraphInstantiate(&c.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHECK(ie);
        c.gkey = std::move(key);
      }
      HIPCHECK(hipGraphLaunch(c.exec, s));
    } else {
      launches(stride);
    }
    HIPCHECK(hipGetLastError());
    if (coarse_only) {
      HIPCHECK(hipStreamSynchronize(s));
      return true;
    }
    HIPCHECK(hipMemcpyAsync(st, w.state, 12, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(st + 4, w.totals, 8, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (st[2]) throw std::invalid_argument("per-query nprobe must be in [1, nprobe]");
    return st[0] == 0;
  };
  // Pool room: the plan's; within fixed scratch, at most what fits there.
  uint64_t allowance = z.allowance;
  uint64_t room_cap = max_rows;
  if (fixed) {
    if (carve(z, pool_stride(z, 0), nullptr, nullptr) > fixed->bytes)
      throw std::invalid_argument("search needs " + std::to_string(carve(z, pool_stride(z, 0), nullptr, nullptr)) +
                                  " bytes of scratch, above the allocation-free search's " +
                                  std::to_string(fixed->bytes));
    const size_t rest = carve(z, 0, nullptr, nullptr) - 256;  // everything but the pool (its one 256-byte slot)
    const uint64_t fit = (fixed->bytes - rest) / 8 / B;
    room_cap = std::min<uint64_t>(room_cap, fit > S ? fit - S : 0);
    allowance = std::min(allowance, room_cap);
  }
  const uint64_t limit = (x && x->max_bytes) ? x->max_bytes : 0;
  // on a pool overflow the batch runs again with 4x the room, up to every
  // candidate; later searches of the shape start from the room that sufficed
  while (!run(allowance)) {
    if (allowance >= room_cap) {
      if (fixed && room_cap < max_rows)
        throw std::runtime_error("survivor pool overflow in the ranking pass: the batch needs more room than the "
                                 "allocation-free search's scratch holds");
      throw std::runtime_error("survivor pool overflow in the ranking pass");
    }
    const uint64_t next = std::min<uint64_t>(room_cap, allowance * 4);
    const size_t need = carve(z, pool_stride(z, next), nullptr, nullptr);
    if (!fixed) {
      std::lock_guard<std::mutex> g(mix.tune_mu);
      mix.allowance_seen[z.akey] = std::max(mix.allowance_seen[z.akey], next);
    }
    if (limit && need > limit) {
      if (x->needed_bytes) *x->needed_bytes = need;
      throw std::runtime_error("survivor pool overflow: the rerun needs " + std::to_string(need) +
                               " bytes of search buffers, above the " + std::to_string(limit) + " allowed");
    }
    allowance = next;
  }
  if (x && x->needed_bytes) *x->needed_bytes = ws_bytes;

  const hipMemcpyKind d2o = dev_io ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (!dev_io && !coarse_only) {
    HIPCHECK(hipMemcpyAsync(out_ids, dids, static_cast<size_t>(B) * K * 4, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(out_dists, ddst, static_cast<size_t>(B) * K * 4, hipMemcpyDeviceToHost, s));
  }
  if (x) {
    if (x->probes) HIPCHECK(hipMemcpyAsync(x->probes, w.probes, static_cast<size_t>(B) * np * 4, d2o, s));
    if (x->probe_dists) HIPCHECK(hipMemcpyAsync(x->probe_dists, w.probe_d, static_cast<size_t>(B) * np * 4, d2o, s));
    if (x->survivors) HIPCHECK(hipMemcpyAsync(x->survivors, w.surv, static_cast<size_t>(B) * S * 8, d2o, s));
    if (x->n_survivors) HIPCHECK(hipMemcpyAsync(x->n_survivors, w.nsurv, B * 4ull, d2o, s));
    if (x->rotated) HIPCHECK(hipMemcpyAsync(x->rotated, yq, static_cast<size_t>(B) * d * 4, d2o, s));
    if (x->codes) HIPCHECK(hipMemcpyAsync(x->codes, w.codes, static_cast<size_t>(B) * dc * 2, d2o, s));
    if (x->nprobe_used) {
      if (npq) {
        HIPCHECK(hipMemcpyAsync(x->nprobe_used, npq, B * 4ull, d2o, s));
      } else if (dev_io) {
        katoann_hip::launch_fill32(x->nprobe_used, B, np, s);
        HIPCHECK(hipGetLastError());
      } else {
        HIPCHECK(hipStreamSynchronize(s));
        std::fill(x->nprobe_used, x->nprobe_used + B, np);
      }
    }
  }
  HIPCHECK(hipStreamSynchronize(s));
  if (x && npq && !dev_io && (x->probes || x->probe_dists)) {  // past a query's own count: no list
    std::vector<uint32_t> hn(B);
    if (caller_npq) std::copy(x->nprobe_per_query, x->nprobe_per_query + B, hn.begin());
    else HIPCHECK(hipMemcpy(hn.data(), npq, B * 4ull, hipMemcpyDeviceToHost));
    for (uint32_t q = 0; q < B; ++q)
      for (uint32_t j = hn[q]; j < np; ++j) {
        if (x->probes) x->probes[static_cast<size_t>(q) * np + j] = UINT32_MAX;
        if (x->probe_dists) x->probe_dists[static_cast<size_t>(q) * np + j] = std::numeric_limits<float>::max();
      }
  }
  {
    std::lock_guard<std::mutex> g(mix.stat_mu);
    if (timing) {
      float ms[6] = {};
      for (int i = 0; i < 6; ++i) (void)hipEventElapsedTime(&ms[i], ctx->ev[i], ctx->ev[i + 1]);
      float rank = 0.0f;
      for (uint32_t b = 0; b < kMaxBands; ++b) {
        float t = 0.0f;
        if (b < nbands && !(z.join && b == 1)) (void)hipEventElapsedTime(&t, ctx->bev[2 * b], ctx->bev[2 * b + 1]);
        mix.band_ms[b] = t;
        rank += t;
      }
      const float out[KATOANN_NSTAGES] = {ms[0], ms[1], ms[2], ms[3], rank, std::max(0.0f, ms[4] - rank), ms[5]};
      for (int i = 0; i < KATOANN_NSTAGES; ++i) mix.stage_ms[i] = out[i];
    }
    mix.work[0] = st[4];  // (query, list) pairs
    mix.work[1] = st[5];  // ranking items
    mix.work[2] = st[1];  // candidates kept
    mix.work[3] = nbands - z.join;  // ranking launches (a joined remainder's band has none)
  }
}

// Row sums of shifted bytes for the kernel-level API's exact unsigned products.
void kernel_pairs(const uint8_t *qh, const uint8_t *ql, const uint8_t *xh, const uint8_t *xl, const int64_t *qn,
                  const int64_t *xn, double step_sq, uint32_t m, uint32_t n, uint32_t dc, uint32_t qstride,
                  uint32_t xstride, bool full, float *out);

}  // namespace

// ============================================================================ C ABI

// The calling thread's error message, shared with build.cpp.
std::string &katoann_error_slot() { return g_last_error; }

#define KATOANN_TRY try {
#define KATOANN_CATCH                                  \
  return KATOANN_OK;                                   \
  }                                                    \
  catch (const std::invalid_argument &e) {             \
    g_last_error = e.what();                           \
    return KATOANN_INVALID_ARGUMENT;                   \
  }                                                    \
  catch (const logic_err &e) {                         \
    g_last_error = e.what();                           \
    return KATOANN_LOGIC_ERROR;                        \
  }                                                    \
  catch (const std::logic_error &e) {                  \
    g_last_error = e.what();                           \
    return KATOANN_LOGIC_ERROR;                        \
  }                                                    \
  catch (const std::exception &e) {                    \
    g_last_error = e.what();                           \
    return KATOANN_RUNTIME_ERROR;                      \
  }

extern "C" {

const char *katoann_last_error(void) { return g_last_error.c_str(); }

void katoann_default_build_options(katoann_build_options *o) {
  *o = katoann_build_options{2048u, 256000u, 12u, 0ull, 0, 100000u, 8u};
}
void katoann_default_load_options(katoann_load_options *o) { *o = katoann_load_options{KATOANN_SQ16, 1, 0u, 1}; }
void katoann_default_search_params(katoann_search_params *p) {
  *p = katoann_search_params{10u, 64u, KATOANN_RANK_PREFIX, 32u, KATOANN_SEL_THRESHOLD};
}

int katoann_set_load_chunk(uint64_t bytes) {
  KATOANN_TRY
  if (bytes < 4096 || bytes > (1ull << 32)) throw std::invalid_argument("load staging chunk must be 4 KiB to 4 GiB");
  g_load_chunk.store(bytes);
  KATOANN_CATCH
}

int katoann_index_create(katoann_index **out) {
  KATOANN_TRY
  *out = new katoann_index();
  KATOANN_CATCH
}

void katoann_index_destroy(katoann_index *h) { delete h; }

int katoann_load(katoann_index *h, const char *prefix_c, const katoann_load_options *opt) {
  KATOANN_TRY
  katoann_load_options o;
  katoann_default_load_options(&o);
  if (opt) o = *opt;
  h->reset();  // Index::load starts from an empty index (src/index.cpp:778)
  check_load_options(o);
  const std::string prefix(prefix_c);
  RawIndex r;
  std::vector<uint64_t> vec_off;
  read_files(prefix, r, &vec_off);
  uint32_t stored = 0;
  bool rotated = false;
  std::vector<float> rotation;
  // bf16 precision reads no .dimc (src/index.cpp:780-786)
  const bool has_dimc = o.precision != KATOANN_BF16 && read_rotation(prefix, r, &stored, &rotated, rotation);
  const uint32_t dims = o.prefix_dims ? o.prefix_dims : stored;
  if (has_dimc && dims && (dims % 64 || dims > round_up(r.dim, 64u)))
    throw std::invalid_argument("prefix_dims must be a multiple of 64 within the dimension");
  DevBuf dvecs;  // every file check has passed: the vectors go to device memory
  stream_vectors(prefix, r, vec_off, dvecs);
  build_device_index(*h, r.dim, r.nlist, r.npts, r.normalized, r.centroids.data(), r.sizes.data(), r.ids.data(),
                     dvecs.as<float>(), true, rotated ? rotation.data() : nullptr, has_dimc, dims,
                     o.batched_coarse != 0, o.precision);
  KATOANN_CATCH
}

int katoann_load_arrays(katoann_index *h, uint32_t dim, uint32_t nlist, uint32_t n, int32_t normalized,
                        const float *centroids, const uint32_t *list_sizes, const uint32_t *ids,
                        const float *vectors, int32_t vectors_on_device, const float *rotation,
                        uint32_t prefix_dims, const katoann_load_options *opt) {
  KATOANN_TRY
  katoann_load_options o;
  katoann_default_load_options(&o);
  if (opt) o = *opt;
  h->reset();
  check_load_options(o);
  if (dim == 0 || dim > kMaxDim || nlist == 0) throw std::runtime_error("bad header in arrays");
  uint64_t total = 0;
  for (uint32_t c = 0; c < nlist; ++c) total += list_sizes[c];
  if (total != n) throw std::runtime_error("list sizes do not add up to n");
  const uint32_t dims = o.prefix_dims ? o.prefix_dims : prefix_dims;
  build_device_index(*h, dim, nlist, n, normalized != 0, centroids, list_sizes, ids, vectors,
                     vectors_on_device != 0, rotation, true, dims, o.batched_coarse != 0, o.precision);
  KATOANN_CATCH
}

namespace {

// First search of a shape with autotuning on: every candidate tile runs three
// times (the first captures its graph), the better of the other two counts;
// the fastest is kept
// for the shape. Every tile gives the same results, so the last run's outputs
// stand.
void maybe_tune(const katoann_index &ix, const float *queries, uint32_t nq, const katoann_search_params &p,
                const katoann_search_extras *x, uint32_t *ids, float *dists) {
  auto &mix = const_cast<katoann_index &>(ix);
  if (!mix.autotune.load() || nq == 0 || ix.nlist == 0 || p.ranking == KATOANN_RANK_FULL) return;
  const TileKey key = tile_key(ix, nq, p);
  {
    std::lock_guard<std::mutex> g(mix.tune_mu);
    if (mix.tuned.count(key)) return;
  }
  double best = 1e30;
  TileConfig pick{1u, 0u, 1024u, 3u, 0u};
  auto trial = [&](const TileConfig &c) {
    if (c.kernel == 3   ? katoann_hip::rank_qs_queries(key[0], c.variant) == 0
        : c.kernel == 5 ? katoann_hip::rank_gemm_queries(key[0]) == 0
                        : katoann_hip::rank_wave_rows(key[0], c.variant) == 0)
      return;
    search_impl(ix, queries, nq, p, x, ids, dists, &c);
    double dt = 1e30;
    for (int rep = 0; rep < 4; ++rep) {  // the best of four timed runs (single runs vary by a few per cent)
      const auto t0 = std::chrono::steady_clock::now();
      search_impl(ix, queries, nq, p, x, ids, dists, &c);
      dt = std::min(dt, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    if (dt < best) {
      best = dt;
      pick = c;
    }
  };
  // the bands first (seed rows, and band factors of distinct band counts),
  // then the tile
  const uint32_t s0 = std::max(64u, std::max(p.survivors, p.k));
  for (uint32_t seed : {0u, 2 * s0, 4 * s0, 8 * s0, 16 * s0, 32 * s0}) {
    uint32_t last_nb = 0;
    for (uint32_t shift : {1u, 2u, 3u, 4u, 5u, 6u, 8u, 10u, 12u}) {
      const uint32_t nb = katoann_hip::plan_bands(p.nprobe, shift, seed);
      if (nb == last_nb || nb > kMaxBands) continue;
      last_nb = nb;
      trial(TileConfig{1u, 0u, 1024u, shift, seed});
      if (seed && nb > katoann_hip::kJoinBand) trial(TileConfig{1u, 0u, 1024u, shift, seed, 1u});
    }
  }
  const uint32_t shift = pick.shift, seed = pick.seed, join = pick.join;
  for (uint32_t variant : {0u, 1u, 2u, 3u})
    for (uint32_t qc : {256u, 512u, 1024u, 4096u}) trial(TileConfig{1u, variant, qc, shift, seed, join});
  for (uint32_t variant : {0u, 1u}) trial(TileConfig{2u, variant, 512u, shift, seed, join});
  // the query-stationary kernel: rows per item
  for (uint32_t variant : {0u, 1u})
    for (uint32_t rows : {256u, 1024u, 4096u}) trial(TileConfig{3u, variant, rows, shift, seed, join});
  // both forms per list (query-stationary for lists with few queries)
  for (uint32_t qc : {256u, 512u, 1024u, 4096u}) trial(TileConfig{4u, 0u, qc, shift, seed, join});
  // the block-tiled kernel: rows per item
  for (uint32_t rows : {512u, 2048u, 8192u}) trial(TileConfig{5u, 0u, rows, shift, seed, join});
  std::lock_guard<std::mutex> g(mix.tune_mu);
  mix.tuned[key] = pick;
}

}  // namespace

int katoann_search(const katoann_index *h, const float *queries, uint32_t nq, const katoann_search_params *p,
                   uint32_t *ids, float *dists) {
  KATOANN_TRY
  maybe_tune(*h, queries, nq, *p, nullptr, ids, dists);
  search_impl(*h, queries, nq, *p, nullptr, ids, dists);
  KATOANN_CATCH
}

// katoann_search_ex (device pointers) in `bytes` of caller-provided device
// memory instead of buffers of its own: no autotuning, graphs or stage timing,
// on x->stream; the caller keeps calls that share the memory apart.
int katoann_search_in_scratch(const katoann_index *h, const float *queries, uint32_t nq,
                              const katoann_search_params *p, const katoann_search_extras *x, uint32_t *ids,
                              float *dists, void *scratch, uint64_t bytes) {
  KATOANN_TRY
  if (!h || !p || !x) throw std::invalid_argument("null argument");
  if (!x->device_io) throw std::invalid_argument("the allocation-free search takes device pointers (device_io)");
  const FixedScratch fs{static_cast<char *>(scratch), bytes};
  search_impl(*h, queries, nq, *p, x, ids, dists, nullptr, &fs);
  KATOANN_CATCH
}

int katoann_search_ex(const katoann_index *h, const float *queries, uint32_t nq, const katoann_search_params *p,
                      const katoann_search_extras *x, uint32_t *ids, float *dists) {
  KATOANN_TRY
  if (x && x->nprobe_per_query && !x->device_io)
    for (uint32_t q = 0; q < nq; ++q)
      if (x->nprobe_per_query[q] == 0 || x->nprobe_per_query[q] > p->nprobe)
        throw std::invalid_argument("per-query nprobe must be in [1, nprobe]");
  maybe_tune(*h, queries, nq, *p, x, ids, dists);
  search_impl(*h, queries, nq, *p, x, ids, dists);
  KATOANN_CATCH
}

int katoann_set_autotune(katoann_index *h, int32_t on) {
  h->autotune.store(on != 0);
  return KATOANN_OK;
}

int katoann_tile_config(const katoann_index *h, uint32_t nq, const katoann_search_params *p, uint32_t *out) {
  KATOANN_TRY
  const TileConfig c = tile_for(*h, nq, *p, nullptr);
  out[0] = c.kernel;
  out[1] = c.variant;
  out[2] = c.qc;
  out[3] = c.shift;
  out[4] = c.seed;
  out[5] = c.join;
  KATOANN_CATCH
}

int katoann_search_footprint(const katoann_index *h, uint32_t nq, const katoann_search_params *p,
                             const katoann_search_extras *x, uint64_t *bytes) {
  KATOANN_TRY
  if (!h || !p || !bytes) throw std::invalid_argument("null argument");
  if (nq) {
    const SearchPlan z = plan_search(*h, nq, *p, x);
    *bytes = carve(z, pool_stride(z, z.allowance), nullptr, nullptr);
  } else {
    *bytes = 0;
  }
  KATOANN_CATCH
}

uint32_t katoann_dim(const katoann_index *h) { return h->dim; }
uint32_t katoann_nlist(const katoann_index *h) { return h->nlist; }
uint32_t katoann_size(const katoann_index *h) { return h->npts; }
int32_t katoann_normalized(const katoann_index *h) { return h->normalized ? 1 : 0; }
int32_t katoann_engine(const katoann_index *) { return KATOANN_ENGINE_GFX950; }
int32_t katoann_precision(const katoann_index *h) { return h->precision; }
uint32_t katoann_prefix_dims(const katoann_index *h) { return h->head_dims; }
int32_t katoann_rotated(const katoann_index *h) { return h->rotated ? 1 : 0; }
int32_t katoann_batched_coarse(const katoann_index *h) { return h->nlist && h->batched ? 1 : 0; }

int katoann_list_sizes(const katoann_index *h, uint32_t *out) {
  KATOANN_TRY
  std::copy(h->sizes.begin(), h->sizes.end(), out);
  KATOANN_CATCH
}

uint64_t katoann_memory_bytes(const katoann_index *h) {
  uint64_t s = 0;
  for (const DevBuf *b : {&h->rot, &h->list_off_d, &h->ids, &h->id_row, &h->hi, &h->lo_plane, &h->pre, &h->norm,
                          &h->hinorm, &h->headnorm, &h->cen_hi, &h->cen_lo, &h->cen_norm, &h->cen_f, &h->bf, &h->bfn})
    s += b->bytes;
  return s;
}

int katoann_index_scale(const katoann_index *h, float *lo, double *inv_step, double *step_sq) {
  KATOANN_TRY
  if (h->nlist == 0) throw logic_err("search before load");
  *lo = h->lo;
  *inv_step = h->inv_step;
  *step_sq = h->step_sq;
  KATOANN_CATCH
}

int katoann_export_codes(const katoann_index *h, uint16_t *codes, uint16_t *centroid_codes, uint32_t *ids) {
  KATOANN_TRY
  if (h->nlist == 0) throw logic_err("search before load");
  if (h->precision == KATOANN_BF16) throw std::invalid_argument("a bf16 index holds no 16-bit codes");
  const uint64_t N = h->nrows;
  const uint32_t dc = h->dc;
  auto unshift = [&](const DevBuf &hb, const DevBuf &lb, uint64_t rows, uint16_t *out) {
    std::vector<int8_t> a(static_cast<size_t>(rows) * dc), b(a.size());
    if (rows) {
      HIPCHECK(hipMemcpy(a.data(), hb.p, a.size(), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(b.data(), lb.p, b.size(), hipMemcpyDeviceToHost));
    }
    const uint32_t d = h->dim;
    for (uint64_t r = 0; r < rows; ++r)
      for (uint32_t k = 0; k < dc; ++k) {
        const size_t i = static_cast<size_t>(r) * dc + k;
        out[i] = k < d ? static_cast<uint16_t>(((a[i] + 128) << 8) | (b[i] + 128)) : 0;
      }
  };
  if (codes) unshift(h->hi, h->lo_plane, N, codes);
  if (centroid_codes) unshift(h->cen_hi, h->cen_lo, h->nlist, centroid_codes);
  if (ids && N) HIPCHECK(hipMemcpy(ids, h->ids.p, N * 4, hipMemcpyDeviceToHost));
  KATOANN_CATCH
}

int katoann_export_lists(const katoann_index *h, const uint32_t *lists, uint32_t nl, uint16_t *codes,
                         uint32_t *ids) {
  KATOANN_TRY
  if (h->nlist == 0) throw logic_err("search before load");
  if (h->precision == KATOANN_BF16) throw std::invalid_argument("a bf16 index holds no 16-bit codes");
  const uint32_t dc = h->dc, d = h->dim;
  uint64_t at = 0;
  std::vector<int8_t> a, b;
  for (uint32_t i = 0; i < nl; ++i) {
    const uint32_t c = lists[i];
    if (c >= h->nlist) throw std::invalid_argument("list index out of range");
    const uint64_t r0 = h->list_off[c], n = h->list_off[c + 1] - r0;
    if (n == 0) continue;
    if (codes) {
      a.resize(static_cast<size_t>(n) * dc);
      b.resize(a.size());
      HIPCHECK(hipMemcpy(a.data(), h->hi.as<int8_t>() + r0 * dc, a.size(), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(b.data(), h->lo_plane.as<int8_t>() + r0 * dc, b.size(), hipMemcpyDeviceToHost));
      for (uint64_t r = 0; r < n; ++r) {
        uint16_t *o = codes + (at + r) * dc;
        const int8_t *ha = a.data() + r * dc, *la = b.data() + r * dc;
        for (uint32_t k = 0; k < d; ++k) o[k] = static_cast<uint16_t>(((ha[k] + 128) << 8) | (la[k] + 128));
        for (uint32_t k = d; k < dc; ++k) o[k] = 0;
      }
    }
    if (ids) HIPCHECK(hipMemcpy(ids + at, h->ids.as<uint32_t>() + r0, n * 4, hipMemcpyDeviceToHost));
    at += n;
  }
  KATOANN_CATCH
}

int katoann_set_timing(katoann_index *h, int32_t on) {
  h->timing.store(on != 0);
  return KATOANN_OK;
}

int katoann_stage_times(const katoann_index *h, float *ms, int32_t n) {
  auto *m = const_cast<katoann_index *>(h);
  std::lock_guard<std::mutex> g(m->stat_mu);
  for (int i = 0; i < n && i < KATOANN_NSTAGES; ++i) ms[i] = h->stage_ms[i];
  return KATOANN_OK;
}

int katoann_band_times(const katoann_index *h, float *ms, int32_t n) {
  auto *m = const_cast<katoann_index *>(h);
  std::lock_guard<std::mutex> g(m->stat_mu);
  for (int i = 0; i < n; ++i) ms[i] = i < static_cast<int>(kMaxBands) ? h->band_ms[i] : 0.0f;
  return KATOANN_OK;
}

int katoann_last_work(const katoann_index *h, uint64_t *out, int32_t n) {
  auto *m = const_cast<katoann_index *>(h);
  std::lock_guard<std::mutex> g(m->stat_mu);
  for (int i = 0; i < n && i < 4; ++i) out[i] = h->work[i];
  return KATOANN_OK;
}

int katoann_set_rank_rows(katoann_index *h, uint32_t rows) {
  KATOANN_TRY
  if (rows == 0 || rows % 64) throw std::invalid_argument("rank rows must be a positive multiple of 64");
  std::lock_guard<std::mutex> g(h->mu);
  h->rank_rows = rows;
  KATOANN_CATCH
}

int katoann_set_band_factor(katoann_index *h, uint32_t log2_factor) {
  KATOANN_TRY
  if (log2_factor < 1 || log2_factor > 12) throw std::invalid_argument("band factor must be 2^1 .. 2^12");
  std::lock_guard<std::mutex> g(h->mu);
  h->band_shift = log2_factor;
  KATOANN_CATCH
}

int katoann_set_pool_allowance(katoann_index *h, uint32_t slots) {
  KATOANN_TRY
  std::lock_guard<std::mutex> g(h->mu);
  h->pool_allowance = slots;
  {
    std::lock_guard<std::mutex> t(h->tune_mu);
    h->allowance_seen.clear();
  }
  KATOANN_CATCH
}

int katoann_set_seed_rows(katoann_index *h, uint32_t rows) {
  KATOANN_TRY
  if (rows > (1u << 20)) throw std::invalid_argument("seed rows must be at most 2^20");
  std::lock_guard<std::mutex> g(h->mu);
  h->seed_rows = rows;
  KATOANN_CATCH
}

int katoann_set_seed_join(katoann_index *h, int32_t on) {
  KATOANN_TRY
  std::lock_guard<std::mutex> g(h->mu);
  h->seed_join = on ? 1u : 0u;
  KATOANN_CATCH
}

int katoann_set_graphs(katoann_index *h, int32_t on) {
  h->graphs.store(on != 0);
  return KATOANN_OK;
}

int katoann_set_rank_kernel(katoann_index *h, uint32_t kernel, uint32_t variant, uint32_t queries_per_item) {
  KATOANN_TRY
  if (kernel > 5 || variant > 3 || (kernel != 1 && variant > 1) || (kernel >= 4 && variant != 0))
    throw std::invalid_argument("rank kernel is 0 to 5; variant 0, 1 or (kernel 1) 2 or 3, 0 for kernels 4 and 5");
  if (queries_per_item == 0 || queries_per_item > 65536)
    throw std::invalid_argument("queries per ranking item must be in 1..65536");
  std::lock_guard<std::mutex> g(h->mu);
  h->rank_kernel = kernel;
  h->wave_variant = variant;
  h->wave_qc = queries_per_item;
  KATOANN_CATCH
}

int katoann_mlp_gemm(uint32_t M, uint32_t N, uint32_t K, const void *A, size_t sam, size_t sak, const void *B,
                     size_t sbn, size_t sbk, const float *bias, const float *gate, size_t ldg, int32_t relu,
                     float *C32, void *C16, size_t ldc, void *stream) {
  KATOANN_TRY
  if (!A || !B || (!C32 && !C16)) throw std::invalid_argument("gemm needs A, B and an output");
  katoann_hip::launch_gemm_bf16(M, N, K, A, sam, sak, B, sbn, sbk, bias, gate, ldg, relu, C32, C16, ldc,
                                static_cast<hipStream_t>(stream));
  HIPCHECK(hipGetLastError());
  KATOANN_CATCH
}

int katoann_mlp_to_bf16(const float *x, void *y, uint64_t n, void *stream) {
  KATOANN_TRY
  katoann_hip::launch_to_bf16(x, y, n, static_cast<hipStream_t>(stream));
  HIPCHECK(hipGetLastError());
  KATOANN_CATCH
}

int katoann_mlp_mse(const float *pred, const float *y, uint32_t n, float *grad, float *loss, void *stream) {
  KATOANN_TRY
  katoann_hip::launch_mse(pred, y, n, grad, loss, static_cast<hipStream_t>(stream));
  HIPCHECK(hipGetLastError());
  KATOANN_CATCH
}

int katoann_mlp_pinball(const float *pred, const float *y, uint32_t n, float tau, float *grad, float *loss,
                        void *stream) {
  KATOANN_TRY
  if (!(tau > 0.0f && tau < 1.0f)) throw std::invalid_argument("pinball quantile must be in (0, 1)");
  katoann_hip::launch_pinball(pred, y, n, tau, grad, loss, static_cast<hipStream_t>(stream));
  HIPCHECK(hipGetLastError());
  KATOANN_CATCH
}

int katoann_mlp_colsum(const float *g, uint32_t rows, uint32_t cols, float *out, void *stream) {
  KATOANN_TRY
  katoann_hip::launch_colsum(g, rows, cols, out, static_cast<hipStream_t>(stream));
  HIPCHECK(hipGetLastError());
  KATOANN_CATCH
}

int katoann_mlp_adam(float *w, const float *g, float *m, float *v, uint32_t n, float lr, float b1, float b2,
                     float eps, uint32_t step, void *w16, void *stream) {
  KATOANN_TRY
  if (step == 0) throw std::invalid_argument("Adam steps count from 1");
  katoann_hip::launch_adam(w, g, m, v, n, lr, b1, b2, eps, step, w16, static_cast<hipStream_t>(stream));
  HIPCHECK(hipGetLastError());
  KATOANN_CATCH
}

int katoann_mlp_features(const float *probe_dists, uint32_t nq, uint32_t nprobe, uint32_t nfeat, const float *mean,
                         const float *inv_std, float *x32, void *x16, void *stream) {
  KATOANN_TRY
  if (nfeat == 0 || nfeat > nprobe) throw std::invalid_argument("need 1 <= nfeat <= nprobe");
  katoann_hip::launch_features(probe_dists, nq, nprobe, nfeat, mean, inv_std, x32, x16,
                               static_cast<hipStream_t>(stream));
  HIPCHECK(hipGetLastError());
  KATOANN_CATCH
}

int katoann_normalize_rows(float *rows, uint64_t n, uint32_t dim) {
  KATOANN_TRY
  DevBuf b;
  upload(b, rows, n * dim);
  katoann_hip::launch_normalize(b.as<float>(), static_cast<uint32_t>(n), dim, nullptr);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(rows, b.p, n * dim * 4, hipMemcpyDeviceToHost));
  KATOANN_CATCH
}

int katoann_rotate_rows(const float *in, float *out, uint64_t n, uint32_t dim, const float *rotation) {
  KATOANN_TRY
  DevBuf a, r, y;
  upload(a, in, n * dim);
  upload(r, rotation, static_cast<size_t>(dim) * dim);
  y.ensure(n * dim * 4);
  katoann_hip::launch_rotate(a.as<float>(), static_cast<uint32_t>(n), dim, r.as<float>(), y.as<float>(), nullptr);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(out, y.p, n * dim * 4, hipMemcpyDeviceToHost));
  KATOANN_CATCH
}

int katoann_sq16_l2(const uint8_t *qhi, const uint8_t *qlo, const uint8_t *xhi, const uint8_t *xlo,
                    const int64_t *qnorm, const int64_t *xnorm, double step_sq, uint32_t m, uint32_t n,
                    uint32_t padded_dim, float *out) {
  KATOANN_TRY
  if (m == 0 || n == 0) return KATOANN_OK;
  if (padded_dim == 0 || padded_dim % 64) throw std::invalid_argument("padded_dim must be a multiple of 64");
  kernel_pairs(qhi, qlo, xhi, xlo, qnorm, xnorm, step_sq, m, n, padded_dim, padded_dim, padded_dim, true, out);
  KATOANN_CATCH
}

int katoann_sq16_high_l2(const uint8_t *qhi, const uint8_t *xhi, const int64_t *qnorm, const int64_t *xnorm,
                         double step_sq, uint32_t m, uint32_t n, uint32_t padded_dim, float *out,
                         uint32_t plane_dim) {
  KATOANN_TRY
  if (m == 0 || n == 0) return KATOANN_OK;
  if (plane_dim == 0) plane_dim = padded_dim;
  if (plane_dim < padded_dim || plane_dim % 64)
    throw std::invalid_argument("plane_dim must be a multiple of 64 and at least padded_dim");
  if (padded_dim == 0 || padded_dim % 64) throw std::invalid_argument("padded_dim must be a multiple of 64");
  kernel_pairs(qhi, nullptr, xhi, nullptr, qnorm, xnorm, step_sq, m, n, padded_dim, padded_dim, plane_dim, false,
               out);
  KATOANN_CATCH
}

int katoann_bf16_l2(const uint16_t *q, const uint16_t *x, const float *qnorm, const float *xnorm, uint32_t m,
                    uint32_t n, uint32_t dim, float *out) {
  KATOANN_TRY
  if (m == 0 || n == 0) return KATOANN_OK;
  if (dim == 0 || dim % 2) throw std::invalid_argument("dim must be a positive multiple of 2");
  DevBuf a, b, c, e, o;
  upload(a, q, static_cast<size_t>(m) * dim);
  upload(b, x, static_cast<size_t>(n) * dim);
  upload(c, qnorm, m);
  upload(e, xnorm, n);
  o.ensure(static_cast<size_t>(m) * n * 4);
  katoann_hip::launch_bf16_pairs(a.as<uint16_t>(), b.as<uint16_t>(), c.as<float>(), e.as<float>(), m, n, dim,
                                 o.as<float>(), nullptr);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(out, o.p, static_cast<size_t>(m) * n * 4, hipMemcpyDeviceToHost));
  KATOANN_CATCH
}

}  // extern "C"

namespace {

// dot of unsigned codes from the shifted-byte MFMA product:
//   full: sum cq cx = dot' + 32896 (sum c'q + sum c'x) + dc * 32896^2
//   high: sum hq hx = dot' + 128 (sum s_q + sum s_x) + dc * 128^2
void kernel_pairs(const uint8_t *qh, const uint8_t *ql, const uint8_t *xh, const uint8_t *xl, const int64_t *qn,
                  const int64_t *xn, double step_sq, uint32_t m, uint32_t n, uint32_t dc, uint32_t qstride,
                  uint32_t xstride, bool full, float *out) {
  auto shift = [&](const uint8_t *src, uint32_t rows, uint32_t stride) {
    std::vector<int8_t> v(static_cast<size_t>(rows) * stride);
    for (size_t i = 0; i < v.size(); ++i) v[i] = static_cast<int8_t>(static_cast<int>(src[i]) - 128);
    return v;
  };
  auto sums = [&](const uint8_t *h, const uint8_t *l, uint32_t rows, uint32_t stride) {
    std::vector<int64_t> s(rows, 0);
    for (uint32_t r = 0; r < rows; ++r)
      for (uint32_t k = 0; k < dc; ++k) {
        const int64_t hv = h[static_cast<size_t>(r) * stride + k];
        s[r] += full ? (hv * 256 + l[static_cast<size_t>(r) * stride + k]) - 32896 : hv - 128;
      }
    return s;
  };
  const auto sqh = shift(qh, m, qstride), sxh = shift(xh, n, xstride);
  std::vector<int8_t> sql, sxl;
  if (full) {
    sql = shift(ql, m, qstride);
    sxl = shift(xl, n, xstride);
  }
  const auto qs = sums(qh, ql, m, qstride), xs = sums(xh, xl, n, xstride);
  const int64_t K = full ? 32896 : 128;
  // fold the bias into the norms: l2 = qn + xn - 2 (dot' + K (sq + sx) + dc K^2)
  std::vector<int64_t> qn2(m), xn2(n);
  for (uint32_t i = 0; i < m; ++i) qn2[i] = qn[i] - 2 * K * qs[i] - static_cast<int64_t>(dc) * K * K;
  for (uint32_t j = 0; j < n; ++j) xn2[j] = xn[j] - 2 * K * xs[j] - static_cast<int64_t>(dc) * K * K;
  DevBuf a, b, c, e, f, g, o;
  upload(a, sqh.data(), sqh.size());
  upload(b, sxh.data(), sxh.size());
  if (full) {
    upload(c, sql.data(), sql.size());
    upload(e, sxl.data(), sxl.size());
  }
  upload(f, qn2.data(), m);
  upload(g, xn2.data(), n);
  o.ensure(static_cast<size_t>(m) * n * 4);
  katoann_hip::launch_sq16_pairs(a.as<int8_t>(), full ? c.as<int8_t>() : a.as<int8_t>(), f.as<int64_t>(),
                                 b.as<int8_t>(), full ? e.as<int8_t>() : b.as<int8_t>(), g.as<int64_t>(), m, n, dc,
                                 qstride, xstride, full ? 1 : 0, step_sq, o.as<float>(), nullptr);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(out, o.p, static_cast<size_t>(m) * n * 4, hipMemcpyDeviceToHost));
}

}  // namespace
