// The reference's C++ API (include/katoann/index.h of KatoANN) for the gfx950
// engine: same namespace, option structs, enums, class, signatures and
// exception types, implemented inline over the C ABI (katoann_c.h) so a
// program that includes this header and links libkatoann.so builds unchanged.
//
//   reference index.h:22-30  BuildOptions        -> katoann_build_options
//   reference index.h:37-43  LoadOptions         -> katoann_load_options (use_amx ignored)
//   reference index.h:45-62  Ranking, Selection, SearchParams -> katoann_search_params
//   reference index.h:74-75  Index::build        -> katoann_build
//   reference index.h:77     Index::load         -> katoann_load
//   reference index.h:80-81  Index::search       -> katoann_search
//   reference index.h:83-93  accessors           -> katoann_dim ... katoann_memory_bytes
// Errors come back as the reference's std::invalid_argument / runtime_error /
// logic_error with its messages.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "katoann/kernels.h"
#include "katoann_c.h"

namespace katoann {

struct BuildOptions {
  uint32_t nlist = 2048;
  uint32_t train_size = 256000;
  uint32_t kmeans_iters = 12;
  uint64_t seed = 0;              // seed of the samples and initial centres; 0 picks one
  bool normalized = false;        // record that the vectors are unit-length
  uint32_t rotation_sample = 100000;
  uint32_t jacobi_sweeps = 8;     // accepted; the eigensolver is rocSOLVER's divide and conquer
};

enum class Precision {
  kSq16,  // 16-bit codes in high and low byte planes (default)
  kBf16,  // bf16 rows, Ranking::kFull only (index.h:32-35)
};

struct LoadOptions {
  Precision precision = Precision::kSq16;
  bool use_amx = true;            // accepted for compatibility; the engine is the GPU
  uint32_t prefix_dims = 0;       // 0 takes the value stored in P.dimc
  bool batched_coarse = true;     // false: the per-query fp32 coarse search
};

enum class Ranking { kFull, kHighBytes, kPrefix };
enum class Selection { kThreshold, kLocal };

struct SearchParams {
  uint32_t k = 10;
  uint32_t nprobe = 64;
  Ranking ranking = Ranking::kPrefix;
  uint32_t survivors = 32;        // candidates refined per query (at least k)
  Selection selection = Selection::kThreshold;
};

namespace detail {
inline void check(int rc) {
  if (rc == KATOANN_OK) return;
  const std::string msg = katoann_last_error();
  if (rc == KATOANN_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  if (rc == KATOANN_LOGIC_ERROR) throw std::logic_error(msg);
  throw std::runtime_error(msg);
}
}  // namespace detail

class Index {
 public:
  Index() { detail::check(katoann_index_create(&h_)); }
  ~Index() {
    if (h_) katoann_index_destroy(h_);
  }
  Index(const Index &) = delete;
  Index &operator=(const Index &) = delete;
  Index(Index &&o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
  Index &operator=(Index &&o) noexcept {
    if (this != &o) {
      if (h_) katoann_index_destroy(h_);
      h_ = std::exchange(o.h_, nullptr);
    }
    return *this;
  }

  static void build(const float *data, uint32_t n, uint32_t dim, const BuildOptions &opt,
                    const std::string &prefix) {
    const katoann_build_options o{opt.nlist, opt.train_size, opt.kmeans_iters, opt.seed,
                                  opt.normalized ? 1 : 0, opt.rotation_sample, opt.jacobi_sweeps};
    detail::check(katoann_build(data, n, dim, &o, prefix.c_str()));
  }

  void load(const std::string &prefix, const LoadOptions &opt = {}) {
    const katoann_load_options o{opt.precision == Precision::kBf16 ? KATOANN_BF16 : KATOANN_SQ16,
                                 opt.use_amx ? 1 : 0, opt.prefix_dims, opt.batched_coarse ? 1 : 0};
    detail::check(katoann_load(h_, prefix.c_str(), &o));
  }

  // Searches nq queries together. ids and dists are nq x k, nearest first.
  void search(const float *queries, uint32_t nq, const SearchParams &p, uint32_t *ids, float *dists) const {
    const katoann_search_params sp{p.k, p.nprobe,
                                   p.ranking == Ranking::kFull ? KATOANN_RANK_FULL
                                   : p.ranking == Ranking::kHighBytes ? KATOANN_RANK_HIGH
                                                                      : KATOANN_RANK_PREFIX,
                                   p.survivors,
                                   p.selection == Selection::kLocal ? KATOANN_SEL_LOCAL : KATOANN_SEL_THRESHOLD};
    detail::check(katoann_search(h_, queries, nq, &sp, ids, dists));
  }

  uint32_t dim() const { return katoann_dim(h_); }
  uint32_t nlist() const { return katoann_nlist(h_); }
  uint32_t size() const { return katoann_size(h_); }
  bool normalized() const { return katoann_normalized(h_) != 0; }
  Engine engine() const { return static_cast<Engine>(katoann_engine(h_)); }
  Precision precision() const { return katoann_precision(h_) == KATOANN_BF16 ? Precision::kBf16 : Precision::kSq16; }
  uint32_t prefix_dims() const { return katoann_prefix_dims(h_); }
  bool rotated() const { return katoann_rotated(h_) != 0; }
  bool batched_coarse() const { return katoann_batched_coarse(h_) != 0; }
  std::vector<uint32_t> list_sizes() const {
    std::vector<uint32_t> v(nlist());
    detail::check(katoann_list_sizes(h_, v.data()));
    return v;
  }
  size_t memory_bytes() const { return static_cast<size_t>(katoann_memory_bytes(h_)); }

  // The C handle, for the extension entry points of katoann_c.h.
  katoann_index *handle() const { return h_; }

 private:
  katoann_index *h_ = nullptr;
};

}  // namespace katoann
