// Driver around the reference library (compiled from its own sources by
// oracle/Makefile into oracle/_ref/). Test infrastructure: it produces the
// golden vectors under tests/golden/ in the container that holds the reference.
//
//   katoann_ref_search dump   PREFIX OUT.bin [prefix_dims]
//       the loaded state: scale, planes (unpacked to rows), norms, centroid codes
//   katoann_ref_search rotate PREFIX OUT.bin
//       the rotated base vectors (f32, list order) from the load path's
//       rotation of every list (index.cpp:613-629)
//   katoann_ref_search search PREFIX QUERIES.fbin OUT.bin k nprobe[,nprobe...] survivors
//                             ranking(full|high|prefix) selection(threshold|local)
//                             [batch=4096] [normalize=0] [mode=batched|flat|bf16]
//       ids (u32) and dists (f32), nq x k, from Index::search in batches (one
//       block of both per nprobe, in the order given, from one load);
//       flat loads with LoadOptions::batched_coarse = false, bf16 with
//       LoadOptions::precision = kBf16
//   katoann_ref_search flat   PREFIX QUERIES.fbin OUT.bin
//       per query, every list in the order of the per-query coarse search
//       (Index::coarse_flat_ with nprobe = nlist), u32, nq x nlist
//
// The loaded state is read through the class's private members; the access
// specifiers are opened for this translation unit only (the object layout of
// the reference's Index does not depend on them).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#define private public
#include "katoann/index.h"
#undef private

using katoann::Index;

namespace {

void put(std::ofstream &f, const void *p, size_t n) {
  f.write(static_cast<const char *>(p), static_cast<std::streamsize>(n));
}

// Inverse of the reference's packed tile order (kernels.h:8-17,
// kernels.cpp:263-275): tile (n/16, k/64), row (k%64)/4, column n%16, byte k%4.
uint8_t packed_at(const std::vector<uint8_t> &pk, uint32_t dc, uint32_t n, uint32_t k) {
  const uint32_t nc = dc / 64;
  const size_t off = ((static_cast<size_t>(n / 16) * nc + k / 64) * 16 + (k % 64) / 4) * 64 +
                     (n % 16) * 4 + k % 4;
  return pk[off];
}

int cmd_dump(const std::string &prefix, const std::string &out, uint32_t pd) {
  Index idx;
  katoann::LoadOptions lo;
  lo.use_amx = false;
  lo.prefix_dims = pd;
  idx.load(prefix, lo);
  std::ofstream f(out, std::ios::binary);
  const uint32_t hdr[8] = {idx.dim_, idx.dc_, idx.nlist_, idx.npts_, idx.head_dims_,
                           idx.rotated_ ? 1u : 0u, idx.coarse_batched_ ? 1u : 0u, 0u};
  put(f, hdr, sizeof(hdr));
  put(f, &idx.sq_lo_, 4);
  put(f, &idx.sq_inv_step_, 8);
  put(f, &idx.sq_step_, 8);
  put(f, &idx.sq_step_sq_, 8);
  const uint32_t dc = idx.dc_;
  for (uint32_t c = 0; c < idx.nlist_; ++c) {
    const uint32_t n = static_cast<uint32_t>(idx.ids_[c].size());
    put(f, &n, 4);
    put(f, idx.ids_[c].data(), n * 4ull);
    std::vector<uint8_t> hi(static_cast<size_t>(n) * dc), lo2(static_cast<size_t>(n) * dc);
    for (uint32_t i = 0; i < n; ++i)
      for (uint32_t k = 0; k < dc; ++k) {
        hi[static_cast<size_t>(i) * dc + k] = packed_at(idx.hi_[c], dc, i, k);
        lo2[static_cast<size_t>(i) * dc + k] = packed_at(idx.lo_[c], dc, i, k);
      }
    put(f, hi.data(), hi.size());
    put(f, lo2.data(), lo2.size());
    put(f, idx.norm_[c].data(), n * 8ull);
    put(f, idx.head_dims_ ? idx.head_norm_[c].data() : idx.hi_norm_[c].data(), n * 8ull);
  }
  std::vector<uint8_t> ch(static_cast<size_t>(idx.nlist_) * dc), cl(ch.size());
  for (uint32_t c = 0; c < idx.nlist_; ++c)
    for (uint32_t k = 0; k < dc; ++k) {
      ch[static_cast<size_t>(c) * dc + k] = packed_at(idx.cen_hi_, dc, c, k);
      cl[static_cast<size_t>(c) * dc + k] = packed_at(idx.cen_lo_, dc, c, k);
    }
  put(f, ch.data(), ch.size());
  put(f, cl.data(), cl.size());
  put(f, idx.cen_norm_.data(), idx.cen_norm_.size() * 8);
  return f ? 0 : 1;
}

int cmd_rotate(const std::string &prefix, const std::string &out) {
  Index idx;
  idx.read_files_(prefix);
  uint32_t stored = 0;
  if (!idx.read_rotation_(prefix, &stored) || !idx.rotated_) return 4;
  idx.rotate_lists_();
  std::ofstream f(out, std::ios::binary);
  for (uint32_t c = 0; c < idx.nlist_; ++c) put(f, idx.raw_[c].data(), idx.raw_[c].size() * 4);
  return f ? 0 : 1;
}

std::vector<float> read_fbin(const std::string &path, uint32_t &n, uint32_t &d) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  f.read(reinterpret_cast<char *>(&n), 4);
  f.read(reinterpret_cast<char *>(&d), 4);
  std::vector<float> v(static_cast<size_t>(n) * d);
  f.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(v.size() * 4));
  if (!f) throw std::runtime_error("truncated " + path);
  return v;
}

int cmd_search(int argc, char **argv) {
  if (argc < 10) return 2;
  const std::string prefix = argv[2], qpath = argv[3], out = argv[4];
  katoann::SearchParams p;
  p.k = static_cast<uint32_t>(std::stoul(argv[5]));
  // nprobe: one value, or several separated by commas (one load, the searches
  // one after another, outputs concatenated in that order)
  std::vector<uint32_t> nprobes;
  for (size_t at = 0, nx; at <= std::string(argv[6]).size(); at = nx + 1) {
    const std::string v(argv[6]);
    nx = v.find(',', at);
    if (nx == std::string::npos) nx = v.size();
    nprobes.push_back(static_cast<uint32_t>(std::stoul(v.substr(at, nx - at))));
  }
  p.survivors = static_cast<uint32_t>(std::stoul(argv[7]));
  const std::string r = argv[8], s = argv[9];
  p.ranking = r == "full" ? katoann::Ranking::kFull
              : r == "high" ? katoann::Ranking::kHighBytes
                            : katoann::Ranking::kPrefix;
  p.selection = s == "local" ? katoann::Selection::kLocal : katoann::Selection::kThreshold;
  const uint32_t batch = argc > 10 ? static_cast<uint32_t>(std::stoul(argv[10])) : 4096u;
  const bool normalize = argc > 11 && std::atoi(argv[11]) != 0;
  Index idx;
  katoann::LoadOptions lo;
  lo.use_amx = false;
  const std::string mode = argc > 12 ? argv[12] : "batched";
  lo.batched_coarse = mode != "flat";
  if (mode == "bf16") lo.precision = katoann::Precision::kBf16;
  idx.load(prefix, lo);
  uint32_t nq = 0, d = 0;
  std::vector<float> q = read_fbin(qpath, nq, d);
  if (normalize) {
    // the CLI and Python module normalise before Index::search; reproduce
    // through the reference's own CLI when that step itself is under test.
    return 3;
  }
  std::ofstream f(out, std::ios::binary);
  for (uint32_t np : nprobes) {
    p.nprobe = np;
    std::vector<uint32_t> ids(static_cast<size_t>(nq) * p.k);
    std::vector<float> dists(ids.size());
    for (uint32_t off = 0; off < nq; off += batch) {
      const uint32_t b = std::min(batch, nq - off);
      idx.search(q.data() + static_cast<size_t>(off) * d, b, p, ids.data() + static_cast<size_t>(off) * p.k,
                 dists.data() + static_cast<size_t>(off) * p.k);
    }
    put(f, ids.data(), ids.size() * 4);
    put(f, dists.data(), dists.size() * 4);
  }
  return f ? 0 : 1;
}

int cmd_flat(const std::string &prefix, const std::string &qpath, const std::string &out) {
  Index idx;
  katoann::LoadOptions lo;
  lo.use_amx = false;
  lo.batched_coarse = false;
  idx.load(prefix, lo);
  uint32_t nq = 0, d = 0;
  std::vector<float> q = read_fbin(qpath, nq, d);
  const uint32_t nlist = idx.nlist();
  std::vector<uint32_t> order(static_cast<size_t>(nq) * nlist);
  for (uint32_t i = 0; i < nq; ++i)
    idx.coarse_flat_(q.data() + static_cast<size_t>(i) * d, nlist, order.data() + static_cast<size_t>(i) * nlist);
  std::ofstream f(out, std::ios::binary);
  put(f, order.data(), order.size() * 4);
  return f ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    if (argc >= 4 && std::string(argv[1]) == "dump")
      return cmd_dump(argv[2], argv[3], argc > 4 ? static_cast<uint32_t>(std::stoul(argv[4])) : 0u);
    if (argc >= 4 && std::string(argv[1]) == "rotate") return cmd_rotate(argv[2], argv[3]);
    if (argc >= 2 && std::string(argv[1]) == "search") return cmd_search(argc, argv);
    if (argc >= 5 && std::string(argv[1]) == "flat") return cmd_flat(argv[2], argv[3], argv[4]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "katoann_ref_search: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr, "usage: katoann_ref_search dump|rotate|search|flat ...\n");
  return 2;
}
