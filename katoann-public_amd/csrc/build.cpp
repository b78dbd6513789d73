// Index::build on the host (src/index.cpp:419-533): the coarse partition and
// the PCA rotation, written in the reference's file format. Not on the search
// path; the GPU only ever reads the files this writes.
//
// Structure as in the reference: strided training sample, k-means++ seeding
// (mt19937 with the given seed; 0 draws one), Lloyd iterations (the reference
// uses Elkan's bounds, which give the same assignments), nearest-centre
// assignment of every vector with ids ascending inside each list, and the PCA
// basis of a strided sample by cyclic Jacobi sweeps, rows sorted by
// decreasing variance.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/katoann_c.h"

namespace {

constexpr uint32_t kDimcMagic = 0x434D4944u;
constexpr uint32_t kDimcVersion = 2u;

float l2(const float *a, const float *b, uint32_t d) {
  float s = 0.0f;
  for (uint32_t k = 0; k < d; ++k) {
    const float t = a[k] - b[k];
    s += t * t;
  }
  return s;
}

void assign(const float *x, size_t n, uint32_t d, const float *cen, uint32_t nc, uint32_t *out) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) {
    const float *v = x + static_cast<size_t>(i) * d;
    float best = std::numeric_limits<float>::max();
    uint32_t bi = 0;
    for (uint32_t c = 0; c < nc; ++c) {
      const float dd = l2(v, cen + static_cast<size_t>(c) * d, d);
      if (dd < best) {
        best = dd;
        bi = c;
      }
    }
    out[i] = bi;
  }
}

void kmeanspp(const float *x, size_t n, uint32_t d, float *cen, uint32_t nc, uint64_t seed) {
  std::mt19937 gen(seed ? static_cast<std::mt19937::result_type>(seed) : std::random_device{}());
  std::uniform_int_distribution<size_t> pick(0, n - 1);
  std::uniform_real_distribution<double> unif(0.0, 1.0);
  size_t first = pick(gen);
  std::memcpy(cen, x + first * d, d * sizeof(float));
  std::vector<float> dist(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) dist[i] = l2(x + static_cast<size_t>(i) * d, cen, d);
  for (uint32_t c = 1; c < nc; ++c) {
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum += dist[i];
    size_t next = pick(gen);
    if (sum > 0.0) {
      const double dart = unif(gen) * sum;
      double acc = 0.0;
      for (size_t i = 0; i < n; ++i) {
        acc += dist[i];
        if (acc > dart) {
          next = i;
          break;
        }
      }
    }
    float *cc = cen + static_cast<size_t>(c) * d;
    std::memcpy(cc, x + next * d, d * sizeof(float));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < static_cast<int64_t>(n); ++i)
      dist[i] = std::min(dist[i], l2(x + static_cast<size_t>(i) * d, cc, d));
  }
}

void lloyd(const float *x, size_t n, uint32_t d, float *cen, uint32_t nc, uint32_t iters) {
  std::vector<uint32_t> a(n);
  std::vector<double> sums(static_cast<size_t>(nc) * d);
  std::vector<size_t> cnt(nc);
  for (uint32_t it = 0; it < iters; ++it) {
    assign(x, n, d, cen, nc, a.data());
    std::fill(sums.begin(), sums.end(), 0.0);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (size_t i = 0; i < n; ++i) {
      double *s = sums.data() + static_cast<size_t>(a[i]) * d;
      const float *v = x + i * d;
      for (uint32_t k = 0; k < d; ++k) s[k] += v[k];
      ++cnt[a[i]];
    }
    for (uint32_t c = 0; c < nc; ++c)
      if (cnt[c])
        for (uint32_t k = 0; k < d; ++k)
          cen[static_cast<size_t>(c) * d + k] = static_cast<float>(sums[static_cast<size_t>(c) * d + k] / cnt[c]);
  }
}

// Eigenvectors of the sample covariance by cyclic Jacobi rotations; row j of
// rot is the axis of the j-th largest variance.
bool pca(std::vector<float> &sample, size_t n, uint32_t d, uint32_t sweeps, std::vector<float> &rot) {
  if (n < 2 || d < 2) return false;
  std::vector<double> mean(d, 0.0);
  for (size_t s = 0; s < n; ++s)
    for (uint32_t k = 0; k < d; ++k) mean[k] += sample[s * d + k];
  for (auto &m : mean) m /= static_cast<double>(n);
  std::vector<double> A(static_cast<size_t>(d) * d, 0.0);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t i = 0; i < static_cast<int64_t>(d); ++i)
    for (size_t s = 0; s < n; ++s) {
      const double xi = sample[s * d + i] - mean[i];
      for (uint32_t j = static_cast<uint32_t>(i); j < d; ++j) A[static_cast<size_t>(i) * d + j] += xi * (sample[s * d + j] - mean[j]);
    }
  for (uint32_t i = 0; i < d; ++i)
    for (uint32_t j = i; j < d; ++j) {
      const double v = A[static_cast<size_t>(i) * d + j] / static_cast<double>(n);
      A[static_cast<size_t>(i) * d + j] = A[static_cast<size_t>(j) * d + i] = v;
    }
  std::vector<double> V(static_cast<size_t>(d) * d, 0.0);
  for (uint32_t i = 0; i < d; ++i) V[static_cast<size_t>(i) * d + i] = 1.0;
  sweeps = std::max(1u, std::min(sweeps, 64u));
  for (uint32_t sw = 0; sw < sweeps; ++sw) {
    double off = 0.0, diag = 0.0;
    for (uint32_t i = 0; i < d; ++i) {
      diag += A[static_cast<size_t>(i) * d + i] * A[static_cast<size_t>(i) * d + i];
      for (uint32_t j = i + 1; j < d; ++j) off += 2.0 * A[static_cast<size_t>(i) * d + j] * A[static_cast<size_t>(i) * d + j];
    }
    if (off <= 1e-18 * std::max(diag, 1e-300)) break;
    for (uint32_t p = 0; p + 1 < d; ++p)
      for (uint32_t q = p + 1; q < d; ++q) {
        const double apq = A[static_cast<size_t>(p) * d + q];
        const double app = A[static_cast<size_t>(p) * d + p], aqq = A[static_cast<size_t>(q) * d + q];
        if (std::fabs(apq) <= 1e-14 * std::sqrt(std::fabs(app * aqq)) + 1e-300) continue;
        // rotation angle that zeroes A[p][q] (Golub & Van Loan, symmetric Schur)
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
        for (uint32_t k = 0; k < d; ++k) {  // rows p, q
          const double a = A[static_cast<size_t>(p) * d + k], b = A[static_cast<size_t>(q) * d + k];
          A[static_cast<size_t>(p) * d + k] = cs * a - sn * b;
          A[static_cast<size_t>(q) * d + k] = sn * a + cs * b;
        }
        for (uint32_t k = 0; k < d; ++k) {  // columns p, q
          const double a = A[static_cast<size_t>(k) * d + p], b = A[static_cast<size_t>(k) * d + q];
          A[static_cast<size_t>(k) * d + p] = cs * a - sn * b;
          A[static_cast<size_t>(k) * d + q] = sn * a + cs * b;
        }
        for (uint32_t k = 0; k < d; ++k) {
          const double a = V[static_cast<size_t>(k) * d + p], b = V[static_cast<size_t>(k) * d + q];
          V[static_cast<size_t>(k) * d + p] = cs * a - sn * b;
          V[static_cast<size_t>(k) * d + q] = sn * a + cs * b;
        }
      }
  }
  std::vector<uint32_t> order(d);
  for (uint32_t i = 0; i < d; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return A[static_cast<size_t>(a) * d + a] > A[static_cast<size_t>(b) * d + b];
  });
  rot.resize(static_cast<size_t>(d) * d);
  for (uint32_t j = 0; j < d; ++j)
    for (uint32_t k = 0; k < d; ++k) rot[static_cast<size_t>(j) * d + k] = static_cast<float>(V[static_cast<size_t>(k) * d + order[j]]);
  return true;
}

void build(const float *data, uint32_t n, uint32_t dim, const katoann_build_options &opt, const std::string &prefix) {
  if (dim == 0 || dim > 4096) throw std::invalid_argument("dimension must be 1 to 4096");
  if (n == 0 || opt.nlist == 0 || opt.nlist > n) throw std::invalid_argument("need at least nlist vectors");
  const uint32_t ts = std::min(opt.train_size, n);
  if (ts > (1u << 23)) throw std::invalid_argument("train_size is limited to 8388608");
  std::vector<float> train(static_cast<size_t>(ts) * dim);
  const uint64_t stride = static_cast<uint64_t>(n) * (1ull << 32) / ts;
  for (uint32_t i = 0; i < ts; ++i) {
    uint32_t src = ts == n ? i : static_cast<uint32_t>((static_cast<uint64_t>(i) * stride) >> 32);
    if (src >= n) src = n - 1;
    std::memcpy(train.data() + static_cast<size_t>(i) * dim, data + static_cast<size_t>(src) * dim, dim * 4);
  }
  std::vector<float> cen(static_cast<size_t>(opt.nlist) * dim);
  kmeanspp(train.data(), ts, dim, cen.data(), opt.nlist, opt.seed);
  lloyd(train.data(), ts, dim, cen.data(), opt.nlist, opt.kmeans_iters);
  train.clear();
  train.shrink_to_fit();
  std::vector<uint32_t> a(n);
  assign(data, n, dim, cen.data(), opt.nlist, a.data());
  std::vector<std::vector<uint32_t>> lists(opt.nlist);
  for (uint32_t i = 0; i < n; ++i) lists[a[i]].push_back(i);
  {
    std::ofstream f(prefix + ".meta", std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + prefix + ".meta");
    const uint32_t meta[5] = {dim, opt.nlist, n, 0u, opt.normalized ? 1u : 0u};
    f.write(reinterpret_cast<const char *>(meta), sizeof(meta));
  }
  {
    std::ofstream f(prefix + ".centroids", std::ios::binary);
    f.write(reinterpret_cast<const char *>(cen.data()), static_cast<std::streamsize>(cen.size() * 4));
  }
  {
    std::ofstream f(prefix + ".invlist", std::ios::binary), g(prefix + ".norms", std::ios::binary);
    for (const auto &l : lists) {
      const uint32_t sz = static_cast<uint32_t>(l.size());
      f.write(reinterpret_cast<const char *>(&sz), 4);
      f.write(reinterpret_cast<const char *>(l.data()), static_cast<std::streamsize>(sz) * 4);
      for (uint32_t id : l) f.write(reinterpret_cast<const char *>(data + static_cast<size_t>(id) * dim), dim * 4);
      g.write(reinterpret_cast<const char *>(&sz), 4);
    }
    if (!f || !g) throw std::runtime_error("cannot write " + prefix + ".invlist");
  }
  // PCA sample: every stride-th vector in list order (src/index.cpp:516-533)
  const size_t target = std::min<size_t>(std::max<uint32_t>(opt.rotation_sample, 1024u), n);
  const size_t step = std::max<size_t>(1, n / target);
  std::vector<float> sample;
  size_t taken = 0, gi = 0;
  for (const auto &l : lists) {
    for (uint32_t id : l) {
      if (taken >= target) break;
      if (gi++ % step) continue;
      sample.insert(sample.end(), data + static_cast<size_t>(id) * dim, data + static_cast<size_t>(id + 1) * dim);
      ++taken;
    }
  }
  std::vector<float> rot;
  const bool rotated = pca(sample, taken, dim, opt.jacobi_sweeps, rot);
  const uint32_t dc = (dim + 63) / 64 * 64;
  const uint32_t head = std::min(std::max(64u, (dim / 4 + 63) / 64 * 64), dc);
  std::ofstream f(prefix + ".dimc", std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + prefix + ".dimc");
  const uint32_t hdr[8] = {kDimcMagic, kDimcVersion, dim, opt.nlist, n, head, rotated ? 1u : 0u, 0u};
  f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
  if (rotated) f.write(reinterpret_cast<const char *>(rot.data()), static_cast<std::streamsize>(rot.size() * 4));
  if (!f) throw std::runtime_error("cannot write " + prefix + ".dimc");
}

}  // namespace

extern std::string &katoann_error_slot();

extern "C" int katoann_build(const float *data, uint32_t n, uint32_t dim, const katoann_build_options *opt,
                             const char *prefix) {
  try {
    katoann_build_options o;
    katoann_default_build_options(&o);
    if (opt) o = *opt;
    build(data, n, dim, o, prefix);
    return KATOANN_OK;
  } catch (const std::invalid_argument &e) {
    katoann_error_slot() = e.what();
    return KATOANN_INVALID_ARGUMENT;
  } catch (const std::exception &e) {
    katoann_error_slot() = e.what();
    return KATOANN_RUNTIME_ERROR;
  }
}
