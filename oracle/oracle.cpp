// CPU oracle for KatoANN's batched IVF search (16-bit split-byte codes, prefix
// ranking, exact refinement). TEST INFRASTRUCTURE ONLY: it is loaded by tests/,
// by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker,
// and is never linked into or called by the product library.
//
// Every function restates the reference algorithm from its description, with
// plain scalar arithmetic, and cites the reference lines it follows
// (src/index.cpp, src/kernels.cpp of GyuyeongKim/KatoANN-public).
//
// Pinning (how each part is checked, see tests/test_oracle_*.py):
//   rotate_rows      golden vectors from the reference build (oracle/_ref) and
//                    an fp64 bound; order pinned bit-exact by the golden files.
//   l2_flat          the reference build's per-query coarse order of every
//                    list (golden), an fp64 bound.
//   quantize/norms   closed form, brute force, golden scale (lo, step_sq).
//   sq16 distances   the reference's own kernel test formula
//                    (tests/test_kernels.cpp:66-108), plus golden outputs.
//   coarse/ranking/  brute force over all candidates with an independent
//   refinement       sort, and golden end-to-end outputs of the reference.
//
// Arithmetic notes
//   f(l2) = float(double(l2) * step_sq)  (kernels.cpp:166-168, 208-210)
//   codes: v = clamp(lround((double(x) - lo) * inv_step), 0, 65535)
//          (index.cpp:665-667, 925-926); the oracle keeps unsigned codes and
//          the direct formula, independent of the GPU's shifted-byte form.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

namespace {

inline float fscale(int64_t l2, double step_sq) {
  return static_cast<float>(static_cast<double>(l2) * step_sq);
}

// lround with the x86 out-of-range result (LONG_MIN, then clamped to 0);
// NaN and |x| >= 2^63 take that path (index.cpp:665-666 via glibc lround).
inline uint32_t quant(float x, float lo, double inv_step) {
  const double t = (static_cast<double>(x) - static_cast<double>(lo)) * inv_step;
  if (!(t > -9.2e18 && t < 9.2e18)) return 0u;
  long r = std::lround(t);
  if (r < 0) r = 0;
  else if (r > 65535) r = 65535;
  return static_cast<uint32_t>(r);
}

struct OIndex {
  uint32_t d = 0, dc = 0, nlist = 0, npts = 0, P = 0;
  bool rotated = false;
  std::vector<float> rotation;     // d x d, row j = basis vector j
  float lo = 0.0f;
  double inv_step = 0.0, step_sq = 0.0;
  std::vector<uint64_t> list_off;  // nlist + 1 row offsets (list order)
  std::vector<uint32_t> ids;       // N
  std::vector<uint16_t> codes;     // N x dc, zero padded
  std::vector<int64_t> norm, hinorm, headnorm;  // N
  std::vector<uint16_t> ccodes;    // nlist x dc
  std::vector<int64_t> cnorm;      // nlist
  std::vector<uint32_t> id_row;    // npts -> row or UINT32_MAX
  std::vector<float> cen;          // nlist x d float centroids, unrotated (per-query coarse search)
  bool flat = false;               // LoadOptions::batched_coarse = false
  // Precision::kBf16: unrotated vectors as bf16 rows dp = roundup(d, 32) apart, and their norms
  uint32_t dp = 0;
  std::vector<uint16_t> bf;
  std::vector<float> bfn;
  bool bf16 = false;
};

// ---- bf16 precision (index.cpp:752-775, 1221-1293; kernels.cpp:318-419)
// to_bf16 (kernels.cpp:318-334): round to nearest even, as
// _mm512_cvtneps_pbh and the scalar tail do for finite values.
inline uint16_t bf16_of(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
inline float f32_of_bf16(uint16_t b) {
  const uint32_t u = static_cast<uint32_t>(b) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
// bf16_norm (kernels.cpp:336-340): s += v * v in order; the GCC build fuses
// each step (scalar FMAs, read from the assembly of the reference built here).
float bf16_norm_o(const uint16_t *v, uint32_t d) {
  float s = 0.0f;
  for (uint32_t k = 0; k < d; ++k) {
    const float x = f32_of_bf16(v[k]);
    s = std::fmaf(x, x, s);
  }
  return s;
}
// bf16_l2, AVX-512 path (kernels.cpp:384-414): eight fp32 accumulators,
// pair k2 (elements 2 k2, 2 k2 + 1) into accumulator k2 % 8 by
// _mm512_dpbf16_ps: the odd element's product added first, then the even
// one's, each product exact (8-bit significands) and each sum rounded; then
// acc[g] += acc[g + s] for s = 4, 2, 1; l2 = (qn + xn) - (dot + dot).
float bf16_dot(const uint16_t *q, const uint16_t *x, uint32_t dp) {
  float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  for (uint32_t k2 = 0; k2 < dp / 2; ++k2) {
    float &a = acc[k2 & 7];
    const float p1 = f32_of_bf16(x[2 * k2 + 1]) * f32_of_bf16(q[2 * k2 + 1]);
    const float p0 = f32_of_bf16(x[2 * k2]) * f32_of_bf16(q[2 * k2]);
    a = a + p1;
    a = a + p0;
  }
  for (int st = 4; st > 0; st >>= 1)
    for (int g = 0; g < st; ++g) acc[g] = acc[g] + acc[g + st];
  return acc[0];
}
// Order-preserving bits of a float (ties of -0 and +0 as +0), for (float, id) keys
// of distances that may be negative.
inline uint32_t ord_bits(float f) {
  if (f == 0.0f) f = 0.0f;
  uint32_t b;
  std::memcpy(&b, &f, 4);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
inline float from_ord(uint32_t o) {
  const uint32_t b = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}

// Reference: rotate_rows / rotate_tile4x4 / rotate_rows_serial
// (index.cpp:139-270). For basis rows j < d & ~3: sixteen fmaf chains, chain l
// over k = l, l+16, ... (k < d & ~15), summed by the GCC 11 _mm512_reduce_add_ps
// tree ((a0+a8)+(a4+a12))+((a2+a10)+(a6+a14)) + ((a1+a9)+(a5+a13))+((a3+a11)+(a7+a15)),
// then an fmaf tail over k >= d & ~15 (the reference's `s += x*r` tail is
// contracted to FMA). For j >= d & ~3, see below.
float rotate_one(const float *x, const float *r, uint32_t d, uint32_t j) {
  const uint32_t d4 = d & ~3u;
  if (j >= d4) {
    // The reference's plain loop `s += x[k] * r[k]` as GCC 11 -O3 compiles it
    // for AVX-512 (index.cpp:171-178, 247-258): products of k < d & ~7 are
    // vectorised (rounded products, added in k order), the last d % 8 terms
    // are scalar FMAs. Read from the assembly of the reference built here and
    // pinned against its rotated vectors.
    const uint32_t d8 = d & ~7u;
    float s = 0.0f;
    uint32_t k = 0;
    for (; k < d8; ++k) {
      const float p = x[k] * r[k];
      s = s + p;
    }
    for (; k < d; ++k) s = std::fmaf(x[k], r[k], s);
    return s;
  }
  float a[16];
  for (int l = 0; l < 16; ++l) a[l] = 0.0f;
  uint32_t k = 0;
  for (; k + 16 <= d; k += 16)
    for (int l = 0; l < 16; ++l) a[l] = std::fmaf(x[k + l], r[k + l], a[l]);
  float t3[8], t6[4];
  for (int i = 0; i < 8; ++i) t3[i] = a[i + 8] + a[i];
  for (int i = 0; i < 4; ++i) t6[i] = t3[i + 4] + t3[i];
  const float t80 = t6[0] + t6[2], t81 = t6[1] + t6[3];
  float s = t80 + t81;
  for (; k < d; ++k) s = std::fmaf(x[k], r[k], s);
  return s;
}

// Reference: l2_fp32 (index.cpp:36-56), the per-query coarse distance
// (index.cpp:832-844), as GCC 11 -O3 compiles it for AVX-512 (read from the
// assembly of the reference built here, pinned against its coarse_flat_ order):
// sixteen fmaf chains of (a - b)^2, chain l over k = l, l+16, ... (k < d & ~15),
// summed by the same _mm512_reduce_add_ps tree as the rotation; the tail
// k >= d & ~15 is a separate sum from 0: when it has 8 or more terms, its first
// 8 are rounded products added in k order (vectorised, no contraction), the
// rest FMAs; then tree + tail.
float l2_flat(const float *a, const float *b, uint32_t d) {
  float acc[16];
  for (int l = 0; l < 16; ++l) acc[l] = 0.0f;
  uint32_t k = 0;
  for (; k + 16 <= d; k += 16)
    for (int l = 0; l < 16; ++l) {
      const float dd = a[k + l] - b[k + l];
      acc[l] = std::fmaf(dd, dd, acc[l]);
    }
  float tail = 0.0f;
  if (d - k >= 8) {
    for (uint32_t e = k + 8; k < e; ++k) {
      const float dd = a[k] - b[k];
      const float p = dd * dd;
      tail = tail + p;
    }
  }
  for (; k < d; ++k) {
    const float dd = a[k] - b[k];
    tail = std::fmaf(dd, dd, tail);
  }
  float t3[8], t6[4];
  for (int i = 0; i < 8; ++i) t3[i] = acc[i + 8] + acc[i];
  for (int i = 0; i < 4; ++i) t6[i] = t3[i + 4] + t3[i];
  const float t80 = t6[0] + t6[2], t81 = t6[1] + t6[3];
  const float tree = t80 + t81;
  return tree + tail;
}

// Each element is rotate_one of its row and basis row; rows go in blocks of 8
// per basis row (the basis row stays in cache), every element computed alike.
void rotate_rows(const float *in, float *out, size_t n, uint32_t d, const float *rot) {
  const int64_t nb = static_cast<int64_t>((n + 7) / 8);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t b = 0; b < nb; ++b) {
    const size_t i0 = static_cast<size_t>(b) * 8, i1 = std::min(n, i0 + 8);
    for (uint32_t j = 0; j < d; ++j) {
      const float *r = rot + static_cast<size_t>(j) * d;
      for (size_t i = i0; i < i1; ++i) out[i * d + j] = rotate_one(in + i * d, r, d, j);
    }
  }
}

// Norms of unsigned codes (index.cpp:668-675, 700-711, 929-941).
void code_norms(const uint16_t *c, uint32_t d, uint32_t dc, uint32_t P, int64_t *n,
                int64_t *hn, int64_t *h0n) {
  int64_t a = 0, b = 0, e = 0;
  for (uint32_t k = 0; k < dc; ++k) {
    const int64_t v = c[k], h = c[k] >> 8;
    a += v * v;
    b += h * h;
    if (k < P) e += h * h;
  }
  (void) d;
  *n = a;
  *hn = b;
  if (h0n) *h0n = e;
}

void finish_index(OIndex &ix) {
  const uint64_t N = ix.list_off[ix.nlist];
  ix.norm.assign(N, 0);
  ix.hinorm.assign(N, 0);
  ix.headnorm.assign(N, 0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < static_cast<int64_t>(N); ++r)
    code_norms(ix.codes.data() + static_cast<size_t>(r) * ix.dc, ix.d, ix.dc, ix.P,
               &ix.norm[r], &ix.hinorm[r], &ix.headnorm[r]);
  ix.cnorm.assign(ix.nlist, 0);
  for (uint32_t c = 0; c < ix.nlist; ++c) {
    int64_t hn = 0;
    code_norms(ix.ccodes.data() + static_cast<size_t>(c) * ix.dc, ix.d, ix.dc, 0, &ix.cnorm[c], &hn,
               nullptr);
  }
  // id -> row, the later occurrence winning like a serial pass (index.cpp:795-806)
  ix.id_row.assign(ix.npts, UINT32_MAX);
  for (uint64_t r = 0; r < N; ++r)
    if (ix.ids[r] < ix.npts) ix.id_row[ix.ids[r]] = static_cast<uint32_t>(r);
}

using Key = std::pair<float, uint32_t>;

struct SearchOut {
  uint32_t *ids;
  float *dists;
  uint32_t *probes;     // nq x nprobe or null
  float *probe_dists;   // nq x nprobe or null
  uint64_t *surv;       // nq x S keys (float bits << 32 | id) ascending, or null
  uint32_t *nsurv;      // nq or null
  uint8_t *flags;       // nq or null: 1 coarse tie, 2 survivor tie, 4 final tie
};

inline uint64_t pack_key(float f, uint32_t id) {
  uint32_t b;
  std::memcpy(&b, &f, 4);
  return (static_cast<uint64_t>(b) << 32) | id;
}

}  // namespace

extern "C" {

void oracle_rotate_rows(const float *in, float *out, uint64_t n, uint32_t d, const float *rot) {
  rotate_rows(in, out, n, d, rot);
}

// Scalar int64 formula of the reference kernel test (tests/test_kernels.cpp:66-81):
// row-major planes, m x n outputs.
void oracle_sq16_l2(const uint8_t *qh, const uint8_t *ql, const uint8_t *xh, const uint8_t *xl,
                    const int64_t *qn, const int64_t *xn, double step_sq, uint32_t m, uint32_t n,
                    uint32_t dc, float *out) {
  for (uint32_t i = 0; i < m; ++i)
    for (uint32_t j = 0; j < n; ++j) {
      int64_t dot = 0;
      for (uint32_t k = 0; k < dc; ++k) {
        const int64_t a = static_cast<int64_t>(qh[static_cast<size_t>(i) * dc + k]) * 256 +
                          ql[static_cast<size_t>(i) * dc + k];
        const int64_t b = static_cast<int64_t>(xh[static_cast<size_t>(j) * dc + k]) * 256 +
                          xl[static_cast<size_t>(j) * dc + k];
        dot += a * b;
      }
      out[static_cast<size_t>(i) * n + j] = fscale(qn[i] + xn[j] - 2 * dot, step_sq);
    }
}

// High-byte distance over the leading pd dims of rows that are plane_dim wide
// (tests/test_kernels.cpp:84-108; kernels.cpp:296-316).
void oracle_sq16_high_l2(const uint8_t *qh, uint32_t qstride, const uint8_t *xh,
                         uint32_t plane_dim, const int64_t *qn, const int64_t *xn,
                         double step_sq, uint32_t m, uint32_t n, uint32_t pd, float *out) {
  for (uint32_t i = 0; i < m; ++i)
    for (uint32_t j = 0; j < n; ++j) {
      int64_t dot = 0;
      for (uint32_t k = 0; k < pd; ++k)
        dot += static_cast<int64_t>(qh[static_cast<size_t>(i) * qstride + k]) *
               xh[static_cast<size_t>(j) * plane_dim + k];
      out[static_cast<size_t>(i) * n + j] = fscale(qn[i] + xn[j] - 2 * dot, step_sq);
    }
}

// Load restated: rotate_lists_ (index.cpp:613-629), quantize_sq16_ (632-689),
// build_head_ (694-715), build_coarse_planes_ (718-750), id maps (795-806).
// vecs: N x d floats in list order (list c holds sizes[c] rows); ids likewise.
void *oracle_index_from_raw(uint32_t d, uint32_t nlist, uint32_t npts, const uint32_t *sizes,
                            const uint32_t *ids, const float *vecs, const float *centroids,
                            const float *rotation, uint32_t prefix_dims) {
  auto *ix = new OIndex();
  ix->d = d;
  ix->dc = (d + 63) / 64 * 64;
  ix->nlist = nlist;
  ix->npts = npts;
  ix->P = prefix_dims;
  ix->list_off.assign(nlist + 1, 0);
  for (uint32_t c = 0; c < nlist; ++c) ix->list_off[c + 1] = ix->list_off[c] + sizes[c];
  const uint64_t N = ix->list_off[nlist];
  ix->ids.assign(ids, ids + N);
  ix->cen.assign(centroids, centroids + static_cast<size_t>(nlist) * d);
  ix->dp = (d + 31) / 32 * 32;
  ix->bf.assign(static_cast<size_t>(N) * ix->dp, 0);
  ix->bfn.assign(N, 0.0f);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < static_cast<int64_t>(N); ++r) {
    uint16_t *o = ix->bf.data() + static_cast<size_t>(r) * ix->dp;
    for (uint32_t k = 0; k < d; ++k) o[k] = bf16_of(vecs[static_cast<size_t>(r) * d + k]);
    ix->bfn[r] = bf16_norm_o(o, d);
  }
  std::vector<float> rot(static_cast<size_t>(N) * d);
  std::vector<float> crot(static_cast<size_t>(nlist) * d);
  if (rotation) {
    ix->rotated = true;
    ix->rotation.assign(rotation, rotation + static_cast<size_t>(d) * d);
    rotate_rows(vecs, rot.data(), N, d, rotation);
    rotate_rows(centroids, crot.data(), nlist, d, rotation);
  } else {
    std::memcpy(rot.data(), vecs, rot.size() * sizeof(float));
    std::memcpy(crot.data(), centroids, crot.size() * sizeof(float));
  }
  float glo = std::numeric_limits<float>::max(), ghi = -std::numeric_limits<float>::max();
  for (float v : rot) {
    glo = std::min(glo, v);
    ghi = std::max(ghi, v);
  }
  ix->lo = glo;
  const double rng = static_cast<double>(ghi) - glo;
  ix->inv_step = 65535.0 / rng;
  const double step = rng / 65535.0;
  ix->step_sq = step * step;
  ix->codes.assign(static_cast<size_t>(N) * ix->dc, 0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < static_cast<int64_t>(N); ++r)
    for (uint32_t k = 0; k < d; ++k)
      ix->codes[static_cast<size_t>(r) * ix->dc + k] =
          static_cast<uint16_t>(quant(rot[static_cast<size_t>(r) * d + k], ix->lo, ix->inv_step));
  ix->ccodes.assign(static_cast<size_t>(nlist) * ix->dc, 0);
  for (uint32_t c = 0; c < nlist; ++c)
    for (uint32_t k = 0; k < d; ++k)
      ix->ccodes[static_cast<size_t>(c) * ix->dc + k] =
          static_cast<uint16_t>(quant(crot[static_cast<size_t>(c) * d + k], ix->lo, ix->inv_step));
  finish_index(*ix);
  return ix;
}

// An index from codes computed elsewhere (the GPU load path), for parity checks
// of the search at sizes where the CPU load is too slow. Codes are unsigned.
void *oracle_index_from_codes(uint32_t d, uint32_t nlist, uint32_t npts, const uint32_t *sizes,
                              const uint32_t *ids, const uint16_t *codes, const uint16_t *ccodes,
                              const float *rotation, float lo, double inv_step, double step_sq,
                              uint32_t prefix_dims) {
  auto *ix = new OIndex();
  ix->d = d;
  ix->dc = (d + 63) / 64 * 64;
  ix->nlist = nlist;
  ix->npts = npts;
  ix->P = prefix_dims;
  ix->list_off.assign(nlist + 1, 0);
  for (uint32_t c = 0; c < nlist; ++c) ix->list_off[c + 1] = ix->list_off[c] + sizes[c];
  const uint64_t N = ix->list_off[nlist];
  ix->ids.assign(ids, ids + N);
  if (rotation) {
    ix->rotated = true;
    ix->rotation.assign(rotation, rotation + static_cast<size_t>(d) * d);
  }
  ix->lo = lo;
  ix->inv_step = inv_step;
  ix->step_sq = step_sq;
  ix->codes.assign(codes, codes + static_cast<size_t>(N) * ix->dc);
  ix->ccodes.assign(ccodes, ccodes + static_cast<size_t>(nlist) * ix->dc);
  finish_index(*ix);
  return ix;
}

void oracle_index_free(void *h) { delete static_cast<OIndex *>(h); }

// Coarse search per query in fp32 (LoadOptions::batched_coarse = false);
// needs the float centroids (an index from raw vectors). Returns 0 on success.
int oracle_index_set_flat(void *h, int flat) {
  auto *ix = static_cast<OIndex *>(h);
  if (flat && ix->cen.empty()) return 1;
  ix->flat = flat != 0;
  return 0;
}

// bf16 precision (LoadOptions::precision = kBf16): full ranking over the bf16
// rows, the per-query coarse search. Returns 0 on success.
int oracle_index_set_bf16(void *h, int on) {
  auto *ix = static_cast<OIndex *>(h);
  if (on && ix->bf.empty()) return 1;
  ix->bf16 = on != 0;
  if (on) ix->flat = true;
  return 0;
}

// The per-query coarse distances of every (query, list), nq x nlist.
int oracle_flat_dists(const void *h, const float *queries, uint32_t nq, float *out) {
  const auto *ix = static_cast<const OIndex *>(h);
  if (ix->cen.empty()) return 1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(nq); ++i)
    for (uint32_t c = 0; c < ix->nlist; ++c)
      out[static_cast<size_t>(i) * ix->nlist + c] =
          l2_flat(queries + static_cast<size_t>(i) * ix->d, ix->cen.data() + static_cast<size_t>(c) * ix->d, ix->d);
  return 0;
}

void oracle_index_scale(const void *h, float *lo, double *inv_step, double *step_sq) {
  const auto *ix = static_cast<const OIndex *>(h);
  *lo = ix->lo;
  *inv_step = ix->inv_step;
  *step_sq = ix->step_sq;
}

uint64_t oracle_index_rows(const void *h) {
  const auto *ix = static_cast<const OIndex *>(h);
  return ix->list_off[ix->nlist];
}

// Codes (N x dc u16), centroid codes, norms: for comparisons with the GPU load.
void oracle_index_export(const void *h, uint16_t *codes, uint16_t *ccodes, int64_t *norm,
                         int64_t *headnorm, int64_t *cnorm) {
  const auto *ix = static_cast<const OIndex *>(h);
  if (codes) std::memcpy(codes, ix->codes.data(), ix->codes.size() * 2);
  if (ccodes) std::memcpy(ccodes, ix->ccodes.data(), ix->ccodes.size() * 2);
  if (norm) std::memcpy(norm, ix->norm.data(), ix->norm.size() * 8);
  if (headnorm) std::memcpy(headnorm, ix->headnorm.data(), ix->headnorm.size() * 8);
  if (cnorm) std::memcpy(cnorm, ix->cnorm.data(), ix->cnorm.size() * 8);
}

// Query rotation + quantisation (index.cpp:895-898, 917-943).
void oracle_queries(const void *h, const float *queries, uint32_t nq, float *qrot,
                    uint16_t *qcodes) {
  const auto *ix = static_cast<const OIndex *>(h);
  const uint32_t d = ix->d, dc = ix->dc;
  std::vector<float> tmp(static_cast<size_t>(nq) * d);
  if (ix->rotated) rotate_rows(queries, tmp.data(), nq, d, ix->rotation.data());
  else std::memcpy(tmp.data(), queries, tmp.size() * 4);
  if (qrot) std::memcpy(qrot, tmp.data(), tmp.size() * 4);
  if (qcodes)
    for (uint32_t i = 0; i < nq; ++i) {
      for (uint32_t k = 0; k < dc; ++k)
        qcodes[static_cast<size_t>(i) * dc + k] =
            k < d ? static_cast<uint16_t>(quant(tmp[static_cast<size_t>(i) * d + k], ix->lo, ix->inv_step))
                  : 0;
    }
}

// The search (index.cpp:879-904, 909-1128, 1131-1219) with canonical ties:
//   coarse     the nprobe smallest (float, list) keys, ascending; with
//              libstdc_ties != 0 the reference's own nth_element + sort over
//              iota by float (index.cpp:862-872), which reproduces its choice
//              among equal floats when built with the same libstdc++.
//   survivors  the S smallest (float, id) ranking keys (index.cpp:1066-1069).
//   final      survivors popped in descending (float, id) order, exact
//              distance, K max-heap admitting a new pair iff its float is
//              strictly below the top's (index.cpp:1081-1124), then ascending
//              (float, id) with FLT_MAX / UINT32_MAX padding (index.cpp:384-398).
// ranking: 0 full (exact top-K of all candidates), 1 high bytes, 2 prefix.
// nprobe_per_query: optional per-query probe counts (<= nprobe).
int oracle_search(const void *h, const float *queries, uint32_t nq, uint32_t k, uint32_t nprobe,
                  uint32_t survivors, int ranking, int libstdc_ties,
                  const uint32_t *nprobe_per_query, uint32_t *out_ids, float *out_dists,
                  uint32_t *out_probes, float *out_probe_dists, uint64_t *out_surv,
                  uint32_t *out_nsurv, uint8_t *out_flags) {
  const auto *ix = static_cast<const OIndex *>(h);
  if (nq == 0) return 0;
  if (k == 0 || nprobe == 0 || nprobe > ix->nlist) return 1;
  const uint32_t d = ix->d, dc = ix->dc, nlist = ix->nlist;
  const uint32_t S = std::max(survivors, k);
  const uint32_t P = ranking == 2 ? ix->P : dc;
  if (ranking == 2 && P == 0) return 1;
  if (ix->bf16 && ranking != 0) return 1;  // bf16 precision supports only Ranking::kFull
  std::vector<uint16_t> qc(static_cast<size_t>(nq) * dc);
  oracle_queries(h, queries, nq, nullptr, qc.data());
  SearchOut so{out_ids, out_dists, out_probes, out_probe_dists, out_surv, out_nsurv, out_flags};

#pragma omp parallel
  {
    std::vector<float> cd(nlist);
    std::vector<uint32_t> ord(nlist);
    std::vector<std::pair<uint64_t, uint32_t>> cand;  // (key, row)
#pragma omp for schedule(dynamic, 4)
    for (int64_t qi = 0; qi < static_cast<int64_t>(nq); ++qi) {
      const uint16_t *q = qc.data() + static_cast<size_t>(qi) * dc;
      int64_t qn, qhn, qh0n;
      code_norms(q, d, dc, ix->P, &qn, &qhn, &qh0n);
      uint8_t flg = 0;
      // coarse (index.cpp:846-877, kernels.cpp:166-168 with M = 16), or per
      // query in fp32 (index.cpp:832-844): its max-heap with strict admission
      // over lists in index order keeps exactly the (float, list) order, ties
      // included
      if (ix->flat)
        for (uint32_t c = 0; c < nlist; ++c)
          cd[c] = l2_flat(queries + static_cast<size_t>(qi) * d, ix->cen.data() + static_cast<size_t>(c) * d, d);
      for (uint32_t c = 0; c < nlist && !ix->flat; ++c) {
        const uint16_t *x = ix->ccodes.data() + static_cast<size_t>(c) * dc;
        int64_t dot = 0;
        for (uint32_t t = 0; t < dc; ++t) dot += static_cast<int64_t>(q[t]) * x[t];
        cd[c] = fscale(qn + ix->cnorm[c] - 2 * dot, ix->step_sq);
      }
      const uint32_t npq = nprobe_per_query ? std::min(nprobe_per_query[qi], nprobe) : nprobe;
      std::iota(ord.begin(), ord.end(), 0u);
      if (libstdc_ties && !ix->flat) {
        auto lt = [&](uint32_t a, uint32_t b) { return cd[a] < cd[b]; };
        if (npq < nlist) std::nth_element(ord.begin(), ord.begin() + npq, ord.end(), lt);
        std::sort(ord.begin(), ord.begin() + npq, lt);
      } else {
        auto lt = [&](uint32_t a, uint32_t b) { return cd[a] < cd[b] || (cd[a] == cd[b] && a < b); };
        std::partial_sort(ord.begin(), ord.begin() + std::min(npq + (npq < nlist ? 1u : 0u), nlist),
                          ord.end(), lt);
      }
      if (npq < nlist) {
        float kth = -1.0f, nxt = -1.0f;
        if (!libstdc_ties) {
          kth = cd[ord[npq - 1]];
          nxt = cd[ord[npq]];
        } else {
          kth = cd[ord[npq - 1]];
          nxt = std::numeric_limits<float>::max();
          for (uint32_t p = npq; p < nlist; ++p) nxt = std::min(nxt, cd[ord[p]]);
        }
        if (kth == nxt && !ix->flat) flg |= 1;
      }
      if (so.probes)
        for (uint32_t p = 0; p < nprobe; ++p)
          so.probes[static_cast<size_t>(qi) * nprobe + p] = p < npq ? ord[p] : UINT32_MAX;
      if (so.probe_dists)
        for (uint32_t p = 0; p < nprobe; ++p)
          so.probe_dists[static_cast<size_t>(qi) * nprobe + p] =
              p < npq ? cd[ord[p]] : std::numeric_limits<float>::max();

      // candidates of every probed list (index.cpp:998-1044)
      cand.clear();
      std::vector<uint16_t> qb;
      float qbn = 0.0f;
      if (ix->bf16) {
        qb.assign(ix->dp, 0);
        for (uint32_t t = 0; t < d; ++t) qb[t] = bf16_of(queries[static_cast<size_t>(qi) * d + t]);
        qbn = bf16_norm_o(qb.data(), d);
      }
      for (uint32_t p = 0; p < npq; ++p) {
        const uint32_t c = ord[p];
        for (uint64_t r = ix->list_off[c]; r < ix->list_off[c + 1]; ++r) {
          if (ix->bf16) {
            const float dot = bf16_dot(qb.data(), ix->bf.data() + static_cast<size_t>(r) * ix->dp, ix->dp);
            const float l2f = (qbn + ix->bfn[r]) - (dot + dot);
            cand.emplace_back((static_cast<uint64_t>(ord_bits(l2f)) << 32) | ix->ids[r], static_cast<uint32_t>(r));
            continue;
          }
          const uint16_t *x = ix->codes.data() + static_cast<size_t>(r) * dc;
          int64_t l2;
          if (ranking == 0) {
            int64_t dot = 0;
            for (uint32_t t = 0; t < dc; ++t) dot += static_cast<int64_t>(q[t]) * x[t];
            l2 = qn + ix->norm[r] - 2 * dot;
          } else {
            int64_t dot = 0;
            for (uint32_t t = 0; t < P; ++t)
              dot += static_cast<int64_t>(q[t] >> 8) * static_cast<int64_t>(x[t] >> 8);
            l2 = (ranking == 2 ? qh0n + ix->headnorm[r] : qhn + ix->hinorm[r]) - 2 * dot;
          }
          cand.emplace_back(pack_key(fscale(l2, ix->step_sq), ix->ids[r]), static_cast<uint32_t>(r));
        }
      }
      uint32_t *oid = so.ids + static_cast<size_t>(qi) * k;
      float *odi = so.dists + static_cast<size_t>(qi) * k;
      std::vector<std::pair<float, uint32_t>> fin;
      if (ranking == 0) {
        // kFull: the K smallest (float, id) of all candidates (index.cpp:1131-1219).
        const size_t m = std::min<size_t>(k, cand.size());
        std::partial_sort(cand.begin(), cand.begin() + std::min(m + 1, cand.size()), cand.end());
        if (m > 0 && m < cand.size() && (cand[m - 1].first >> 32) == (cand[m].first >> 32)) flg |= 4;
        for (size_t t = 0; t < m; ++t) {
          float f;
          const uint32_t b = static_cast<uint32_t>(cand[t].first >> 32);
          if (ix->bf16) f = from_ord(b);
          else std::memcpy(&f, &b, 4);
          fin.emplace_back(f, static_cast<uint32_t>(cand[t].first));
        }
        if (so.nsurv) so.nsurv[qi] = 0;
      } else {
        const size_t ns = std::min<size_t>(S, cand.size());
        std::partial_sort(cand.begin(), cand.begin() + std::min(ns + 1, cand.size()), cand.end());
        if (ns > 0 && ns < cand.size() && (cand[ns - 1].first >> 32) == (cand[ns].first >> 32))
          flg |= 2;
        if (so.nsurv) so.nsurv[qi] = static_cast<uint32_t>(ns);
        if (so.surv)
          for (uint32_t t = 0; t < S; ++t)
            so.surv[static_cast<size_t>(qi) * S + t] = t < ns ? cand[t].first : ~0ull;
        // refinement in arrival order: descending (float, id)
        std::vector<std::pair<float, uint32_t>> heap;  // max-heap by pair
        std::vector<float> ex;
        for (size_t t = ns; t-- > 0;) {
          const uint32_t r = cand[t].second;
          const uint16_t *x = ix->codes.data() + static_cast<size_t>(r) * dc;
          int64_t dot = 0;
          for (uint32_t u = 0; u < dc; ++u) dot += static_cast<int64_t>(q[u]) * x[u];
          const float e = fscale(qn + ix->norm[r] - 2 * dot, ix->step_sq);
          ex.push_back(e);
          const std::pair<float, uint32_t> kv(e, ix->ids[r]);
          if (heap.size() < k) {
            heap.push_back(kv);
            std::push_heap(heap.begin(), heap.end());
          } else if (e < heap.front().first) {
            std::pop_heap(heap.begin(), heap.end());
            heap.back() = kv;
            std::push_heap(heap.begin(), heap.end());
          }
        }
        if (ex.size() > k) {
          std::vector<float> sorted = ex;
          std::nth_element(sorted.begin(), sorted.begin() + k, sorted.end());
          const float kp1 = sorted[k];
          const float kth = *std::max_element(sorted.begin(), sorted.begin() + k);
          if (kth == kp1) flg |= 4;
        }
        fin = heap;
      }
      std::sort(fin.begin(), fin.end());
      for (uint32_t t = 0; t < k; ++t) {
        if (t < fin.size()) {
          odi[t] = fin[t].first;
          oid[t] = fin[t].second;
        } else {
          odi[t] = std::numeric_limits<float>::max();
          oid[t] = UINT32_MAX;
        }
      }
      if (so.flags) so.flags[qi] = flg;
    }
  }
  return 0;
}

// Query normalisation of the Python module and CLI (python/katoann_module.cpp:97-110,
// tools/katoann.cpp:102-111): s = sum of x*x in double (one fma chain), then
// x *= float(1 / sqrt(s)) unless s == 0.
void oracle_normalize_rows(float *v, uint64_t n, uint32_t d) {
  for (uint64_t i = 0; i < n; ++i) {
    float *x = v + i * d;
    double s = 0;
    for (uint32_t k = 0; k < d; ++k) s = std::fma(static_cast<double>(x[k]), static_cast<double>(x[k]), s);
    if (s > 0) {
      const float inv = static_cast<float>(1.0 / std::sqrt(s));
      for (uint32_t k = 0; k < d; ++k) x[k] *= inv;
    }
  }
}

}  // extern "C"
