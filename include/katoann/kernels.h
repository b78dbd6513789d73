// Kernel-level API of the reference (include/katoann/kernels.h of KatoANN) for
// the gfx950 engine. The distance kernels run on the GPU (v_mfma_i32_16x16x64_i8)
// and take row-major planes: the reference's packed AMX tile layout
// (pack_plane) is a CPU layout and not part of this engine.
//
//   reference kernels.h:19     Engine          (+ kGfx950)
//   reference kernels.h:22     amx_available   (false: the engine is the GPU)
//   reference kernels.h:24     kMaxDim
//   reference kernels.h:26-31  plane_bytes, pack_plane (row-major planes here)
//   reference kernels.h:34-40  sq16_l2         -> katoann_sq16_l2
//   reference kernels.h:42-52  sq16_high_l2    -> katoann_sq16_high_l2
//   reference kernels.h:54-67  bf16 helpers, bf16_l2 -> katoann_bf16_l2
// Programs written against the reference's packed layouts keep working: they
// treat a packed plane as opaque, and pack_plane / pack_bf16 here produce the
// row-major planes the GPU kernels read.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "katoann_c.h"

namespace katoann {

enum class Engine { kAmx, kAvx512, kGfx950 };

inline bool amx_available() { return false; }

constexpr uint32_t kMaxDim = 4096;

// Bytes of one row-major plane of n vectors padded to padded_dim dimensions.
inline size_t plane_bytes(uint32_t n, uint32_t padded_dim) { return static_cast<size_t>(n) * padded_dim; }

// rows: n x padded_dim bytes, row-major; the plane this engine reads is the
// same layout.
inline void pack_plane(const uint8_t *rows, uint32_t n, uint32_t padded_dim, uint8_t *out) {
  std::memcpy(out, rows, plane_bytes(n, padded_dim));
}

namespace detail {
inline void check_kernel(int rc) {
  if (rc == KATOANN_OK) return;
  if (rc == KATOANN_INVALID_ARGUMENT) throw std::invalid_argument(katoann_last_error());
  throw std::runtime_error(katoann_last_error());
}
}  // namespace detail

// Squared L2 of m queries against n vectors from the full codes (row-major
// hi / lo planes, m x padded_dim and n x padded_dim); norms of the unsigned
// codes; out is m x n, scaled by step_sq.
inline void sq16_l2(Engine, const uint8_t *qhi, const uint8_t *qlo, const uint8_t *xhi, const uint8_t *xlo,
                    const int64_t *qnorm, const int64_t *xnorm, double step_sq, uint32_t m, uint32_t n,
                    uint32_t padded_dim, float *out) {
  detail::check_kernel(katoann_sq16_l2(qhi, qlo, xhi, xlo, qnorm, xnorm, step_sq, m, n, padded_dim, out));
}

// Same from the high bytes of the leading padded_dim dimensions of rows that
// are plane_dim wide (0 means padded_dim).
inline void sq16_high_l2(Engine, const uint8_t *qhi, const uint8_t *xhi, const int64_t *qnorm,
                         const int64_t *xnorm, double step_sq, uint32_t m, uint32_t n, uint32_t padded_dim,
                         float *out, uint32_t plane_dim = 0) {
  detail::check_kernel(
      katoann_sq16_high_l2(qhi, xhi, qnorm, xnorm, step_sq, m, n, padded_dim, out, plane_dim));
}

// bf16 vectors (reference kernels.h:54-67): rows padded to a multiple of 32
// values; conversion rounds to nearest even; norms are in-order FMA sums.
inline size_t bf16_packed_elems(uint32_t n, uint32_t dim) {
  return static_cast<size_t>(n) * ((dim + 31u) / 32u) * 32u;
}
inline void to_bf16(const float *src, uint16_t *dst, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    uint32_t u;
    std::memcpy(&u, &src[i], 4);
    dst[i] = static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
  }
}
inline float bf16_norm(const uint16_t *v, uint32_t dim) {
  float s = 0.0f;
  for (uint32_t i = 0; i < dim; ++i) {
    const uint32_t u = static_cast<uint32_t>(v[i]) << 16;
    float x;
    std::memcpy(&x, &u, 4);
    s = std::fma(x, x, s);
  }
  return s;
}
inline void pack_bf16(const uint16_t *rows, uint32_t n, uint32_t dim, uint16_t *out) {
  const uint32_t dp = (dim + 31u) / 32u * 32u;
  for (uint32_t r = 0; r < n; ++r)
    for (uint32_t k = 0; k < dp; ++k) out[static_cast<size_t>(r) * dp + k] = k < dim ? rows[static_cast<size_t>(r) * dim + k] : 0;
}

// Distances of m bf16 queries to n bf16 rows (row-major, dim values each; the
// output of pack_bf16 when dim is a multiple of 32), on the GPU in the
// reference's AVX-512 arithmetic; out is m x n.
inline void bf16_l2(Engine, const uint16_t *q, const uint16_t *x, const float *qnorm, const float *xnorm,
                    uint32_t m, uint32_t n, uint32_t dim, float *out) {
  detail::check_kernel(katoann_bf16_l2(q, x, qnorm, xnorm, m, n, dim, out));
}

}  // namespace katoann
