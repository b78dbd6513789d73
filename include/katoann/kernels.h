// Kernel-level API of the reference (include/katoann/kernels.h of KatoANN) for
// the gfx950 engine. The distance kernels run on the GPU (v_mfma_i32_16x16x64_i8)
// and take row-major planes: the reference's packed AMX tile layout
// (pack_plane) is a CPU layout and not part of this engine.
//
//   reference kernels.h:19     Engine          (+ kGfx950)
//   reference kernels.h:22     amx_available   (false: the engine is the GPU)
//   reference kernels.h:24     kMaxDim
//   reference kernels.h:34-40  sq16_l2         -> katoann_sq16_l2
//   reference kernels.h:42-52  sq16_high_l2    -> katoann_sq16_high_l2
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "katoann_c.h"

namespace katoann {

enum class Engine { kAmx, kAvx512, kGfx950 };

inline bool amx_available() { return false; }

constexpr uint32_t kMaxDim = 4096;

// Bytes of one row-major plane of n vectors padded to padded_dim dimensions.
inline size_t plane_bytes(uint32_t n, uint32_t padded_dim) { return static_cast<size_t>(n) * padded_dim; }

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

}  // namespace katoann
