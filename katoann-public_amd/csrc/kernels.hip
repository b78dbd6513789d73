// gfx950 kernels of the batched IVF search. Host launchers at the bottom
// (declared in kernels.hpp). Compiled without fast-math, with FP contraction
// off: every fused multiply-add below is an explicit fmaf.
#include <float.h>

#include "device.hpp"
#include "kernels.hpp"

using namespace katoann_dev;

namespace {

// --------------------------------------------------------------------------
// Query normalisation of the reference front ends (python/katoann_module.cpp:
// 97-110, tools/katoann.cpp:102-111): s = sum x*x in double, one FMA chain in
// k order; x *= float(1 / sqrt(s)) unless s == 0.
__global__ void k_normalize(float *x, uint32_t n, uint32_t d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float *r = x + static_cast<size_t>(i) * d;
  double s = 0.0;
  for (uint32_t k = 0; k < d; ++k) {
    const double v = static_cast<double>(r[k]);
    s = fma(v, v, s);
  }
  if (s > 0.0) {
    const float inv = static_cast<float>(1.0 / sqrt(s));
    for (uint32_t k = 0; k < d; ++k) r[k] = r[k] * inv;
  }
}

// --------------------------------------------------------------------------
// Rotation y[i][j] = <x_i, R_j> in the reference's exact arithmetic
// (src/index.cpp:139-270), basis rows j < d & ~3: sixteen FMA chains, chain l
// over k = l, l+16, ... (k < d & ~15); the 16 partial sums added as the GCC
// _mm512_reduce_add_ps tree, which is the xor-8/4/2/1 butterfly below (float
// addition is commutative, so every lane of the butterfly ends with the same
// bits); then FMAs over k >= d & ~15 in order.
//
// The same kernel with L2 = true computes the per-query coarse distances
// l2_fp32(x_i, c_j) (src/index.cpp:36-56, 832-844) of queries and float
// centroids: chains of fmaf(x - c, x - c, acc), the same tree, and a tail
// k >= d & ~15 summed separately from 0 (as the reference's AVX-512 build
// evaluates it: when it has 8 or more terms, the first 8 are rounded products
// added in k order, the rest FMAs), then tree + tail.
//
// Tile: 64 rows x 32 basis rows per 512-thread block. Thread = (chain l,
// slot); the 4 slots of a wave share 8 input rows (LDS broadcast) and own 8
// basis rows each, so each thread keeps 8 x 8 chain accumulators. The next
// 64-wide chunk of k is loaded into registers while the current one is
// multiplied out of LDS (double buffer, one barrier per chunk).
constexpr int RT_ROWS = 64;
constexpr int RT_COLS = 32;
constexpr int RT_K = 64;
constexpr int RT_LD = RT_K + 2;  // 66 floats: the four slots of a wave hit disjoint banks

template <bool L2>
__global__ __launch_bounds__(512, 4) void k_rotate(const float *__restrict__ X, uint32_t n, uint32_t d,
                                                const float *__restrict__ R, uint32_t ncols,
                                                float *__restrict__ Y, uint32_t ldy) {
  __shared__ float xs[2][RT_ROWS][RT_LD];
  __shared__ float rs[2][RT_COLS][RT_LD];
  const int tid = threadIdx.x;
  const int l = tid & 15;
  const int slot = tid >> 4;
  const int rg = slot >> 2, cg = slot & 3;
  const uint32_t i0 = blockIdx.x * RT_ROWS, j0 = blockIdx.y * RT_COLS;  // rows on x: no 65535 limit
  const uint32_t d16 = d & ~15u;
  // loader: thread t moves 4 consecutive k of input rows t / 16 and 32 + t / 16
  // and of basis row t / 16 per chunk
  const int lrow = tid >> 4, lk = 4 * (tid & 15);
  const float *xrow0 = X + static_cast<size_t>(min(i0 + lrow, n - 1)) * d;
  const float *xrow1 = X + static_cast<size_t>(min(i0 + 32 + lrow, n - 1)) * d;
  const float *rrow = R + static_cast<size_t>(min(j0 + lrow, ncols - 1)) * d;
  const bool vec = (d & 3u) == 0;
  float4 p0, p1, pr;
  auto ld4 = [&](const float *r, uint32_t k) {
    return vec ? *reinterpret_cast<const float4 *>(r + k) : float4{r[k], r[k + 1], r[k + 2], r[k + 3]};
  };
  auto fetch = [&](uint32_t k0) {
    const uint32_t k = k0 + lk;  // k < d16 for all four or for none (d16 % 16 == 0)
    if (k >= d16) {
      p0 = p1 = pr = float4{0.0f, 0.0f, 0.0f, 0.0f};
    } else {
      p0 = ld4(xrow0, k);
      p1 = ld4(xrow1, k);
      pr = ld4(rrow, k);
    }
  };
  auto put4 = [&](float *dst, float4 v) {
    dst[0] = v.x;
    dst[1] = v.y;
    dst[2] = v.z;
    dst[3] = v.w;
  };
  auto stash = [&](int b) {
    put4(&xs[b][lrow][lk], p0);
    put4(&xs[b][32 + lrow][lk], p1);
    put4(&rs[b][lrow][lk], pr);
  };
  float acc[8][8];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[a][b] = 0.0f;

  if (d16) {
    fetch(0);
    stash(0);
  }
  __syncthreads();
  for (uint32_t k0 = 0; k0 < d16; k0 += RT_K) {
    const int buf = (k0 / RT_K) & 1;
    const bool more = k0 + RT_K < d16;
    if (more) fetch(k0 + RT_K);
#pragma unroll
    for (int m = 0; m < RT_K / 16; ++m) {
      if (k0 + 16 * m < d16) {  // uniform: d16 is a multiple of 16
        const int kk = 16 * m + l;
        float xv[8], rv[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) xv[a] = xs[buf][rg * 8 + a][kk];
#pragma unroll
        for (int b = 0; b < 8; ++b) rv[b] = rs[buf][cg * 8 + b][kk];
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
          for (int b = 0; b < 8; ++b) {
            if (L2) {
              const float dd = xv[a] - rv[b];
              acc[a][b] = fmaf(dd, dd, acc[a][b]);
            } else {
              acc[a][b] = fmaf(xv[a], rv[b], acc[a][b]);
            }
          }
      }
    }
    if (more) {
      stash(buf ^ 1);
      __syncthreads();
    }
  }
  // butterfly = reduce tree; every one of the 16 lanes ends with the sum
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      float v = acc[a][b];
      v = v + __shfl_xor(v, 8);
      v = v + __shfl_xor(v, 4);
      v = v + __shfl_xor(v, 2);
      v = v + __shfl_xor(v, 1);
      acc[a][b] = v;
    }
  // lane l writes outputs l, l+16, l+32, l+48 of the slot's 8 x 8, with the tail
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int o = l + 16 * t;
    const int a = o >> 3, b = o & 7;
    float s = 0.0f;
#pragma unroll
    for (int aa = 0; aa < 8; ++aa)
#pragma unroll
      for (int bb = 0; bb < 8; ++bb)
        if (aa == a && bb == b) s = acc[aa][bb];
    const uint32_t i = i0 + rg * 8 + a, j = j0 + cg * 8 + b;
    if (i < n && j < ncols) {
      const float *x = X + static_cast<size_t>(i) * d;
      const float *r = R + static_cast<size_t>(j) * d;
      if (L2) {
        float t = 0.0f;
        uint32_t k = d16;
        if (d - d16 >= 8) {
          for (; k < d16 + 8; ++k) {
            const float dd = x[k] - r[k];
            const float p = dd * dd;
            t = t + p;
          }
        }
        for (; k < d; ++k) {
          const float dd = x[k] - r[k];
          t = fmaf(dd, dd, t);
        }
        s = s + t;
      } else {
        for (uint32_t k = d16; k < d; ++k) s = fmaf(x[k], r[k], s);
      }
      Y[static_cast<size_t>(i) * ldy + j] = s;
    }
  }
}

// Basis rows j >= d & ~3 (only when d % 4 != 0): the reference's plain loop as
// its GCC -O3 AVX-512 build evaluates it (src/index.cpp:171-178, 247-258):
// rounded products added in k order for k < d & ~7, FMAs for the last d % 8.
__global__ void k_rotate_tail(const float *__restrict__ X, uint32_t n, uint32_t d,
                              const float *__restrict__ R, float *__restrict__ Y) {
  const uint32_t d4 = d & ~3u, nt = d - d4, d8 = d & ~7u;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nt) return;
  const uint32_t i = t / nt, j = d4 + t % nt;
  const float *x = X + static_cast<size_t>(i) * d;
  const float *r = R + static_cast<size_t>(j) * d;
  float s = 0.0f;
  uint32_t k = 0;
  for (; k < d8; ++k) {
    const float p = x[k] * r[k];
    s = s + p;
  }
  for (; k < d; ++k) s = fmaf(x[k], r[k], s);
  Y[static_cast<size_t>(i) * d + j] = s;
}

// --------------------------------------------------------------------------
// bf16 precision (src/index.cpp:752-775, 1230-1237; kernels.cpp:318-340): rows
// rounded to bf16 (nearest even; the reference's _mm512_cvtneps_pbh and its
// scalar tail agree on finite values), zero padded to dp = roundup(d, 32), and
// their norms s = fmaf(v, v, s) in element order (the reference's scalar loop,
// fused by its GCC build). One wave per row: lanes convert 64 elements at a
// time, then every lane runs the same in-order chain over them.
__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ X, uint32_t n, uint32_t d, uint32_t dp,
                                                 uint16_t *__restrict__ out, float *__restrict__ norms) {
  const uint32_t row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  const int lane = lane_id();
  float s = 0.0f;
  for (uint32_t k0 = 0; k0 < dp; k0 += 64) {
    const uint32_t k = k0 + lane;
    uint32_t b = 0;
    if (k < d) {
      const uint32_t u = __float_as_uint(X[static_cast<size_t>(row) * d + k]);
      b = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    }
    if (k < dp) out[static_cast<size_t>(row) * dp + k] = static_cast<uint16_t>(b);
    const float v = __uint_as_float(b << 16);
    const uint32_t m = min(64u, d > k0 ? d - k0 : 0u);
    for (uint32_t l = 0; l < m; ++l) {
      const float x = __shfl(v, static_cast<int>(l));
      s = fmaf(x, x, s);
    }
  }
  if (lane == 0) norms[row] = s;
}

// Kernel-level bf16 distances (kernels.cpp:384-414, row-major rows, dim a
// multiple of 2): one thread per (query, row), the eight accumulators over
// pairs k2 (k2 % 8), odd element first, then the 4/2/1 tree and
// (qn + xn) - (dot + dot), as the ranking kernel computes them.
__global__ void k_bf16_pairs(const uint32_t *__restrict__ q, const uint32_t *__restrict__ x,
                             const float *__restrict__ qn, const float *__restrict__ xn, uint32_t m, uint32_t n,
                             uint32_t d2, float *__restrict__ out) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= m * n) return;
  const uint32_t i = t / n, j = t % n;
  const uint32_t *qr = q + static_cast<size_t>(i) * d2, *xr = x + static_cast<size_t>(j) * d2;
  float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  for (uint32_t k2 = 0; k2 < d2; ++k2) {
    const uint32_t qw = qr[k2], xw = xr[k2];
    float &a = acc[k2 & 7];
    a = fmaf(__uint_as_float(xw & 0xffff0000u), __uint_as_float(qw & 0xffff0000u), a);
    a = fmaf(__uint_as_float(xw << 16), __uint_as_float(qw << 16), a);
  }
  for (int st = 4; st > 0; st >>= 1)
    for (int g = 0; g < st; ++g) acc[g] = acc[g] + acc[g + st];
  out[static_cast<size_t>(i) * n + j] = (qn[i] + xn[j]) - (acc[0] + acc[0]);
}

// --------------------------------------------------------------------------
// 16-bit codes (src/index.cpp:655-675, 917-943): v = clamp(lround((double(y) -
// lo) * inv_step), 0, 65535), round half away from zero; NaN and |t| >= 2^63
// give 0 like the reference's x86 conversion. One wave per row; writes the
// shifted bytes, the shifted norms and (optionally) the prefix plane.
__device__ __forceinline__ uint32_t quantize(float y, float lo, double inv_step) {
  const double t = (static_cast<double>(y) - static_cast<double>(lo)) * inv_step;
  if (!(t > -9.2e18 && t < 9.2e18)) return 0u;
  double r = round(t);
  if (r < 0.0) r = 0.0;
  if (r > 65535.0) r = 65535.0;
  return static_cast<uint32_t>(r);
}

__device__ __forceinline__ int64_t wave_sum64(int64_t v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

__global__ __launch_bounds__(256) void k_quantize(const float *__restrict__ Y, uint32_t n, uint32_t d,
                                                  uint32_t dc, uint32_t P, float lo, double inv_step,
                                                  int8_t *hi, int8_t *lop, int64_t *norm,
                                                  int32_t *hinorm, int32_t *headnorm, int8_t *pre,
                                                  uint16_t *codes_out) {
  const uint32_t row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  const int lane = lane_id();
  const uint32_t ph = min(P, d);
  int64_t s_full = 0, s_hi = 0, s_head = 0;
  for (uint32_t k = lane; k < dc; k += 64) {
    int32_t sh = 0, sl = 0;
    uint32_t v = 0;
    if (k < d) {
      v = quantize(Y[static_cast<size_t>(row) * d + k], lo, inv_step);
      sh = static_cast<int32_t>(v >> 8) - 128;
      sl = static_cast<int32_t>(v & 255u) - 128;
      const int64_t c = static_cast<int64_t>(v) - 32896;
      s_full += c * c;
      s_hi += sh * sh;
      if (k < ph) s_head += sh * sh;
    }
    hi[static_cast<size_t>(row) * dc + k] = static_cast<int8_t>(sh);
    lop[static_cast<size_t>(row) * dc + k] = static_cast<int8_t>(sl);
    if (pre && k < P) pre[static_cast<size_t>(row) * P + k] = static_cast<int8_t>(sh);
    if (codes_out) codes_out[static_cast<size_t>(row) * dc + k] = static_cast<uint16_t>(v);
  }
  s_full = wave_sum64(s_full);
  s_hi = wave_sum64(s_hi);
  s_head = wave_sum64(s_head);
  if (lane == 0) {
    norm[row] = s_full;
    if (hinorm) hinorm[row] = static_cast<int32_t>(s_hi);
    if (headnorm) headnorm[row] = static_cast<int32_t>(s_head);
  }
}

// Global min / max of the rotated database (src/index.cpp:634-646), two passes.
__global__ void k_minmax_partial(const float *__restrict__ y, uint64_t n, float *part) {
  __shared__ float smin[256], smax[256];
  float mn = FLT_MAX, mx = -FLT_MAX;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) {
    const float v = y[i];
    mn = fminf(mn, v);
    mx = fmaxf(mx, v);
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = smin[0];
    part[2 * blockIdx.x + 1] = smax[0];
  }
}

// --------------------------------------------------------------------------
// Coarse distances (src/index.cpp:846-861): all queries x all centroids from
// the full codes, dot = 65536 hh + 256 (hl + lh) + ll with int32 MFMA partials
// combined in int64, then f(qn + cn - 2 dot).
// v_mfma_i32_16x16x64_i8: lane l holds 16 bytes of row l % 16 at k-slice l / 16
// for both operands; C[i][j] sits in lane 16 (i / 4) + j, element i % 4.
__device__ __forceinline__ v4i load16(const int8_t *p) { return *reinterpret_cast<const v4i *>(p); }

__device__ __forceinline__ v4i mfma(v4i a, v4i b, v4i c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// Per 64-byte chunk of the codes the four operand planes (queries and
// centroids, high and low) are staged in LDS in MFMA fragment order (lane l's
// 16 bytes at l * 16 of a 1 KB tile fragment, so every read is conflict-free).
constexpr int CO_T = 128;                    // queries and centroids per block tile
constexpr int CO_PLANES = 4 * CO_T * 64;     // bytes of one chunk: 4 planes x 128 rows x 64
constexpr int CO_NORMS = 2 * CO_T * 8;       // the tile's query and centroid norms (int64)
constexpr int CO_STAGE = CO_PLANES + CO_NORMS;
constexpr int CO_NS = 4;                     // LDS stages: copies run CO_NS - 1 chunks ahead
constexpr size_t CO_SMEM = static_cast<size_t>(CO_NS) * CO_STAGE;

typedef __attribute__((address_space(3))) void co_lds_void;

// Asynchronous copy of 16 / 4 bytes per lane from global memory into LDS: lane
// l's bytes land at l * size past the wave-uniform LDS address.
__device__ __forceinline__ void co_dma16(const void *g, void *l) {
  __builtin_amdgcn_global_load_lds(g, (co_lds_void *)l, 16, 0, 0);
}
__device__ __forceinline__ void co_dma4(const void *g, void *l) {
  __builtin_amdgcn_global_load_lds(g, (co_lds_void *)l, 4, 0, 0);
}

__device__ __forceinline__ uint32_t co_lds_addr(const void *p) {
  return static_cast<uint32_t>(reinterpret_cast<size_t>((const __attribute__((address_space(3))) char *)p));
}

// s_waitcnt immediate (gfx9 encoding) for the vector memory counter only: at
// most n of the wave's vector memory loads in flight (loads complete in issue
// order, so the older ones are done).
constexpr int co_vm_imm(int n) { return (n & 0xf) | (((n >> 4) & 3) << 14) | (0x7 << 4) | (0xf << 8); }

__device__ __forceinline__ void co_wait_vm(uint32_t n) {
  switch (n) {
    case 0: __builtin_amdgcn_s_waitcnt(co_vm_imm(0)); break;
    case 4: __builtin_amdgcn_s_waitcnt(co_vm_imm(4)); break;
    case 5: __builtin_amdgcn_s_waitcnt(co_vm_imm(5)); break;
    case 8: __builtin_amdgcn_s_waitcnt(co_vm_imm(8)); break;
    case 9: __builtin_amdgcn_s_waitcnt(co_vm_imm(9)); break;
    case 10: __builtin_amdgcn_s_waitcnt(co_vm_imm(10)); break;
    default: __builtin_amdgcn_s_waitcnt(co_vm_imm(0)); break;
  }
}

// LDS reads of one chunk's operands, written as asm that waits for its own
// results: the compiler would otherwise make every LDS read wait for all the
// copies into LDS in flight, the chunks ahead included. a: the wave's first
// query fragment (high plane; the low plane 8 KB on), b: its first centroid
// fragment (high plane; the low plane 8 KB on).
__device__ __forceinline__ void co_read_hi(uint32_t a, uint32_t b, v4i (&ah)[4], v4i (&bh)[2], v4i (&bl)[2]) {
  asm volatile(
      "ds_read_b128 %0, %8\n\t"
      "ds_read_b128 %1, %8 offset:1024\n\t"
      "ds_read_b128 %2, %8 offset:2048\n\t"
      "ds_read_b128 %3, %8 offset:3072\n\t"
      "ds_read_b128 %4, %9\n\t"
      "ds_read_b128 %5, %9 offset:1024\n\t"
      "ds_read_b128 %6, %9 offset:8192\n\t"
      "ds_read_b128 %7, %9 offset:9216\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(ah[0]), "=&v"(ah[1]), "=&v"(ah[2]), "=&v"(ah[3]), "=&v"(bh[0]), "=&v"(bh[1]), "=&v"(bl[0]),
        "=&v"(bl[1])
      : "v"(a), "v"(b));
}
__device__ __forceinline__ void co_read_lo(uint32_t a, v4i (&al)[4]) {
  asm volatile(
      "ds_read_b128 %0, %4 offset:8192\n\t"
      "ds_read_b128 %1, %4 offset:9216\n\t"
      "ds_read_b128 %2, %4 offset:10240\n\t"
      "ds_read_b128 %3, %4 offset:11264\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(al[0]), "=&v"(al[1]), "=&v"(al[2]), "=&v"(al[3])
      : "v"(a));
}
__device__ __forceinline__ int64_t co_read_i64(uint32_t a) {
  int64_t x;
  asm volatile(
      "ds_read_b64 %0, %1\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(x)
      : "v"(a));
  return x;
}

// Coarse distances dist[q][c] = f(qn + cn - 2 dot(q, c)) of the 16-bit codes
// (reference src/index.cpp:846-877), dot = 65536 hh + 256 (hl + lh) + ll from
// four int8 MFMA products in int32 (|hh| <= 128 * 128 * Dc fits), combined in
// int64.
//
// Persistent blocks of 8 waves, one per CU (the stages take 136 KB of LDS),
// tiles of 128 queries x 128 centroids dealt round-robin over the blocks
// (block b takes tiles b, b + grid, ...). Wave (wq, wc) owns 64 queries x 32
// centroids = 4 x 2 MFMA tiles with three accumulators each (hh, hl + lh, ll);
// two waves per SIMD, so one hides the other's LDS reads and barrier waits.
// The four operand planes of every 64-byte chunk are copied straight from
// HBM into an LDS ring of CO_NS stages (global_load_lds) CO_NS - 1 chunks
// ahead, across tile boundaries, so the copies of the chunks ahead are in
// flight while one is multiplied and the next tile's first chunks while a
// tile's results are written. Wave w copies row block w (16 rows) of each
// plane, and at a tile's first chunk also 64 dwords of the tile's norms: 4
// copies per wave and chunk, 5 at a tile's first. Before a chunk, each wave
// waits for its own copies of it (vector memory counter, those of the chunks
// issued after it may stay in flight), then a barrier makes all eight waves'
// copies visible and frees the stage the next copy overwrites.
__global__ __launch_bounds__(512, 1) void k_coarse8(const int8_t *__restrict__ qhi, const int8_t *__restrict__ qlo,
                                                    const int64_t *__restrict__ qnorm, uint32_t B,
                                                    const int8_t *__restrict__ chi, const int8_t *__restrict__ clo,
                                                    const int64_t *__restrict__ cnorm, uint32_t nlist, uint32_t dc,
                                                    double step_sq, float *__restrict__ dist, uint32_t nct,
                                                    uint32_t ntiles) {
  extern __shared__ __attribute__((aligned(16))) int8_t co_st[];
  const int lane = lane_id(), w = threadIdx.x >> 6;
  const int wq = w >> 2, wc = w & 3;  // 64 queries (wq) x 32 centroids (wc)
  const uint32_t nk = dc / 64;
  const uint32_t ntile_b = blockIdx.x < ntiles ? (ntiles - blockIdx.x + gridDim.x - 1) / gridDim.x : 0u;
  const uint32_t total = ntile_b * nk;  // chunks of this block
  // Chunk bookkeeping by increments (the tile's row and column blocks advance
  // by gridDim.x tiles with a carry): no integer division per chunk, whose
  // scalar sequences would delay every chunk's copies and products.
  const uint32_t gq = gridDim.x / nct, gr = gridDim.x % nct;
  struct Cursor {
    uint32_t kc, tq, tc;  // chunk within the tile, the tile's query and centroid blocks
  };
  auto advance = [&](Cursor &u) {
    if (++u.kc == nk) {
      u.kc = 0;
      u.tq += gq;
      u.tc += gr;
      if (u.tc >= nct) {
        u.tc -= nct;
        ++u.tq;
      }
    }
  };
  Cursor ci{0u, blockIdx.x / nct, blockIdx.x % nct};  // next chunk to copy
  Cursor cc = ci;                                     // next chunk to multiply
  uint32_t issued = 0, firsts = 0;  // chunks copied; tiles' first chunks among those not yet multiplied
  // copies of chunk `issued` into stage issued % CO_NS
  auto issue = [&]() {
    const uint32_t kc = ci.kc, q0 = ci.tq * CO_T, c0 = ci.tc * CO_T;
    const uint32_t r = 16u * w + (lane & 15);
    const uint32_t qr = min(q0 + r, B - 1), cr = min(c0 + r, nlist - 1);
    const uint32_t off = kc * 64 + 16 * (lane >> 4);
    int8_t *S = co_st + (issued % CO_NS) * CO_STAGE + w * 1024;
    co_dma16(qhi + static_cast<size_t>(qr) * dc + off, S);
    co_dma16(qlo + static_cast<size_t>(qr) * dc + off, S + CO_T * 64);
    co_dma16(chi + static_cast<size_t>(cr) * dc + off, S + 2 * CO_T * 64);
    co_dma16(clo + static_cast<size_t>(cr) * dc + off, S + 3 * CO_T * 64);
    if (kc == 0) {  // dword j = 64 w + lane of [query norms | centroid norms]
      const uint32_t j = 64u * w + lane, e = (j & 255u) >> 1;
      const int64_t *src = j < 256 ? qnorm + min(q0 + e, B - 1) : cnorm + min(c0 + e, nlist - 1);
      co_dma4(reinterpret_cast<const uint32_t *>(src) + (j & 1u),
              co_st + (issued % CO_NS) * CO_STAGE + CO_PLANES + 256 * w);
      ++firsts;
    }
    advance(ci);
    ++issued;
  };
  const uint32_t pro = min(static_cast<uint32_t>(CO_NS - 1), total);
  while (issued < pro) issue();
  v4i hh[4][2], mid[4][2], ll[4][2];
  int64_t qn[4][4], cn[2];
  for (uint32_t g = 0; g < total; ++g) {
    const uint32_t kc = cc.kc;
    // copies of the chunks issued after g: 4 each, 5 for a tile's first
    if (kc == 0) --firsts;
    co_wait_vm(4u * (issued - g - 1) + firsts);
    __builtin_amdgcn_s_barrier();
    if (issued < total) issue();
    const uint32_t S = co_lds_addr(co_st + (g % CO_NS) * CO_STAGE);
    if (kc == 0) {
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) hh[a][b] = mid[a][b] = ll[a][b] = v4i{0, 0, 0, 0};
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          qn[a][r] = co_read_i64(S + CO_PLANES + 8 * (64 * wq + 16 * a + 4 * (lane >> 4) + r));
#pragma unroll
      for (int t = 0; t < 2; ++t) cn[t] = co_read_i64(S + CO_PLANES + 1024 + 8 * (32 * wc + 16 * t + (lane & 15)));
    }
    v4i bh[2], bl[2], ah[4], al[4];
    co_read_hi(S + (4 * wq) * 1024 + lane * 16, S + 2 * CO_T * 64 + (2 * wc) * 1024 + lane * 16, ah, bh, bl);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        hh[a][t] = mfma(ah[a], bh[t], hh[a][t]);
        mid[a][t] = mfma(ah[a], bl[t], mid[a][t]);
      }
    co_read_lo(S + (4 * wq) * 1024 + lane * 16, al);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        mid[a][t] = mfma(al[a], bh[t], mid[a][t]);
        ll[a][t] = mfma(al[a], bl[t], ll[a][t]);
      }
    if (kc == nk - 1) {
      const uint32_t q0 = cc.tq * CO_T, c0 = cc.tc * CO_T;
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const uint32_t c = c0 + 32 * wc + 16 * t + (lane & 15);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const uint32_t q = q0 + 64 * wq + 16 * a + 4 * (lane >> 4) + r;
            const int64_t dot = static_cast<int64_t>(hh[a][t][r]) * 65536 +
                                static_cast<int64_t>(mid[a][t][r]) * 256 + static_cast<int64_t>(ll[a][t][r]);
            if (q < B && c < nlist)
              dist[static_cast<size_t>(q) * nlist + c] = fscale(qn[a][r] + cn[t] - 2 * dot, step_sq);
          }
      }
    }
    advance(cc);
  }
}

// --------------------------------------------------------------------------
// The np smallest (float, list) keys of each query's coarse row, ascending
// (canonical form of src/index.cpp:862-874). One 256-thread block per query:
//  1. radix select of the np-th smallest value on the float bits relative to
//     the row minimum (non-negative floats order like their bits), 11-bit
//     digits from the highest set bit of the row's range: at most 3 passes,
//     and the histogram bins spread over the row's own value range;
//  2. the values below the pivot, then the ties at the pivot in list order;
//  3. a bitonic sort of the np keys in LDS.
constexpr uint32_t SEL_CACHE = 16384;  // coarse row cached in LDS up to this nlist
constexpr uint32_t SEL_MAXNP = 4096;
constexpr int SEL_BINS = 2048;

// Block-wide exclusive scan of 8 consecutive bins per thread (256 threads,
// 2048 bins): finds the bin where the running count reaches rank.
__device__ inline void sel_find(const uint32_t *hist, uint32_t rank, uint32_t *wsum, uint32_t *sh_digit,
                                uint32_t *sh_rank) {
  const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
  uint32_t h[8], sum = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    h[i] = hist[8 * tid + i];
    sum += h[i];
  }
  uint32_t inc = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(inc, o);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t cum = inc - sum;
  for (int ww = 0; ww < w; ++ww) cum += wsum[ww];
  if (cum < rank && rank <= cum + sum) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (cum < rank && rank <= cum + h[i]) {
        *sh_digit = 8 * tid + i;
        *sh_rank = rank - cum;
      }
      cum += h[i];
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_select(const float *__restrict__ dist, uint32_t B, uint32_t nlist,
                                                uint32_t np, uint32_t *__restrict__ probes,
                                                float *__restrict__ probe_d) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const bool cached = nlist <= SEL_CACHE;
  uint64_t *sel = reinterpret_cast<uint64_t *>(smem);
  uint32_t M = 1;
  while (M < np) M <<= 1;
  uint32_t *cache = reinterpret_cast<uint32_t *>(smem + static_cast<size_t>(M) * 8);
  __shared__ uint32_t hist[SEL_BINS];
  __shared__ uint32_t sh_digit, sh_rank, sh_cnt, sh_taken, wsum[4], wred[8];
  const uint32_t q = blockIdx.x;
  const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
  const uint32_t *row = reinterpret_cast<const uint32_t *>(dist) + static_cast<size_t>(q) * nlist;
  uint32_t vmin = 0xffffffffu, vmax = 0;
  for (uint32_t i = tid; i < nlist; i += 256) {
    const uint32_t v = row[i];
    if (cached) cache[i] = v;
    vmin = min(vmin, v);
    vmax = max(vmax, v);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    vmin = min(vmin, static_cast<uint32_t>(__shfl_xor(static_cast<int>(vmin), o)));
    vmax = max(vmax, static_cast<uint32_t>(__shfl_xor(static_cast<int>(vmax), o)));
  }
  if (lane == 0) {
    wred[w] = vmin;
    wred[4 + w] = vmax;
  }
  __syncthreads();
  vmin = min(min(wred[0], wred[1]), min(wred[2], wred[3]));
  vmax = max(max(wred[4], wred[5]), max(wred[6], wred[7]));
  auto val = [&](uint32_t i) -> uint32_t { return cached ? cache[i] : row[i]; };

  // radix select on r = v - vmin
  const uint32_t range = vmax - vmin;
  const int nbits = range ? 32 - __clz(range) : 0;
  uint32_t prefix = 0, mask = 0, rank = np;
  for (int top = nbits; top > 0; top -= 11) {
    const int shift = max(0, top - 11);
    const uint32_t dmask = (1u << (top - shift)) - 1u;
    for (int i = tid; i < SEL_BINS; i += 256) hist[i] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < nlist; i += 256) {
      const uint32_t r = val(i) - vmin;
      if ((r & mask) == prefix) atomicAdd(&hist[(r >> shift) & dmask], 1u);
    }
    __syncthreads();
    sel_find(hist, rank, wsum, &sh_digit, &sh_rank);
    prefix |= sh_digit << shift;
    mask |= dmask << shift;
    rank = sh_rank;
  }
  const uint32_t T = vmin + prefix, need_eq = rank, n_lt = np - need_eq;
  if (tid == 0) {
    sh_cnt = 0;
    sh_taken = 0;
  }
  __syncthreads();
  for (uint32_t i = tid; i < nlist; i += 256) {
    const uint32_t v = val(i);
    if (v < T) sel[atomicAdd(&sh_cnt, 1u)] = (static_cast<uint64_t>(v) << 32) | i;
  }
  // the need_eq lowest-numbered lists at the pivot value, in list order
  for (uint32_t base = 0; base < nlist; base += 256) {
    __syncthreads();
    const uint32_t taken = sh_taken;
    if (taken >= need_eq) break;
    const uint32_t i = base + tid;
    const bool f = i < nlist && val(i) == T;
    const uint64_t bal = ballot(f);
    const uint32_t before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(bal);
    __syncthreads();
    uint32_t off = taken;
    for (int ww = 0; ww < w; ++ww) off += wsum[ww];
    if (f && off + before < need_eq) sel[n_lt + off + before] = (static_cast<uint64_t>(T) << 32) | i;
    __syncthreads();
    if (tid == 0) sh_taken = taken + wsum[0] + wsum[1] + wsum[2] + wsum[3];
  }
  __syncthreads();
  for (uint32_t i = np + tid; i < M; i += 256) sel[i] = kNoKey;
  __syncthreads();
  // bitonic sort: every thread takes M / 512 compare-exchange pairs per step
  for (uint32_t k = 2; k <= M; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t pr = tid; pr < M / 2; pr += 256) {
        const uint32_t i = 2 * pr - (pr & (j - 1));
        const uint32_t p = i + j;
        const uint64_t a = sel[i], b = sel[p];
        const bool up = (i & k) == 0;
        if ((a > b) == up) {
          sel[i] = b;
          sel[p] = a;
        }
      }
      __syncthreads();
    }
  for (uint32_t p = tid; p < np; p += 256) {
    const uint64_t k = sel[p];
    probes[static_cast<size_t>(q) * np + p] = static_cast<uint32_t>(k);
    probe_d[static_cast<size_t>(q) * np + p] = __uint_as_float(static_cast<uint32_t>(k >> 32));
  }
}

// Probe selection with the coarse row in registers (nlist <= 256 V): thread t
// owns lists [t V, t V + V). The np-th smallest value is found by radix select
// on range-relative float bits (per-wave LDS histograms), the kept keys are
// collected in list order through block scans (values below the pivot, then
// the lowest-numbered lists at the pivot), each wave sorts a quarter of them
// in registers, and two merge-path levels finish the sort.
constexpr uint32_t SEL_FAST = 1024;  // candidates of the small-np path at most

template <int V, bool COAL>
__global__ __launch_bounds__(256, 2) void k_select_reg(const float *__restrict__ dist, uint32_t B, uint32_t nlist,
                                                    uint32_t np, uint32_t *__restrict__ probes,
                                                    float *__restrict__ probe_d) {
  extern __shared__ __attribute__((aligned(16))) uint64_t keys[];  // 2 x max(M, SEL_FAST if np <= 512)
  __shared__ uint32_t hist[SEL_BINS];  // one for the four waves: LDS for four blocks per CU
  __shared__ uint32_t tmp[8], sh_digit, sh_rank;
  const uint32_t q = blockIdx.x;
  const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
  uint32_t M = 64;
  while (M < np) M <<= 1;
  const uint32_t *row = reinterpret_cast<const uint32_t *>(dist) + static_cast<size_t>(q) * nlist;
  const uint32_t c0 = static_cast<uint32_t>(tid) * V;
  uint32_t v[V];
  // np <= 512 on a full row: the wave reads 1 KB runs (lane-contiguous 16-byte
  // pieces; thread t's value i is list 4 (256 (i / 4) + t) + i % 4): a few
  // cache lines per load instead of one per lane. The small-np path only
  // compares (value, list) keys, so any ownership serves; the radix path
  // below reloads the row thread-contiguous.
  // (COAL: np <= 512 and nlist = 256 V, chosen at launch)
  auto lid = [&](int i) -> uint32_t {
    return COAL ? 4u * (256u * static_cast<uint32_t>(i >> 2) + static_cast<uint32_t>(tid)) + static_cast<uint32_t>(i & 3)
                : c0 + static_cast<uint32_t>(i);
  };
  auto load_row = [&](bool lanes) {
    if (lanes) {
#pragma unroll
      for (int i = 0; i < V; i += 4) {
        const uint4 x = *reinterpret_cast<const uint4 *>(row + lid(i));
        v[i] = x.x;
        v[i + 1] = x.y;
        v[i + 2] = x.z;
        v[i + 3] = x.w;
      }
    } else if ((nlist & 3u) == 0 && c0 + V <= nlist) {
#pragma unroll
      for (int i = 0; i < V; i += 4) {
        const uint4 x = *reinterpret_cast<const uint4 *>(row + c0 + i);
        v[i] = x.x;
        v[i + 1] = x.y;
        v[i + 2] = x.z;
        v[i + 3] = x.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < V; ++i) v[i] = c0 + i < nlist ? row[c0 + i] : 0xffffffffu;
    }
  };
  load_row(COAL);
  // the kept keys, sorted: runs of W keys by the waves, then merge levels;
  // the first np go out
  auto sort_and_emit = [&](uint32_t M) {
    const uint64_t *src = block_sort_keys(keys, M);
    for (uint32_t p = tid; p < np; p += 256) {
      const uint64_t k = src[p];
      probes[static_cast<size_t>(q) * np + p] = static_cast<uint32_t>(k);
      probe_d[static_cast<size_t>(q) * np + p] = __uint_as_float(static_cast<uint32_t>(k >> 32));
    }
  };
  if (np <= 512) {
    // Each wave finds the ceil(np / 4)-th smallest of its threads' R
    // smallest distances (R = 1, 2 or 4, so that they number at least twice
    // that; bit by bit from the top, one ballot per bit and value); they are
    // distinct lists, so at least np keys lie at or below the largest of the
    // four bounds, and for rows without long runs of small values in few
    // threads not many more: those candidates hold the answer. Past SEL_FAST
    // candidates the radix path below runs instead.
    const uint32_t K = (np + 3) / 4, R = K <= 32 ? 1u : K <= 64 ? 2u : 4u;
    uint32_t m0 = 0xffffffffu, m1 = 0xffffffffu, m2 = 0xffffffffu, m3 = 0xffffffffu;  // ascending
    if (R == 1) {
#pragma unroll
      for (int i = 0; i < V; ++i)
        if (lid(i) < nlist) m0 = min(m0, v[i]);
    } else {
#pragma unroll
      for (int i = 0; i < V; ++i)
        if (lid(i) < nlist) {
          m3 = min(m3, max(m2, v[i]));
          m2 = min(m2, max(m1, v[i]));
          m1 = min(m1, max(m0, v[i]));
          m0 = min(m0, v[i]);
        }
    }
    uint32_t T = 0;  // bits above b of the K-th smallest
    for (int b = 31; b >= 0; --b) {
      const uint32_t hi0 = T | ((1u << b) - 1u);  // largest value with bit b clear under the prefix
      uint32_t c = static_cast<uint32_t>(__popcll(ballot(m0 <= hi0)));
      if (R > 1) c += static_cast<uint32_t>(__popcll(ballot(m1 <= hi0)));
      if (R > 2) c += static_cast<uint32_t>(__popcll(ballot(m2 <= hi0))) + static_cast<uint32_t>(__popcll(ballot(m3 <= hi0)));
      if (c < K) T |= 1u << b;
    }
    const uint64_t wk = (static_cast<uint64_t>(T) << 32) | 0xffffffffull;
    if (lane == 0) keys[w] = wk;
    __syncthreads();
    const uint64_t tk = umax64(umax64(keys[0], keys[1]), umax64(keys[2], keys[3]));
    uint32_t cnt = 0;
#pragma unroll
    for (int i = 0; i < V; ++i)
      if (lid(i) < nlist) cnt += ((static_cast<uint64_t>(v[i]) << 32) | lid(i)) <= tk ? 1u : 0u;
    uint32_t off = block_exscan256(cnt, tmp);
    if (tid == 255) sh_rank = off + cnt;
    __syncthreads();
    const uint32_t total = sh_rank;
    if (total <= SEL_FAST) {
      uint32_t C = 64;
      while (C < total) C <<= 1;
#pragma unroll
      for (int i = 0; i < V; ++i)
        if (lid(i) < nlist) {
          const uint64_t key = (static_cast<uint64_t>(v[i]) << 32) | lid(i);
          if (key <= tk) keys[off++] = key;
        }
      for (uint32_t p = total + tid; p < C; p += 256) keys[p] = kNoKey;
      __syncthreads();
      sort_and_emit(C);
      return;
    }
    __syncthreads();
    if (COAL) load_row(false);
  }
  // range of the row (lists past nlist hold 0xffffffff, above every distance)
  uint32_t vmin = 0xffffffffu, vmax = 0;
#pragma unroll
  for (int i = 0; i < V; ++i)
    if (c0 + i < nlist) {
      vmin = min(vmin, v[i]);
      vmax = max(vmax, v[i]);
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    vmin = min(vmin, static_cast<uint32_t>(__shfl_xor(static_cast<int>(vmin), o)));
    vmax = max(vmax, static_cast<uint32_t>(__shfl_xor(static_cast<int>(vmax), o)));
  }
  if (lane == 0) {
    tmp[w] = vmin;
    tmp[4 + w] = vmax;
  }
  __syncthreads();
  vmin = min(min(tmp[0], tmp[1]), min(tmp[2], tmp[3]));
  vmax = max(max(tmp[4], tmp[5]), max(tmp[6], tmp[7]));
  __syncthreads();
  // radix select of the np-th smallest on r = v - vmin, 11-bit digits
  const uint32_t range = vmax - vmin;
  const int nbits = range ? 32 - __clz(range) : 0;
  uint32_t prefix = 0, mask = 0, rank = np;
  for (int top = nbits; top > 0; top -= 11) {
    const int shift = max(0, top - 11);
    const uint32_t dmask = (1u << (top - shift)) - 1u;
    for (int i = tid; i < SEL_BINS; i += 256) hist[i] = 0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const uint32_t r = v[i] - vmin;
      if (c0 + i < nlist && (r & mask) == prefix) atomicAdd(&hist[(r >> shift) & dmask], 1u);
    }
    __syncthreads();
    // bins 8 tid .. 8 tid + 7, summed over the waves; the bin holding the rank-th
    uint32_t h[8], sum = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int b = 8 * tid + i;
      h[i] = hist[b];
      sum += h[i];
    }
    uint32_t cum = block_exscan256(sum, tmp);
    if (cum < rank && rank <= cum + sum) {
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        if (cum < rank && rank <= cum + h[i]) {
          sh_digit = 8 * tid + i;
          sh_rank = rank - cum;
        }
        cum += h[i];
      }
    }
    __syncthreads();
    prefix |= sh_digit << shift;
    mask |= dmask << shift;
    rank = sh_rank;
    __syncthreads();
  }
  // the kept keys in list order: below the pivot, then the first need_eq at it
  const uint32_t T = vmin + prefix, need_eq = rank, n_lt = np - need_eq;
  uint32_t clt = 0, ceq = 0;
#pragma unroll
  for (int i = 0; i < V; ++i)
    if (c0 + i < nlist) {
      clt += v[i] < T ? 1u : 0u;
      ceq += v[i] == T ? 1u : 0u;
    }
  uint32_t olt = block_exscan256(clt, tmp);
  uint32_t oeq = block_exscan256(ceq, tmp);
#pragma unroll
  for (int i = 0; i < V; ++i)
    if (c0 + i < nlist) {
      const uint64_t key = (static_cast<uint64_t>(v[i]) << 32) | (c0 + i);
      if (v[i] < T) keys[olt++] = key;
      else if (v[i] == T) {
        if (oeq < need_eq) keys[n_lt + oeq] = key;
        ++oeq;
      }
    }
  for (uint32_t p = np + tid; p < M; p += 256) keys[p] = kNoKey;
  __syncthreads();
  sort_and_emit(M);
}

// --------------------------------------------------------------------------
// Inversion (src/index.cpp:403-412) and the work plan, list-major and without
// global atomics. Work is grouped by (probe-rank band, list): for band factor
// F = 2^shift, band b holds ranks [(F^b - 1) / (F - 1), (F^(b+1) - 1) / (F - 1)),
// so a query's nearest lists come first in the queue and its threshold is tight
// when its farther lists are ranked. Pair (q, c) is probed iff the key
// (coarse distance bits, c) is at most the key of q's last probe, and its band
// is the number of band-start keys of q at or below it; the coarse matrix is
// read in tiles of 256 lists x 256 queries:
//   k_inv_pass<false> per-tile counts of every (band, list) slot,
//   k_slot_totals     per slot: the tiles' offsets within it, its query count
//                     and its items; per 1024 slots: their sums,
//   k_slot_scan       slot and item offsets (exclusive scans),
//   k_inv_pass<true>  the query lists of every slot (ascending queries).
// When nprobe is well below nlist (inv_from_probes) the two passes read the
// probe lists instead: k_inv_probe, one thread per probed pair, its band from
// its probe rank, counts and query-list cursors by atomics.
__global__ void k_fill32(uint32_t *p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// Per-query probe counts from the caller: a count outside [1, np] sets
// state[2] (the host turns it into std::invalid_argument) and is clamped
// into range, so the kernels after it read only probed ranks.
__global__ void k_check_npq(uint32_t *npq, uint32_t B, uint32_t np, uint32_t *state) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= B) return;
  const uint32_t v = npq[q];
  if (v == 0 || v > np) {
    atomicOr(&state[2], 1u);
    npq[q] = v == 0 ? 1u : np;
  }
}

constexpr int INV_QT = 64;    // queries per tile
// The probe-list inversion counts a slot's pairs on INV_PG counters (the
// pair's block picks one), so that the thousands of pairs of a popular list
// do not all serialise on one address; the counters' offsets within the slot
// are then scanned like the matrix path's per-tile counts.
constexpr uint32_t INV_PG = 8;
constexpr int INV_MAXB = 16;  // bands at most (nprobe 4096, F = 2: 13)

__device__ __forceinline__ uint64_t probe_key(const uint32_t *probes, const float *probe_d, size_t i) {
  return (static_cast<uint64_t>(__float_as_uint(probe_d[i])) << 32) | probes[i];
}

struct InvQuery {
  uint64_t kb[INV_MAXB];  // kb[b]: key of the first probe of band b (b >= 1)
  uint64_t kend;          // key of the last probe
  bool valid;
};

// Band-start keys of query q: kb[b] = key of the first probe of key band b
// (b >= 1). Without a seed the key bands are the slot bands; with one, key
// band 0 is the seed (slot bands 0 and 1) and key band b >= 1 is slot band
// b + 1, starting at rank r0 + F (F^(b-1) - 1) / (F - 1) (r0 = seed[q]).
__device__ inline InvQuery inv_query(const uint32_t *probes, const float *probe_d, const uint32_t *npq, uint32_t q,
                                     uint32_t B, uint32_t np, uint32_t nb, uint32_t shift, const uint32_t *seed) {
  InvQuery r;
  r.kend = 0;
  r.valid = q < B;
#pragma unroll
  for (int b = 0; b < INV_MAXB; ++b) r.kb[b] = kNoKey;
  if (q >= B) return r;
  const uint32_t n = npq ? min(npq[q], np) : np;
  const size_t base = static_cast<size_t>(q) * np;
  r.kend = probe_key(probes, probe_d, base + n - 1);
  if (seed) {
    uint64_t sb = min(seed[q], n), step = 1ull << shift;
#pragma unroll
    for (int b = 1; b < INV_MAXB; ++b) {
      if (static_cast<uint32_t>(b) + 1 >= nb) break;
      if (sb < n) r.kb[b] = probe_key(probes, probe_d, base + sb);
      sb += step;
      step = min(step << shift, 1ull << 32);
    }
    return r;
  }
  const uint32_t F1 = (1u << shift) - 1u;
#pragma unroll
  for (int b = 1; b < INV_MAXB; ++b) {
    if (static_cast<uint32_t>(b) >= nb) break;
    const uint32_t sb = ((1u << (shift * b)) - 1u) / F1;
    if (sb < n) r.kb[b] = probe_key(probes, probe_d, base + sb);
  }
  return r;
}

// Seed ranks: per query, the fewest nearest lists whose sizes, each counted up
// to R0 rows, reach R0 (all its probes if they do not). Probes are read eight
// at a time so the loads of a step are independent.
__global__ void k_seed_ranks(const uint32_t *__restrict__ probes, const uint32_t *__restrict__ npq,
                             const uint64_t *__restrict__ list_off, uint32_t B, uint32_t np, uint32_t seed,
                             uint32_t *__restrict__ r0) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= B) return;
  const uint32_t n = npq ? min(npq[q], np) : np;
  const uint32_t *pr = probes + static_cast<size_t>(q) * np;
  uint64_t cum = 0;
  uint32_t r = 0;
  while (r < n && cum < seed) {
    uint64_t sz[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t c = r + u < n ? pr[r + u] : 0u;
      sz[u] = r + u < n ? list_off[c + 1] - list_off[c] : 0u;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (r < n && cum < seed) {
        cum += min(sz[u], static_cast<uint64_t>(seed));
        ++r;
      }
    }
  }
  r0[q] = r;
}

// One thread per list, walking a tile of INV_QT queries in ascending order
// (reads of the coarse matrix coalesced across the lists): counts of each
// (band, list) slot in the tile (FILL false), or the tile's queries appended to
// the slots' query lists from the tile's offsets (FILL true). The queries'
// band-start keys are staged in LDS and read as broadcasts.
constexpr int INV_LB = 256;  // lists per block

template <bool FILL, int NB>
__global__ __launch_bounds__(256) void k_inv_pass(const float *__restrict__ dist, const uint32_t *__restrict__ probes,
                                                  const float *__restrict__ probe_d, const uint32_t *__restrict__ npq,
                                                  uint32_t B, uint32_t np, uint32_t nlist, uint32_t shift,
                                                  const uint32_t *__restrict__ seed_r0, uint32_t seed,
                                                  const uint64_t *__restrict__ list_off, uint32_t *__restrict__ part,
                                                  const uint32_t *__restrict__ qoff, uint32_t *__restrict__ qlist) {
  constexpr uint32_t nb = NB;
  __shared__ uint64_t s_kend[INV_QT];
  __shared__ uint64_t s_kb[NB][INV_QT];
  const uint32_t qt = blockIdx.y, q0 = qt * INV_QT, c = blockIdx.x * INV_LB + threadIdx.x;
  const uint32_t nslots = nb * nlist;
  if (threadIdx.x < INV_QT) {
    const InvQuery r = inv_query(probes, probe_d, npq, q0 + threadIdx.x, B, np, nb, shift, seed ? seed_r0 : nullptr);
    s_kend[threadIdx.x] = r.kend;
#pragma unroll
    for (int b = 0; b < NB; ++b) s_kb[b][threadIdx.x] = r.kb[b];
  }
  __syncthreads();
  if (c >= nlist) return;
  const uint32_t nq = min(static_cast<uint32_t>(INV_QT), B - q0);
  // with a seed, key band 0 also feeds slot band 1 when the list has rows past R0
  const bool rest = seed && list_off[c + 1] - list_off[c] > seed;
  uint32_t acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const size_t slot = static_cast<size_t>(b) * nlist + c;
    acc[b] = FILL ? qoff[slot] + part[static_cast<size_t>(qt) * nslots + slot] : 0u;
  }
  const float *col = dist + static_cast<size_t>(q0) * nlist + c;
  constexpr int U = 8;
  for (uint32_t i0 = 0; i0 < nq; i0 += U) {
    uint32_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = i0 + u < nq ? __float_as_uint(col[static_cast<size_t>(i0 + u) * nlist]) : 0u;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + u;
      if (i >= nq) break;
      const uint64_t k = (static_cast<uint64_t>(v[u]) << 32) | c;
      if (k > s_kend[i]) continue;
      uint32_t band = 0;
#pragma unroll
      for (int b = 1; b < NB; ++b) band += k >= s_kb[b][i] ? 1u : 0u;
      if (seed && band) ++band;  // key band b >= 1 -> slot band b + 1
#pragma unroll
      for (int b = 0; b < NB; ++b)
        if (band == static_cast<uint32_t>(b) || (b == 1 && band == 0 && rest)) {
          if (FILL) qlist[acc[b]] = q0 + i;
          ++acc[b];
        }
    }
  }
  if (!FILL)
#pragma unroll
    for (int b = 0; b < NB; ++b) part[static_cast<size_t>(qt) * nslots + static_cast<size_t>(b) * nlist + c] = acc[b];
}

// Inversion from the probe lists (nprobe well below nlist: B x nprobe pairs
// instead of the B x nlist matrix). Slot band of probe rank r (r0: the query's seed ranks when seed > 0).
__device__ inline uint32_t rank_band(uint32_t r, uint32_t shift, uint32_t seed, uint32_t r0) {
  const uint64_t F = 1ull << shift, cap = 1ull << 40;
  if (seed) {
    if (r < r0) return 0;
    const uint64_t rr = r - r0;
    uint64_t s = 0, size = F;
    uint32_t b = 2;
    while (rr >= s + size) {
      s += size;
      size = min(size * F, cap);
      ++b;
    }
    return b;
  }
  uint64_t s = 0, size = 1;
  uint32_t b = 0;
  while (r >= s + size) {
    s += size;
    size = min(size * F, cap);
    ++b;
  }
  return b;
}

// One thread per (query, probe rank) pair of the batch.
template <bool FILL>
__global__ __launch_bounds__(256) void k_inv_probe(const uint32_t *__restrict__ probes, const uint32_t *__restrict__ npq,
                                                   uint32_t B, uint32_t np, uint32_t nlist, uint32_t shift,
                                                   const uint32_t *__restrict__ seed_r0, uint32_t seed,
                                                   const uint64_t *__restrict__ list_off, uint32_t *__restrict__ part,
                                                   const uint32_t *__restrict__ qoff, uint32_t *__restrict__ qlist,
                                                   uint32_t join, uint32_t *__restrict__ acnt, uint32_t nslots) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  uint32_t *pc = part + static_cast<size_t>(blockIdx.x % INV_PG) * nslots;  // this block's counters
  if (e >= static_cast<uint64_t>(B) * np) return;
  const uint32_t q = static_cast<uint32_t>(e / np), r = static_cast<uint32_t>(e % np);
  const uint32_t n = npq ? min(npq[q], np) : np;
  if (r >= n) return;
  const uint32_t c = probes[e];
  const uint32_t b = rank_band(r, shift, seed, seed ? seed_r0[q] : 0u);
  const size_t slot = static_cast<size_t>(b) * nlist + c;
  if (FILL) qlist[qoff[slot] + atomicAdd(&pc[slot], 1u)] = q;
  else atomicAdd(&pc[slot], 1u);
  if (seed && b == 0 && list_off[c + 1] - list_off[c] > seed) {
    if (join) {  // after the slot's own queries of the first probe-rank band (acnt: count, then cursor)
      const size_t rest = static_cast<size_t>(katoann_hip::kJoinBand) * nlist + c;
      if (FILL) qlist[qoff[rest] + atomicAdd(&acnt[c], 1u)] = q;
      else atomicAdd(&acnt[c], 1u);
    } else {
      const size_t rest = static_cast<size_t>(nlist) + c;
      if (FILL) qlist[qoff[rest] + atomicAdd(&pc[rest], 1u)] = q;
      else atomicAdd(&pc[rest], 1u);
    }
  }
}

// Block-wide exclusive scan of one value per thread (1024 threads).
__device__ inline uint64_t block_exscan(uint64_t v, uint64_t *tmp, uint64_t *total) {
  const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
  uint64_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) tmp[w] = x;
  __syncthreads();
  if (tid == 0) {
    uint64_t s = 0;
    for (int i = 0; i < 16; ++i) {
      const uint64_t t = tmp[i];
      tmp[i] = s;
      s += t;
    }
    tmp[16] = s;
  }
  __syncthreads();
  const uint64_t r = x - v + tmp[w];
  if (total) *total = tmp[16];
  __syncthreads();
  return r;
}

// Per slot: the tiles' offsets within it (in place), its queries and items;
// per block of 1024 slots: their sums.
__global__ __launch_bounds__(1024) void k_slot_totals(uint32_t *part, uint32_t qtiles, uint32_t nslots,
                                                      uint32_t nlist, const uint64_t *__restrict__ list_off,
                                                      uint32_t qc, uint32_t rt, uint32_t qsq, uint32_t qsrt,
                                                      uint32_t seed, uint32_t *gcnt, uint32_t *nit, uint32_t *bsum,
                                                      uint32_t *ihist, uint32_t join, uint32_t *acnt, uint32_t *gbv) {
  __shared__ uint64_t tmp[17];
  __shared__ uint32_t hist[katoann_hip::kPlanBands * katoann_hip::kItemClasses];
  for (uint32_t i = threadIdx.x; i < katoann_hip::kPlanBands * katoann_hip::kItemClasses; i += 1024) hist[i] = 0;
  __syncthreads();
  const uint32_t sl = blockIdx.x * 1024 + threadIdx.x;
  uint32_t g = 0, it = 0;
  if (sl < nslots) {
    for (uint32_t t = 0; t < qtiles; ++t) {
      const size_t k = static_cast<size_t>(t) * nslots + sl;
      const uint32_t v = part[k];
      part[k] = g;
      g += v;
    }
    const uint32_t c = sl % nlist;
    uint64_t first, end;
    katoann_hip::slot_rows(sl / nlist, list_off[c + 1] - list_off[c], seed, first, end);
    const uint32_t n = static_cast<uint32_t>(end - first);
    uint32_t *hb = hist + (sl / nlist) * katoann_hip::kItemClasses;
    // the items per cost class (k_items places them by class) of g queries
    // in chunks of qc over `rows` rows
    auto classes = [&](uint32_t gq, uint32_t rows) {
      if (!gq || !rows) return;
      const uint32_t nch = (gq + qc - 1) / qc, nt = (rows + rt - 1) / rt;
      if (nch > 1) atomicAdd(&hb[katoann_hip::item_class(qc)], (nch - 1) * nt);
      atomicAdd(&hb[katoann_hip::item_class(gq - (nch - 1) * qc)], nt);
    };
    const uint32_t ga = join && sl / nlist == katoann_hip::kJoinBand ? acnt[c] : 0u;
    if (join && sl / nlist == katoann_hip::kJoinBand) {
      gbv[c] = g;
      acnt[c] = g;  // the joined queries' cursor starts after the slot's own
    }
    if (ga) {  // joined seed remainder: rows [0, n1) x g queries, [n1, n) x g + ga
      const uint32_t n1 = min(seed, n);
      it = katoann_hip::chunk_items(g, n1, qc, rt) + katoann_hip::chunk_items(g + ga, n - n1, qc, rt);
      classes(g, n1);
      classes(g + ga, n - n1);
      g += ga;
    } else if (g && n && katoann_hip::slot_qs(g, qsq)) {
      it = (n + qsrt - 1) / qsrt;
      atomicAdd(&hb[katoann_hip::item_class(g)], it);
    } else {
      it = katoann_hip::chunk_items(g, n, qc, rt);
      classes(g, n);
    }
    gcnt[sl] = g;
    nit[sl] = it;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < katoann_hip::kPlanBands * katoann_hip::kItemClasses; i += 1024)
    if (hist[i]) atomicAdd(&ihist[i], hist[i]);
  uint64_t tg = 0, ti = 0;
  (void)block_exscan(g, tmp, &tg);
  (void)block_exscan(it, tmp, &ti);
  if (threadIdx.x == 0) {
    bsum[2 * blockIdx.x] = static_cast<uint32_t>(tg);
    bsum[2 * blockIdx.x + 1] = static_cast<uint32_t>(ti);
  }
}

// Slot offsets into the query lists and item offsets (exclusive scans); the
// last entry of item_off and totals = (pairs, items).
__global__ __launch_bounds__(1024) void k_slot_scan(const uint32_t *__restrict__ gcnt, const uint32_t *__restrict__ nit,
                                                    const uint32_t *__restrict__ bsum, uint32_t nslots, uint32_t *qoff,
                                                    uint32_t *item_off, uint32_t *totals) {
  __shared__ uint64_t tmp[17];
  __shared__ uint32_t sbase[2];
  const int tid = threadIdx.x, lane = lane_id();
  if (tid < 64) {
    uint64_t a = 0, b = 0;
    for (uint32_t k = lane; k < blockIdx.x; k += 64) {
      a += bsum[2 * k];
      b += bsum[2 * k + 1];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      a += __shfl_xor(a, o);
      b += __shfl_xor(b, o);
    }
    if (lane == 0) {
      sbase[0] = static_cast<uint32_t>(a);
      sbase[1] = static_cast<uint32_t>(b);
    }
  }
  __syncthreads();
  const uint32_t sl = blockIdx.x * 1024 + tid;
  const uint32_t g = sl < nslots ? gcnt[sl] : 0u, it = sl < nslots ? nit[sl] : 0u;
  uint64_t tg = 0, ti = 0;
  const uint64_t og = block_exscan(g, tmp, &tg);
  const uint64_t oi = block_exscan(it, tmp, &ti);
  if (sl < nslots) {
    qoff[sl] = sbase[0] + static_cast<uint32_t>(og);
    item_off[sl] = sbase[1] + static_cast<uint32_t>(oi);
  }
  if (blockIdx.x == gridDim.x - 1 && tid == 0) {
    item_off[nslots] = sbase[1] + static_cast<uint32_t>(ti);
    totals[0] = sbase[0] + static_cast<uint32_t>(tg);
    totals[1] = sbase[1] + static_cast<uint32_t>(ti);
  }
}

// Per-batch resets of each query's pool and threshold (none yet: every
// candidate of the first band is kept).
__global__ void k_query_init(uint32_t B, uint32_t *pool_cnt, uint32_t *pool_sorted, uint64_t *thr,
                             int32_t *lbound) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= B) return;
  pool_cnt[static_cast<size_t>(q) * katoann_hip::kCntStride] = 0;
  pool_sorted[q] = 0;
  thr[q] = kNoKey;
  lbound[q] = 0x7fffffff;
}

// --------------------------------------------------------------------------
// Kernel-level distances for the known-answer tests (reference kernels.h:34-52):
// m queries x n vectors, row-major shifted planes.
__global__ void k_sq16_pairs(const int8_t *qh, const int8_t *ql, const int64_t *qn, const int8_t *xh,
                             const int8_t *xl, const int64_t *xn, uint32_t m, uint32_t n, uint32_t dc,
                             uint32_t qstride, uint32_t xstride, int full, double step_sq, float *out) {
  // one wave per 16 x 16 output tile
  const int lane = lane_id();
  const uint32_t tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t tn = (n + 15) / 16;
  if (tile >= ((m + 15) / 16) * tn) return;
  const uint32_t i0 = (tile / tn) * 16, j0 = (tile % tn) * 16;
  const uint32_t qi = min(i0 + (lane & 15), m - 1), xj = min(j0 + (lane & 15), n - 1);
  const int koff = 16 * (lane >> 4);
  v4i hh = {0, 0, 0, 0}, mid = {0, 0, 0, 0}, ll = {0, 0, 0, 0};
  for (uint32_t kc = 0; kc < dc; kc += 64) {
    const v4i ah = load16(qh + static_cast<size_t>(qi) * qstride + kc + koff);
    const v4i bh = load16(xh + static_cast<size_t>(xj) * xstride + kc + koff);
    hh = mfma(ah, bh, hh);
    if (full) {
      const v4i al = load16(ql + static_cast<size_t>(qi) * qstride + kc + koff);
      const v4i bl = load16(xl + static_cast<size_t>(xj) * xstride + kc + koff);
      mid = mfma(ah, bl, mid);
      mid = mfma(al, bh, mid);
      ll = mfma(al, bl, ll);
    }
  }
  const uint32_t j = j0 + (lane & 15);
  if (j >= n) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t i = i0 + 4 * (lane >> 4) + r;
    if (i >= m) continue;
    const int64_t dot = full ? static_cast<int64_t>(hh[r]) * 65536 + static_cast<int64_t>(mid[r]) * 256 +
                                   static_cast<int64_t>(ll[r])
                             : static_cast<int64_t>(hh[r]);
    out[static_cast<size_t>(i) * n + j] = fscale(qn[i] + xn[j] - 2 * dot, step_sq);
  }
}

}  // namespace

// ============================================================================
// Launchers

namespace katoann_hip {

void launch_normalize(float *x, uint32_t n, uint32_t d, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_normalize, dim3((n + 255) / 256), dim3(256), 0, s, x, n, d);
}

void launch_rotate(const float *X, uint32_t n, uint32_t d, const float *R, float *Y, hipStream_t s) {
  if (!n) return;
  const uint32_t d4 = d & ~3u;
  if (d4) {
    dim3 grid((n + RT_ROWS - 1) / RT_ROWS, (d4 + RT_COLS - 1) / RT_COLS);
    hipLaunchKernelGGL(k_rotate<false>, grid, dim3(512), 0, s, X, n, d, R, d4, Y, d);
  }
  if (d4 < d) {
    const uint64_t t = static_cast<uint64_t>(n) * (d - d4);
    hipLaunchKernelGGL(k_rotate_tail, dim3(static_cast<uint32_t>((t + 255) / 256)), dim3(256), 0, s, X, n, d, R,
                       Y);
  }
}

void launch_coarse_flat(const float *Q, uint32_t B, uint32_t d, const float *cen, uint32_t nlist, float *dist,
                        hipStream_t s) {
  if (!B || !nlist) return;
  dim3 grid((B + RT_ROWS - 1) / RT_ROWS, (nlist + RT_COLS - 1) / RT_COLS);
  hipLaunchKernelGGL(k_rotate<true>, grid, dim3(512), 0, s, Q, B, d, cen, nlist, dist, nlist);
}

void launch_bf16_pairs(const uint16_t *q, const uint16_t *x, const float *qn, const float *xn, uint32_t m,
                       uint32_t n, uint32_t dim, float *out, hipStream_t s) {
  const uint64_t t = static_cast<uint64_t>(m) * n;
  if (!t) return;
  hipLaunchKernelGGL(k_bf16_pairs, dim3(static_cast<uint32_t>((t + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<const uint32_t *>(q), reinterpret_cast<const uint32_t *>(x), qn, xn, m, n,
                     dim / 2, out);
}

void launch_to_bf16(const float *X, uint32_t n, uint32_t d, uint32_t dp, uint16_t *out, float *norms,
                    hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_to_bf16, dim3((n + 3) / 4), dim3(256), 0, s, X, n, d, dp, out, norms);
}

void launch_quantize(const float *Y, uint32_t n, uint32_t d, uint32_t dc, uint32_t P, float lo,
                     double inv_step, int8_t *hi, int8_t *lop, int64_t *norm, int32_t *hinorm,
                     int32_t *headnorm, int8_t *pre, uint16_t *codes, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_quantize, dim3((n + 3) / 4), dim3(256), 0, s, Y, n, d, dc, P, lo, inv_step, hi, lop,
                     norm, hinorm, headnorm, pre, codes);
}

void launch_minmax(const float *y, uint64_t n, float *part, uint32_t nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_minmax_partial, dim3(nblocks), dim3(256), 0, s, y, n, part);
}

void launch_coarse(const int8_t *qhi, const int8_t *qlo, const int64_t *qnorm, uint32_t B, const int8_t *chi,
                   const int8_t *clo, const int64_t *cnorm, uint32_t nlist, uint32_t dc, double step_sq,
                   float *dist, int cus, hipStream_t s) {
  static const bool attr = [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_coarse8),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(CO_SMEM));
    return true;
  }();
  (void)attr;
  const uint32_t nct = (nlist + CO_T - 1) / CO_T, ntiles = nct * ((B + CO_T - 1) / CO_T);
  const uint32_t grid = min(ntiles, static_cast<uint32_t>(cus));
  hipLaunchKernelGGL(k_coarse8, dim3(grid), dim3(512), CO_SMEM, s, qhi, qlo, qnorm, B, chi, clo, cnorm, nlist, dc,
                     step_sq, dist, nct, ntiles);
}

size_t select_smem(uint32_t nlist, uint32_t np) {
  uint32_t M = 1;
  while (M < np) M <<= 1;
  return (nlist <= SEL_CACHE ? static_cast<size_t>(nlist) * 4 : 0) + static_cast<size_t>(M) * 8;
}

template <int V>
static void select_reg_launch(const float *dist, uint32_t B, uint32_t nlist, uint32_t np, uint32_t *probes,
                              float *probe_d, hipStream_t s) {
  uint32_t M = 64;
  while (M < np) M <<= 1;
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_select_reg<V, false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(2 * SEL_MAXNP * 8));
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_select_reg<V, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(2 * SEL_MAXNP * 8));
    attr = true;
  }
  const uint32_t words = np <= 512 && M < SEL_FAST ? 2 * SEL_FAST : 2 * M;  // the sort's two buffers
  if (np <= 512 && nlist == 256u * V)
    hipLaunchKernelGGL((k_select_reg<V, true>), dim3(B), dim3(256), words * 8, s, dist, B, nlist, np, probes, probe_d);
  else
    hipLaunchKernelGGL((k_select_reg<V, false>), dim3(B), dim3(256), words * 8, s, dist, B, nlist, np, probes, probe_d);
}

void launch_select(const float *dist, uint32_t B, uint32_t nlist, uint32_t np, uint32_t *probes, float *probe_d,
                   hipStream_t s) {
  if (nlist <= 256 * 16) return select_reg_launch<16>(dist, B, nlist, np, probes, probe_d, s);
  if (nlist <= 256 * 32) return select_reg_launch<32>(dist, B, nlist, np, probes, probe_d, s);
  if (nlist <= 256 * 64) return select_reg_launch<64>(dist, B, nlist, np, probes, probe_d, s);
  const size_t sm = select_smem(nlist, np);
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_select),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(select_smem(SEL_CACHE, SEL_MAXNP)));
    attr = true;
  }
  hipLaunchKernelGGL(k_select, dim3(B), dim3(256), sm, s, dist, B, nlist, np, probes, probe_d);
}

uint32_t plan_slots(uint32_t nlist, uint32_t np, uint32_t shift, uint32_t seed) {
  return plan_bands(np, shift, seed) * nlist;
}

size_t plan_scratch_words(uint32_t B, uint32_t nlist, uint32_t np, uint32_t shift, uint32_t seed) {
  const size_t nslots = plan_slots(nlist, np, shift, seed);
  const size_t qtiles = (B + INV_QT - 1) / INV_QT, nblk = (nslots + 1023) / 1024;
  // tile partial counts, item counts, block sums, seed ranks, and for a
  // joined seed remainder its counts / cursors and the slots' own counts
  return (inv_from_probes(nlist, np) ? INV_PG : qtiles) * nslots + nslots + 2 * nblk + B + 2ull * nlist;
}

template <bool FILL>
static void inv_pass_launch(uint32_t nb, dim3 tiles, hipStream_t s, const float *dist, const uint32_t *probes,
                            const float *probe_d, const uint32_t *npq, uint32_t B, uint32_t np, uint32_t nlist,
                            uint32_t shift, const uint32_t *seed_r0, uint32_t seed, const uint64_t *list_off,
                            uint32_t *part, const uint32_t *qoff, uint32_t *qlist) {
  switch (nb) {
#define NB_CASE(n)                                                                                              \
  case n:                                                                                                       \
    hipLaunchKernelGGL((k_inv_pass<FILL, n>), tiles, dim3(256), 0, s, dist, probes, probe_d, npq, B, np, nlist, \
                       shift, seed_r0, seed, list_off, part, qoff, qlist);                                      \
    break;
    NB_CASE(1) NB_CASE(2) NB_CASE(3) NB_CASE(4) NB_CASE(5) NB_CASE(6) NB_CASE(7) NB_CASE(8)
    NB_CASE(9) NB_CASE(10) NB_CASE(11) NB_CASE(12) NB_CASE(13) NB_CASE(14) NB_CASE(15) NB_CASE(16)
#undef NB_CASE
    default:
      break;
  }
}

void launch_fill32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_fill32, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
}

void launch_plan(const float *dist, const uint32_t *probes, const float *probe_d, const uint32_t *npq, bool check_npq,
                 uint32_t B, uint32_t np, const uint64_t *list_off, uint32_t nlist, uint32_t qc, uint32_t rt,
                 uint32_t qsq, uint32_t qsrt, uint32_t *scratch, uint32_t *gcnt, uint32_t *qoff, uint32_t *item_off,
                 uint32_t *qlist, uint32_t *totals, uint32_t *pool_cnt, uint32_t *pool_sorted, uint64_t *thr,
                 int32_t *lbound, uint32_t *state, uint32_t nstate, uint32_t shift, uint32_t seed, uint32_t join,
                 hipStream_t s) {
  const uint32_t nb = plan_bands(np, shift, seed), nslots = nb * nlist;
  const bool from_probes = inv_from_probes(nlist, np);
  const uint32_t qtiles = from_probes ? INV_PG : (B + INV_QT - 1) / INV_QT, nblk = (nslots + 1023) / 1024;
  uint32_t *part = scratch, *nit = scratch + static_cast<size_t>(qtiles) * nslots, *bsum = nit + nslots;
  uint32_t *seed_r0 = bsum + 2 * nblk, *acnt = seed_r0 + B, *gbv = acnt + nlist;
  const dim3 tiles((nlist + INV_LB - 1) / INV_LB, qtiles);
  // the state and, after it, the per (band, class) item counts and cursors
  const uint32_t nclear = nstate + katoann_hip::kItemHistWords;
  hipLaunchKernelGGL(k_fill32, dim3((nclear + 255) / 256), dim3(256), 0, s, state, nclear, 0u);
  if (npq && check_npq)
    hipLaunchKernelGGL(k_check_npq, dim3((B + 255) / 256), dim3(256), 0, s, const_cast<uint32_t *>(npq), B, np, state);
  if (seed)
    hipLaunchKernelGGL(k_seed_ranks, dim3((B + 255) / 256), dim3(256), 0, s, probes, npq, list_off, B, np, seed,
                       seed_r0);
  if (from_probes) {
    const uint64_t pairs = static_cast<uint64_t>(B) * np;
    const dim3 pgrid(static_cast<uint32_t>((pairs + 255) / 256));
    hipLaunchKernelGGL(k_fill32, dim3((qtiles * nslots + 255) / 256), dim3(256), 0, s, part, qtiles * nslots, 0u);
    if (join) hipLaunchKernelGGL(k_fill32, dim3((nlist + 255) / 256), dim3(256), 0, s, acnt, nlist, 0u);
    hipLaunchKernelGGL(k_inv_probe<false>, pgrid, dim3(256), 0, s, probes, npq, B, np, nlist, shift, seed_r0, seed,
                       list_off, part, nullptr, nullptr, join, acnt, nslots);
    // (INV_PG partial counts per slot: the counts become the counters'
    // cursors, at their offsets within the slot)
    hipLaunchKernelGGL(k_slot_totals, dim3(nblk), dim3(1024), 0, s, part, qtiles, nslots, nlist, list_off, qc, rt, qsq, qsrt, seed,
                       gcnt, nit, bsum, state + nstate, join, acnt, gbv);
    hipLaunchKernelGGL(k_slot_scan, dim3(nblk), dim3(1024), 0, s, gcnt, nit, bsum, nslots, qoff, item_off, totals);
    hipLaunchKernelGGL(k_inv_probe<true>, pgrid, dim3(256), 0, s, probes, npq, B, np, nlist, shift, seed_r0, seed,
                       list_off, part, qoff, qlist, join, acnt, nslots);
    hipLaunchKernelGGL(k_query_init, dim3((B + 255) / 256), dim3(256), 0, s, B, pool_cnt, pool_sorted, thr, lbound);
    return;
  }
  inv_pass_launch<false>(nb, tiles, s, dist, probes, probe_d, npq, B, np, nlist, shift, seed_r0, seed, list_off, part,
                         nullptr, nullptr);
  hipLaunchKernelGGL(k_slot_totals, dim3(nblk), dim3(1024), 0, s, part, qtiles, nslots, nlist, list_off, qc, rt, qsq, qsrt, seed,
                     gcnt, nit, bsum, state + nstate, 0u, nullptr, nullptr);
  hipLaunchKernelGGL(k_slot_scan, dim3(nblk), dim3(1024), 0, s, gcnt, nit, bsum, nslots, qoff, item_off, totals);
  inv_pass_launch<true>(nb, tiles, s, dist, probes, probe_d, npq, B, np, nlist, shift, seed_r0, seed, list_off, part,
                        qoff, qlist);
  hipLaunchKernelGGL(k_query_init, dim3((B + 255) / 256), dim3(256), 0, s, B, pool_cnt, pool_sorted, thr,
                     lbound);
}

void launch_sq16_pairs(const int8_t *qh, const int8_t *ql, const int64_t *qn, const int8_t *xh, const int8_t *xl,
                       const int64_t *xn, uint32_t m, uint32_t n, uint32_t dc, uint32_t qstride, uint32_t xstride,
                       int full, double step_sq, float *out, hipStream_t s) {
  const uint32_t tiles = ((m + 15) / 16) * ((n + 15) / 16);
  hipLaunchKernelGGL(k_sq16_pairs, dim3((tiles + 3) / 4), dim3(256), 0, s, qh, ql, qn, xh, xl, xn, m, n, dc,
                     qstride, xstride, full, step_sq, out);
}

}  // namespace katoann_hip
