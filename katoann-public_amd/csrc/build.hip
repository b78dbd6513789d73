// Index build on the GPU: the partition of the vectors into nlist lists by
// k-means and the PCA rotation, written as the reference's index files
// (Index::build, src/index.cpp:419-545; k-means in src/kmeans.cpp). The search
// path reads only those files (or the arrays katoann_build_device returns), so
// the method here is this library's own; what is fixed is the file format, the
// options and the errors.
//
// Method:
//  * training sample: train_size distinct rows, uniform, in ascending order, by
//    selection sampling (Knuth, TAOCP vol. 2, 3.4.2, Algorithm S) over a
//    SplitMix64 stream of the seed (seed 0 draws a seed from the OS);
//  * initial centres: nlist distinct training rows (partial Fisher-Yates);
//  * Lloyd iterations: every training row goes to its nearest centre (k_assign:
//    fp32 MFMA dot products fused with the arg-min, ties to the lowest centre
//    index); a centre becomes the fp64 mean of its rows, summed in ascending row
//    order, so the result does not depend on scheduling; a centre left without
//    rows stays where it was;
//  * every vector is assigned with the same kernel; the lists hold their ids in
//    ascending order (stable radix sort of the ids by list);
//  * rotation: eigenvectors of the fp64 covariance of a uniform sample of
//    max(rotation_sample, 1024) rows (rocSOLVER dsyevd, loaded on first use),
//    rows by decreasing eigenvalue, each signed so that its largest-magnitude
//    component (the first such) is positive.
//
// Distances in the assignment: |c|^2 - 2 x.c in fp32 (|x|^2 is common to the
// row). A row whose distances are all NaN or infinite goes to list 0.
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocsolver/rocsolver.h>

#include "../../include/katoann_c.h"

extern std::string &katoann_error_slot();

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

void ck(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
#define CK(x) ck((x), #x)

template <class T>
struct Dev {
  T *p = nullptr;
  Dev() = default;
  explicit Dev(size_t n) { alloc(n); }
  Dev(const Dev &) = delete;
  Dev &operator=(const Dev &) = delete;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t n) {
    if (p) (void)hipFree(p);
    p = nullptr;
    CK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  }
};

struct Pinned {
  void *p = nullptr;
  explicit Pinned(size_t bytes) { CK(hipHostMalloc(&p, bytes, hipHostMallocDefault)); }
  Pinned(const Pinned &) = delete;
  Pinned &operator=(const Pinned &) = delete;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};

struct Stream {
  hipStream_t s = nullptr;
  Stream() { CK(hipStreamCreate(&s)); }
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

uint32_t round_up(uint32_t v, uint32_t m) { return (v + m - 1) / m * m; }

// ---- random streams --------------------------------------------------------

// SplitMix64 (Steele, Lea and Flood, OOPSLA 2014).
struct SplitMix64 {
  uint64_t x;
  explicit SplitMix64(uint64_t seed) : x(seed) {}
  uint64_t next() {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double unit() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0, 1)
  uint64_t below(uint64_t m) { return static_cast<uint64_t>((static_cast<unsigned __int128>(next()) * m) >> 64); }
};

// k distinct indices of [0, n), uniform, ascending (Algorithm S).
std::vector<uint32_t> choose_sorted(uint64_t n, uint64_t k, SplitMix64 &g) {
  std::vector<uint32_t> out;
  out.reserve(k);
  for (uint64_t i = 0; i < n && out.size() < k; ++i)
    if (static_cast<double>(n - i) * g.unit() < static_cast<double>(k - out.size()))
      out.push_back(static_cast<uint32_t>(i));
  return out;
}

// k distinct indices of [0, n), uniform, in draw order.
std::vector<uint32_t> choose_unordered(uint32_t n, uint32_t k, SplitMix64 &g) {
  std::vector<uint32_t> p(n);
  for (uint32_t i = 0; i < n; ++i) p[i] = i;
  for (uint32_t i = 0; i < k; ++i) std::swap(p[i], p[i + g.below(n - i)]);
  p.resize(k);
  return p;
}

// ---- nearest centre: fp32 MFMA distances fused with the arg-min -------------
//
// One block of 256 threads takes 128 rows and walks all centres in blocks of
// 128, the dimension in stages of 32 (double-buffered LDS, one barrier per
// stage). Wave w computes rows (w & 1) * 64 .. +64 against centres (w >> 1) * 64
// .. +64 of the block as 4 x 4 tiles of v_mfma_f32_16x16x4f32 with the centres
// as the A operand, so lane l holds, for row (l & 15) of each row tile, four
// centres (4 (l >> 4) + r) of each centre tile and keeps its running minimum
// without any exchange until the end. Inside a stage, lane group g = l >> 4
// feeds dimension g * 8 + j to the j-th product, so each lane reads 8
// consecutive floats per row with two 16-byte LDS loads (the sum over the 32
// dimensions is the same set of terms).
//
// Centres are padded to a multiple of 128 rows (norm +inf: never chosen) and of
// 32 dimensions (zeros); rows past n and dimensions past d read as zero.
constexpr int kAR = 128, kAC = 128, kAK = 32, kALD = kAK + 4;

template <bool V4>
__global__ __launch_bounds__(256, 2) void k_assign(const float *__restrict__ x, uint64_t ldx, uint32_t n, uint32_t d,
                                                   const float *__restrict__ cen, uint32_t ldc,
                                                   const float *__restrict__ cn, uint32_t ncb, uint32_t nks,
                                                   uint32_t *__restrict__ out) {
  __shared__ __align__(16) float xs[2][kAR * kALD];
  __shared__ __align__(16) float cs[2][kAC * kALD];
  __shared__ float red_d[kAR];
  __shared__ uint32_t red_i[kAR];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int g = lane >> 4, li = lane & 15;
  const int wr = (w & 1) * 64, wc = (w >> 1) * 64;
  const uint64_t row0 = static_cast<uint64_t>(blockIdx.x) * kAR;
  const int sr = t >> 3, sq = (t & 7) * 4;  // staging: rows sr + 32 i, floats sq .. sq + 3
  v4f gx[4], gc[4];

  auto fetch = [&](uint32_t s) {
    const uint32_t cb = s / nks, ks = s - cb * nks;
    const uint32_t k0 = ks * kAK + sq;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint64_t row = row0 + sr + 32 * i;
      v4f v = {0.f, 0.f, 0.f, 0.f};
      if (row < n) {
        const float *p = x + row * ldx + k0;
        if (V4) {
          if (k0 < d) v = *reinterpret_cast<const v4f *>(p);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (k0 + e < d) v[e] = p[e];
        }
      }
      gx[i] = v;
      gc[i] = *reinterpret_cast<const v4f *>(cen + static_cast<uint64_t>(cb * kAC + sr + 32 * i) * ldc + k0);
    }
  };
  auto stash = [&](int b) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      *reinterpret_cast<v4f *>(&xs[b][(sr + 32 * i) * kALD + sq]) = gx[i];
      *reinterpret_cast<v4f *>(&cs[b][(sr + 32 * i) * kALD + sq]) = gc[i];
    }
  };

  float bd[4];
  uint32_t bi[4];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    bd[rt] = INFINITY;
    bi[rt] = 0;
  }
  v4f acc[4][4];
  const uint32_t S = ncb * nks;
  fetch(0);
  stash(0);
  __syncthreads();
  for (uint32_t s = 0; s < S; ++s) {
    const int b = s & 1;
    if (s + 1 < S) fetch(s + 1);
    const uint32_t cb = s / nks, ks = s - cb * nks;
    if (ks == 0) {
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) acc[ct][rt] = v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      v4f a[4], y[4];
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
        a[ct] = *reinterpret_cast<const v4f *>(&cs[b][(wc + ct * 16 + li) * kALD + g * 8 + 4 * h]);
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
        y[rt] = *reinterpret_cast<const v4f *>(&xs[b][(wr + rt * 16 + li) * kALD + g * 8 + 4 * h]);
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
          for (int rt = 0; rt < 4; ++rt)
            acc[ct][rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ct][e], y[rt][e], acc[ct][rt], 0, 0, 0);
    }
    if (ks == nks - 1) {
      // centres in increasing index order per lane: the strict < keeps the lowest
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t c = cb * kAC + wc + ct * 16 + 4 * g + r;
          const float cnv = cn[c];
#pragma unroll
          for (int rt = 0; rt < 4; ++rt) {
            const float dd = cnv - 2.0f * acc[ct][rt][r];
            if (dd < bd[rt]) {
              bd[rt] = dd;
              bi[rt] = c;
            }
          }
        }
    }
    if (s + 1 < S) stash(b ^ 1);
    __syncthreads();
  }
  // lanes l, l^16, l^32, l^48 hold the same row
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int off = 16; off <= 32; off <<= 1) {
      const float od = __shfl_xor(bd[rt], off);
      const uint32_t oi = __shfl_xor(bi[rt], off);
      if (od < bd[rt] || (od == bd[rt] && oi < bi[rt])) {
        bd[rt] = od;
        bi[rt] = oi;
      }
    }
  // waves w and w ^ 2 hold the same rows over the two halves of every centre block
  if ((w >> 1) == 1 && g == 0)
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      red_d[wr + rt * 16 + li] = bd[rt];
      red_i[wr + rt * 16 + li] = bi[rt];
    }
  __syncthreads();
  if ((w >> 1) == 0 && g == 0)
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      const int r = wr + rt * 16 + li;
      const float od = red_d[r];
      const uint32_t oi = red_i[r];
      uint32_t best = bi[rt];
      if (od < bd[rt] || (od == bd[rt] && oi < best)) best = oi;
      if (row0 + r < n) out[row0 + r] = best;
    }
}

// |c|^2 of the padded centres; +inf for the padding rows.
__global__ void k_centre_norms(const float *__restrict__ cen, uint32_t ldc, uint32_t nlist, uint32_t nlp,
                               float *__restrict__ cn) {
  const uint32_t c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= nlp) return;
  if (c >= nlist) {
    if (lane == 0) cn[c] = INFINITY;
    return;
  }
  float s = 0.f;
  for (uint32_t k = lane; k < ldc; k += 64) {
    const float v = cen[static_cast<uint64_t>(c) * ldc + k];
    s += v * v;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) cn[c] = s;
}

__global__ void k_iota(uint32_t *__restrict__ v, uint64_t n) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] = static_cast<uint32_t>(i);
}

// offs[c] = first position of list c in the sorted keys; offs[nlist] = n.
__global__ void k_bounds(const uint32_t *__restrict__ key, uint64_t n, uint32_t nlist, uint32_t *__restrict__ offs) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i > n) return;
  const int64_t prev = i == 0 ? -1 : static_cast<int64_t>(key[i - 1]);
  const int64_t cur = i == n ? static_cast<int64_t>(nlist) : static_cast<int64_t>(key[i]);
  for (int64_t c = prev + 1; c <= cur; ++c) offs[c] = static_cast<uint32_t>(i);
}

// rows[ids[i]] -> dst[i]
__global__ void k_gather(const float *__restrict__ src, uint64_t ld, const uint32_t *__restrict__ ids, uint32_t m,
                         uint32_t d, float *__restrict__ dst) {
  const uint32_t i = blockIdx.x;
  if (i >= m) return;
  const float *s = src + static_cast<uint64_t>(ids[i]) * ld;
  float *o = dst + static_cast<uint64_t>(i) * d;
  for (uint32_t k = threadIdx.x; k < d; k += blockDim.x) o[k] = s[k];
}

// padded centres -> nlist x d
__global__ void k_unpad(const float *__restrict__ cen, uint32_t ldc, uint32_t nlist, uint32_t d, float *__restrict__ out) {
  const uint32_t c = blockIdx.x;
  if (c >= nlist) return;
  for (uint32_t k = threadIdx.x; k < d; k += blockDim.x)
    out[static_cast<uint64_t>(c) * d + k] = cen[static_cast<uint64_t>(c) * ldc + k];
}

// nlist x d -> padded centres (zeros past d)
__global__ void k_pad(const float *__restrict__ in, uint32_t nlist, uint32_t d, float *__restrict__ cen, uint32_t ldc) {
  const uint32_t c = blockIdx.x;
  if (c >= nlist) return;
  for (uint32_t k = threadIdx.x; k < ldc; k += blockDim.x)
    cen[static_cast<uint64_t>(c) * ldc + k] = k < d ? in[static_cast<uint64_t>(c) * d + k] : 0.f;
}

__global__ void k_sizes(const uint32_t *__restrict__ offs, uint32_t nlist, uint32_t *__restrict__ sizes) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nlist) sizes[c] = offs[c + 1] - offs[c];
}

// ---- centre update: fp64 sums in ascending row order ------------------------
//
// A list is cut into segments of at most kSeg rows. A list of one segment is
// averaged by that segment's block; a longer one writes each segment's sums
// (slot >= 0) and k_merge_segments adds them in segment order.
constexpr uint32_t kSeg = 1024;

struct Segment {
  uint32_t list, lo, hi;
  int32_t slot;
};
struct Merge {
  uint32_t list, first, count, rows;
};

__global__ __launch_bounds__(256) void k_segment_sums(const float *__restrict__ x, uint64_t ld, uint32_t d,
                                                      const uint32_t *__restrict__ order,
                                                      const Segment *__restrict__ segs, float *__restrict__ cen,
                                                      uint32_t ldc, double *__restrict__ part) {
  const Segment sg = segs[blockIdx.x];
  const double inv = 1.0 / static_cast<double>(sg.hi - sg.lo);
  for (uint32_t k0 = 0; k0 < d; k0 += 4 * 256) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    uint32_t i = sg.lo;
    for (; i + 4 <= sg.hi; i += 4) {  // four rows in flight, added in row order
      float v[4][4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float *row = x + static_cast<uint64_t>(order[i + u]) * ld;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t k = k0 + threadIdx.x + 256 * q;
          v[u][q] = k < d ? row[k] : 0.f;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] += static_cast<double>(v[u][q]);
    }
    for (; i < sg.hi; ++i) {
      const float *row = x + static_cast<uint64_t>(order[i]) * ld;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t k = k0 + threadIdx.x + 256 * q;
        if (k < d) acc[q] += static_cast<double>(row[k]);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t k = k0 + threadIdx.x + 256 * q;
      if (k >= d) continue;
      if (sg.slot < 0) cen[static_cast<uint64_t>(sg.list) * ldc + k] = static_cast<float>(acc[q] * inv);
      else part[static_cast<uint64_t>(sg.slot) * d + k] = acc[q];
    }
  }
}

__global__ void k_merge_segments(const double *__restrict__ part, uint32_t d, const Merge *__restrict__ merges,
                                 float *__restrict__ cen, uint32_t ldc) {
  const Merge m = merges[blockIdx.x];
  const double inv = 1.0 / static_cast<double>(m.rows);
  for (uint32_t k = threadIdx.x; k < d; k += blockDim.x) {
    double s = 0.0;
    for (uint32_t j = 0; j < m.count; ++j) s += part[static_cast<uint64_t>(m.first + j) * d + k];
    cen[static_cast<uint64_t>(m.list) * ldc + k] = static_cast<float>(s * inv);
  }
}

// ---- PCA ---------------------------------------------------------------------

// part[p * d + k] = sum of column k over rows [p R, (p + 1) R)
__global__ void k_column_sums(const float *__restrict__ x, uint32_t m, uint32_t d, uint32_t R,
                              double *__restrict__ part) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x, p = blockIdx.y;
  if (k >= d) return;
  double s = 0.0;
  const uint32_t hi = std::min<uint64_t>(m, static_cast<uint64_t>(p + 1) * R);
  for (uint32_t r = p * R; r < hi; ++r) s += static_cast<double>(x[static_cast<uint64_t>(r) * d + k]);
  part[static_cast<uint64_t>(p) * d + k] = s;
}

__global__ void k_column_means(const double *__restrict__ part, uint32_t P, uint32_t m, uint32_t d,
                               double *__restrict__ mean) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d) return;
  double s = 0.0;
  for (uint32_t p = 0; p < P; ++p) s += part[static_cast<uint64_t>(p) * d + k];
  mean[k] = s / static_cast<double>(m);
}

// part[p][i][j] = sum over rows [p R, (p + 1) R) of (x_i - mean_i)(x_j - mean_j):
// 64 x 64 output tiles, 16 x 16 threads of 4 x 4 outputs each (strided, so the
// LDS reads of a wave spread over the banks). Both triangles are computed, each
// entry and its mirror from the same products in the same order, so the matrix
// is exactly symmetric.
__global__ __launch_bounds__(256) void k_covariance(const float *__restrict__ x, uint32_t m, uint32_t d,
                                                    const double *__restrict__ mean, uint32_t R,
                                                    double *__restrict__ part) {
  __shared__ double si[16][64];
  __shared__ double sj[16][64];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const uint32_t i0 = blockIdx.y * 64, j0 = blockIdx.x * 64, p = blockIdx.z;
  const uint32_t lo = p * R, hi = std::min<uint64_t>(m, static_cast<uint64_t>(p + 1) * R);
  double acc[4][4] = {};
  for (uint32_t r0 = lo; r0 < hi; r0 += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int rr = e >> 6, cc = e & 63;
      const uint32_t r = r0 + rr;
      const uint32_t ci = i0 + cc, cj = j0 + cc;
      si[rr][cc] = (r < hi && ci < d) ? static_cast<double>(x[static_cast<uint64_t>(r) * d + ci]) - mean[ci] : 0.0;
      sj[rr][cc] = (r < hi && cj < d) ? static_cast<double>(x[static_cast<uint64_t>(r) * d + cj]) - mean[cj] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = si[rr][ty + 16 * u];
        b[u] = sj[rr][tx + 16 * u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const uint32_t i = i0 + ty + 16 * u, j = j0 + tx + 16 * v;
      if (i < d && j < d) part[(static_cast<uint64_t>(p) * d + i) * d + j] = acc[u][v];
    }
}

__global__ void k_covariance_reduce(const double *__restrict__ part, uint32_t P, uint32_t m, uint64_t dd,
                                    double *__restrict__ a) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= dd) return;
  double s = 0.0;
  for (uint32_t p = 0; p < P; ++p) s += part[p * dd + e];
  a[e] = s / static_cast<double>(m);
}

// Row r of the rotation = eigenvector of the (d - 1 - r)-th smallest eigenvalue
// (column d - 1 - r of dsyevd's column-major output), signed so that its first
// largest-magnitude component is positive.
__global__ __launch_bounds__(256) void k_rotation_rows(const double *__restrict__ vec, uint32_t d,
                                                       float *__restrict__ rot) {
  __shared__ double best_v[256];
  __shared__ uint32_t best_k[256];
  const uint32_t r = blockIdx.x;
  const double *v = vec + static_cast<uint64_t>(d - 1 - r) * d;
  double bv = -1.0;
  uint32_t bk = 0;
  for (uint32_t k = threadIdx.x; k < d; k += 256) {
    const double a = fabs(v[k]);
    if (a > bv) {
      bv = a;
      bk = k;
    }
  }
  best_v[threadIdx.x] = bv;
  best_k[threadIdx.x] = bk;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const double ov = best_v[threadIdx.x + s];
      const uint32_t ok = best_k[threadIdx.x + s];
      if (ov > best_v[threadIdx.x] || (ov == best_v[threadIdx.x] && ok < best_k[threadIdx.x])) {
        best_v[threadIdx.x] = ov;
        best_k[threadIdx.x] = ok;
      }
    }
    __syncthreads();
  }
  const double sign = v[best_k[0]] < 0.0 ? -1.0 : 1.0;
  for (uint32_t k = threadIdx.x; k < d; k += 256)
    rot[static_cast<uint64_t>(r) * d + k] = static_cast<float>(sign * v[k]);
}

// ---- rocSOLVER, loaded on first use -------------------------------------------

struct Solver {
  rocblas_status (*create)(rocblas_handle *) = nullptr;
  rocblas_status (*destroy)(rocblas_handle) = nullptr;
  rocblas_status (*set_stream)(rocblas_handle, hipStream_t) = nullptr;
  rocblas_status (*dsyevd)(rocblas_handle, rocblas_evect, rocblas_fill, rocblas_int, double *, rocblas_int, double *,
                           double *, rocblas_int *) = nullptr;
  std::string error;
};

const Solver &solver() {
  static Solver s;
  static std::once_flag once;
  std::call_once(once, [] {
    void *blas = dlopen("librocblas.so.5", RTLD_NOW | RTLD_GLOBAL);
    if (!blas) blas = dlopen("librocblas.so", RTLD_NOW | RTLD_GLOBAL);
    void *sol = dlopen("librocsolver.so.0", RTLD_NOW | RTLD_GLOBAL);
    if (!sol) sol = dlopen("librocsolver.so", RTLD_NOW | RTLD_GLOBAL);
    if (!blas || !sol) {
      s.error = std::string("cannot load rocBLAS/rocSOLVER: ") + dlerror();
      return;
    }
    s.create = reinterpret_cast<decltype(s.create)>(dlsym(blas, "rocblas_create_handle"));
    s.destroy = reinterpret_cast<decltype(s.destroy)>(dlsym(blas, "rocblas_destroy_handle"));
    s.set_stream = reinterpret_cast<decltype(s.set_stream)>(dlsym(blas, "rocblas_set_stream"));
    s.dsyevd = reinterpret_cast<decltype(s.dsyevd)>(dlsym(sol, "rocsolver_dsyevd"));
    if (!s.create || !s.destroy || !s.set_stream || !s.dsyevd) s.error = "rocBLAS/rocSOLVER lack a needed symbol";
  });
  return s;
}

// ---- the build -----------------------------------------------------------------

// Where the vectors are: host memory (streamed through pinned buffers) or HBM.
struct Rows {
  const float *host = nullptr;
  const float *dev = nullptr;
  uint32_t n = 0, d = 0;
};

constexpr size_t kStageBytes = 64ull << 20;

// rows ids (ascending or not) -> dst (m x d, device)
void gather_rows(const Rows &src, const std::vector<uint32_t> &ids, float *dst, hipStream_t s) {
  const uint32_t m = static_cast<uint32_t>(ids.size()), d = src.d;
  if (!m) return;
  if (src.dev) {
    Dev<uint32_t> di(m);
    CK(hipMemcpyAsync(di.p, ids.data(), m * 4ull, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_gather, dim3(m), dim3(256), 0, s, src.dev, static_cast<uint64_t>(d), di.p, m, d, dst);
    CK(hipGetLastError());
    CK(hipStreamSynchronize(s));
    return;
  }
  const size_t per = std::max<size_t>(1, kStageBytes / (d * 4ull));
  Pinned buf[2] = {Pinned(per * d * 4), Pinned(per * d * 4)};
  hipEvent_t ev[2];
  CK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  CK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  bool used[2] = {false, false};
  int b = 0;
  for (size_t i0 = 0; i0 < m; i0 += per, b ^= 1) {
    const size_t cnt = std::min<size_t>(per, m - i0);
    if (used[b]) CK(hipEventSynchronize(ev[b]));
    float *h = static_cast<float *>(buf[b].p);
    for (size_t i = 0; i < cnt; ++i)
      std::memcpy(h + i * d, src.host + static_cast<uint64_t>(ids[i0 + i]) * d, d * 4ull);
    CK(hipMemcpyAsync(dst + i0 * d, h, cnt * d * 4ull, hipMemcpyHostToDevice, s));
    CK(hipEventRecord(ev[b], s));
    used[b] = true;
  }
  CK(hipStreamSynchronize(s));
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
}

struct Centres {
  uint32_t nlist = 0, d = 0, nlp = 0, ldc = 0;
  Dev<float> c, cn;
  Centres(uint32_t nl, uint32_t dim) : nlist(nl), d(dim), nlp(round_up(nl, kAC)), ldc(round_up(dim, kAK)) {
    c.alloc(static_cast<size_t>(nlp) * ldc);
    cn.alloc(nlp);
    CK(hipMemset(c.p, 0, static_cast<size_t>(nlp) * ldc * 4));
  }
  void norms(hipStream_t s) {
    hipLaunchKernelGGL(k_centre_norms, dim3((nlp + 3) / 4), dim3(256), 0, s, c.p, ldc, nlist, nlp, cn.p);
    CK(hipGetLastError());
  }
};

void launch_assign(const float *x, uint64_t ldx, uint64_t n, uint32_t d, const Centres &C, uint32_t *out,
                   hipStream_t s) {
  const bool v4 = d % 4 == 0 && ldx % 4 == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0;
  const uint32_t ncb = C.nlp / kAC, nks = C.ldc / kAK;
  const uint64_t max_rows = static_cast<uint64_t>(1u << 30) * kAR;  // grid.x stays far below 2^31
  for (uint64_t r0 = 0; r0 < n; r0 += max_rows) {
    const uint32_t nr = static_cast<uint32_t>(std::min<uint64_t>(max_rows, n - r0));
    const dim3 grid((nr + kAR - 1) / kAR);
    if (v4)
      hipLaunchKernelGGL(k_assign<true>, grid, dim3(256), 0, s, x + r0 * ldx, ldx, nr, d, C.c.p, C.ldc, C.cn.p, ncb,
                         nks, out + r0);
    else
      hipLaunchKernelGGL(k_assign<false>, grid, dim3(256), 0, s, x + r0 * ldx, ldx, nr, d, C.c.p, C.ldc, C.cn.p, ncb,
                         nks, out + r0);
    CK(hipGetLastError());
  }
}

// Stable sort of row ids by list: ids (ascending inside each list) and offsets.
struct ListSort {
  uint64_t n = 0;
  uint32_t nlist = 0, bits = 1;
  Dev<uint32_t> keys_out, vals_in, vals_out, offs;
  Dev<char> temp;
  size_t temp_bytes = 0;
  ListSort(uint64_t rows, uint32_t nl, hipStream_t s) : n(rows), nlist(nl) {
    while (bits < 32 && (1ull << bits) < nl) ++bits;
    keys_out.alloc(n);
    vals_in.alloc(n);
    vals_out.alloc(n);
    offs.alloc(static_cast<size_t>(nl) + 1);
    if (n) {
      hipLaunchKernelGGL(k_iota, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, s, vals_in.p, n);
      CK(hipGetLastError());
      CK(rocprim::radix_sort_pairs(nullptr, temp_bytes, vals_in.p, keys_out.p, vals_in.p, vals_out.p, n, 0, bits, s));
      temp.alloc(temp_bytes);
    }
  }
  void run(const uint32_t *keys, hipStream_t s) {
    if (n) CK(rocprim::radix_sort_pairs(temp.p, temp_bytes, keys, keys_out.p, vals_in.p, vals_out.p, n, 0, bits, s));
    hipLaunchKernelGGL(k_bounds, dim3(static_cast<uint32_t>((n + 1 + 255) / 256)), dim3(256), 0, s, keys_out.p, n, nlist,
                       offs.p);
    CK(hipGetLastError());
  }
};

// Centres <- mean of their rows (x: the rows the sort ran over).
void update_centres(const float *x, uint32_t d, const ListSort &ls, Centres &C, hipStream_t s) {
  std::vector<uint32_t> offs(ls.nlist + 1);
  CK(hipMemcpyAsync(offs.data(), ls.offs.p, offs.size() * 4, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  std::vector<Segment> segs;
  std::vector<Merge> merges;
  int32_t slots = 0;
  for (uint32_t c = 0; c < ls.nlist; ++c) {
    const uint32_t lo = offs[c], hi = offs[c + 1];
    if (lo == hi) continue;
    if (hi - lo <= kSeg) {
      segs.push_back({c, lo, hi, -1});
      continue;
    }
    merges.push_back({c, static_cast<uint32_t>(slots), 0, hi - lo});
    for (uint32_t a = lo; a < hi; a += kSeg, ++merges.back().count)
      segs.push_back({c, a, std::min(hi, a + kSeg), slots++});
  }
  if (segs.empty()) return;
  Dev<Segment> dsegs(segs.size());
  Dev<Merge> dmerges(merges.size());
  Dev<double> part(static_cast<size_t>(slots) * d);
  CK(hipMemcpyAsync(dsegs.p, segs.data(), segs.size() * sizeof(Segment), hipMemcpyHostToDevice, s));
  if (!merges.empty())
    CK(hipMemcpyAsync(dmerges.p, merges.data(), merges.size() * sizeof(Merge), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_segment_sums, dim3(static_cast<uint32_t>(segs.size())), dim3(256), 0, s, x,
                     static_cast<uint64_t>(d), d, ls.vals_out.p, dsegs.p, C.c.p, C.ldc, part.p);
  CK(hipGetLastError());
  if (!merges.empty()) {
    hipLaunchKernelGGL(k_merge_segments, dim3(static_cast<uint32_t>(merges.size())), dim3(256), 0, s, part.p, d,
                       dmerges.p, C.c.p, C.ldc);
    CK(hipGetLastError());
  }
  CK(hipStreamSynchronize(s));
}

// Lloyd iterations on device-resident training rows (ts x d, contiguous).
void lloyd(const float *train, uint32_t ts, uint32_t d, Centres &C, uint32_t iters, hipStream_t s) {
  if (!iters) return;
  Dev<uint32_t> a(ts);
  ListSort ls(ts, C.nlist, s);
  for (uint32_t it = 0; it < iters; ++it) {
    C.norms(s);
    launch_assign(train, d, ts, d, C, a.p, s);
    ls.run(a.p, s);
    update_centres(train, d, ls, C, s);
  }
}

// Nearest centre of every vector -> out (device, n).
void assign_all(const Rows &src, Centres &C, uint32_t *out, hipStream_t s) {
  C.norms(s);
  const uint32_t d = src.d;
  if (src.dev) {
    launch_assign(src.dev, d, src.n, d, C, out, s);
    CK(hipStreamSynchronize(s));
    return;
  }
  const size_t per = std::max<size_t>(kAR, kStageBytes / (d * 4ull) / kAR * kAR);
  Pinned hb[2] = {Pinned(per * d * 4), Pinned(per * d * 4)};
  Dev<float> db[2];
  db[0].alloc(per * d);
  db[1].alloc(per * d);
  hipEvent_t ev[2];
  CK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  CK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  bool used[2] = {false, false};
  int b = 0;
  for (uint64_t r0 = 0; r0 < src.n; r0 += per, b ^= 1) {
    const uint64_t cnt = std::min<uint64_t>(per, src.n - r0);
    if (used[b]) CK(hipEventSynchronize(ev[b]));
    std::memcpy(hb[b].p, src.host + r0 * d, cnt * d * 4);
    CK(hipMemcpyAsync(db[b].p, hb[b].p, cnt * d * 4, hipMemcpyHostToDevice, s));
    launch_assign(db[b].p, d, cnt, d, C, out + r0, s);
    CK(hipEventRecord(ev[b], s));
    used[b] = true;
  }
  CK(hipStreamSynchronize(s));
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
}

// PCA rotation of a device sample (m x d) -> rot (device, d x d); eigenvalues
// in decreasing order -> eig (host, optional). False when m < 2 or d < 2.
bool pca_rotation(const float *sample, uint32_t m, uint32_t d, float *rot, double *eig, hipStream_t s) {
  if (m < 2 || d < 2) return false;
  const Solver &sv = solver();
  if (!sv.error.empty()) throw std::runtime_error(sv.error);
  const uint32_t P1 = std::max<uint32_t>(1, std::min<uint32_t>(256, m / 256));
  const uint32_t R1 = (m + P1 - 1) / P1;
  Dev<double> csum(static_cast<size_t>(P1) * d), mean(d);
  hipLaunchKernelGGL(k_column_sums, dim3((d + 255) / 256, P1), dim3(256), 0, s, sample, m, d, R1, csum.p);
  hipLaunchKernelGGL(k_column_means, dim3((d + 255) / 256), dim3(256), 0, s, csum.p, P1, m, d, mean.p);
  CK(hipGetLastError());
  const uint32_t nt = (d + 63) / 64;
  const uint32_t P2 = std::max<uint32_t>(1, std::min<uint32_t>((m + 2047) / 2048, 4096 / (nt * nt)));
  const uint32_t R2 = (m + P2 - 1) / P2;
  const uint64_t dd = static_cast<uint64_t>(d) * d;
  Dev<double> part(P2 * dd), a(dd), w(d), e(d);
  Dev<rocblas_int> info(1);
  hipLaunchKernelGGL(k_covariance, dim3(nt, nt, P2), dim3(256), 0, s, sample, m, d, mean.p, R2, part.p);
  hipLaunchKernelGGL(k_covariance_reduce, dim3(static_cast<uint32_t>((dd + 255) / 256)), dim3(256), 0, s, part.p, P2,
                     m, dd, a.p);
  CK(hipGetLastError());
  rocblas_handle h = nullptr;
  if (sv.create(&h) != rocblas_status_success) throw std::runtime_error("rocblas_create_handle failed");
  rocblas_status st = sv.set_stream(h, s);
  if (st == rocblas_status_success)
    st = sv.dsyevd(h, rocblas_evect_original, rocblas_fill_upper, static_cast<rocblas_int>(d), a.p,
                   static_cast<rocblas_int>(d), w.p, e.p, info.p);
  rocblas_int hinfo = 0;
  if (st == rocblas_status_success) {
    CK(hipMemcpyAsync(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
  }
  sv.destroy(h);
  if (st != rocblas_status_success) throw std::runtime_error("rocsolver_dsyevd failed");
  if (hinfo != 0) throw std::runtime_error("eigensolver did not converge");
  hipLaunchKernelGGL(k_rotation_rows, dim3(d), dim3(256), 0, s, a.p, d, rot);
  CK(hipGetLastError());
  if (eig) {
    std::vector<double> hw(d);
    CK(hipMemcpyAsync(hw.data(), w.p, d * 8ull, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < d; ++i) eig[i] = hw[d - 1 - i];
  }
  CK(hipStreamSynchronize(s));
  return true;
}

uint64_t seed_of(uint64_t seed) {
  if (seed) return seed;
  std::random_device rd;
  return (static_cast<uint64_t>(rd()) << 32) ^ rd();
}

void check_build_args(uint32_t n, uint32_t dim, const katoann_build_options &o) {
  if (dim == 0 || dim > 4096) throw std::invalid_argument("dimension must be 1 to 4096");
  if (n == 0 || o.nlist == 0 || o.nlist > n) throw std::invalid_argument("need at least nlist vectors");
  if (std::min(o.train_size, n) > (1u << 23)) throw std::invalid_argument("train_size is limited to 8388608");
}

// k-means + full assignment + PCA. Outputs on the device: centroids (nlist x
// d), the sorted lists (ls.vals_out: ids grouped by list, ascending inside
// each; ls.offs), the rotation (d x d) when rot is given.
void partition(const Rows &src, const katoann_build_options &o, Centres &C, ListSort *&ls_out, Dev<uint32_t> &assign,
               float *rot, bool *rotated, hipStream_t s) {
  const uint32_t n = src.n, d = src.d;
  SplitMix64 g(seed_of(o.seed));
  const uint32_t ts = std::max(std::min(o.train_size, n), o.nlist);
  const std::vector<uint32_t> tids = choose_sorted(n, ts, g);
  const std::vector<uint32_t> init = choose_unordered(ts, o.nlist, g);
  {
    Dev<float> train(static_cast<size_t>(ts) * d);
    gather_rows(src, tids, train.p, s);
    {
      Dev<uint32_t> di(o.nlist);
      CK(hipMemcpyAsync(di.p, init.data(), o.nlist * 4ull, hipMemcpyHostToDevice, s));
      Dev<float> c0(static_cast<size_t>(o.nlist) * d);
      hipLaunchKernelGGL(k_gather, dim3(o.nlist), dim3(256), 0, s, train.p, static_cast<uint64_t>(d), di.p, o.nlist,
                         d, c0.p);
      hipLaunchKernelGGL(k_pad, dim3(o.nlist), dim3(256), 0, s, c0.p, o.nlist, d, C.c.p, C.ldc);
      CK(hipGetLastError());
      CK(hipStreamSynchronize(s));
    }
    lloyd(train.p, ts, d, C, o.kmeans_iters, s);
  }
  assign.alloc(n);
  assign_all(src, C, assign.p, s);
  ls_out = new ListSort(n, o.nlist, s);
  ls_out->run(assign.p, s);
  CK(hipStreamSynchronize(s));
  *rotated = false;
  if (rot) {
    const uint32_t m = std::min<uint32_t>(std::max<uint32_t>(o.rotation_sample, 1024u), n);
    const std::vector<uint32_t> pids = choose_sorted(n, m, g);
    Dev<float> sample(static_cast<size_t>(m) * d);
    gather_rows(src, pids, sample.p, s);
    *rotated = pca_rotation(sample.p, m, d, rot, nullptr, s);
  }
}

void write_or_throw(FILE *f, const void *p, size_t bytes, const std::string &name) {
  if (bytes && std::fwrite(p, 1, bytes, f) != bytes) {
    std::fclose(f);
    throw std::runtime_error("cannot write " + name);
  }
}

FILE *open_or_throw(const std::string &name) {
  FILE *f = std::fopen(name.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + name);
  std::setvbuf(f, nullptr, _IOFBF, 8 << 20);
  return f;
}

void close_or_throw(FILE *f, const std::string &name) {
  if (std::fclose(f) != 0) throw std::runtime_error("cannot write " + name);
}

void build_files(const float *data, uint32_t n, uint32_t dim, const katoann_build_options &o,
                 const std::string &prefix) {
  check_build_args(n, dim, o);
  Stream st;
  Rows src;
  src.host = data;
  src.n = n;
  src.d = dim;
  Centres C(o.nlist, dim);
  ListSort *lsp = nullptr;
  Dev<uint32_t> assign;
  Dev<float> rot(static_cast<size_t>(dim) * dim);
  bool rotated = false;
  partition(src, o, C, lsp, assign, rot.p, &rotated, st.s);
  std::unique_ptr<ListSort> ls(lsp);
  std::vector<float> cen(static_cast<size_t>(o.nlist) * dim), hrot;
  std::vector<uint32_t> order(n), offs(o.nlist + 1);
  {
    Dev<float> c(static_cast<size_t>(o.nlist) * dim);
    hipLaunchKernelGGL(k_unpad, dim3(o.nlist), dim3(256), 0, st.s, C.c.p, C.ldc, o.nlist, dim, c.p);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(cen.data(), c.p, cen.size() * 4, hipMemcpyDeviceToHost, st.s));
    CK(hipMemcpyAsync(order.data(), ls->vals_out.p, n * 4ull, hipMemcpyDeviceToHost, st.s));
    CK(hipMemcpyAsync(offs.data(), ls->offs.p, offs.size() * 4, hipMemcpyDeviceToHost, st.s));
    if (rotated) {
      hrot.resize(static_cast<size_t>(dim) * dim);
      CK(hipMemcpyAsync(hrot.data(), rot.p, hrot.size() * 4, hipMemcpyDeviceToHost, st.s));
    }
    CK(hipStreamSynchronize(st.s));
  }
  // file layouts: include/katoann/index.h:14-20, src/index.cpp:482-505, 535-545
  {
    const std::string name = prefix + ".meta";
    FILE *f = open_or_throw(name);
    const uint32_t meta[5] = {dim, o.nlist, n, 0u, o.normalized ? 1u : 0u};
    write_or_throw(f, meta, sizeof(meta), name);
    close_or_throw(f, name);
  }
  {
    const std::string name = prefix + ".centroids";
    FILE *f = open_or_throw(name);
    write_or_throw(f, cen.data(), cen.size() * 4, name);
    close_or_throw(f, name);
  }
  {
    const std::string name = prefix + ".invlist", nname = prefix + ".norms";
    FILE *f = open_or_throw(name);
    FILE *g = std::fopen(nname.c_str(), "wb");
    if (!g) {
      std::fclose(f);
      throw std::runtime_error("cannot write " + name);
    }
    for (uint32_t c = 0; c < o.nlist; ++c) {
      const uint32_t lo = offs[c], sz = offs[c + 1] - offs[c];
      write_or_throw(f, &sz, 4, name);
      write_or_throw(f, order.data() + lo, sz * 4ull, name);
      for (uint32_t i = 0; i < sz; ++i)
        write_or_throw(f, data + static_cast<uint64_t>(order[lo + i]) * dim, dim * 4ull, name);
      if (std::fwrite(&sz, 4, 1, g) != 1) {
        std::fclose(f);
        std::fclose(g);
        throw std::runtime_error("cannot write " + name);
      }
    }
    if (std::fclose(g) != 0) {
      std::fclose(f);
      throw std::runtime_error("cannot write " + name);
    }
    close_or_throw(f, name);
  }
  {
    const std::string name = prefix + ".dimc";
    FILE *f = open_or_throw(name);
    const uint32_t dc = round_up(dim, 64);
    const uint32_t head = std::min(std::max(64u, round_up(dim / 4, 64)), dc);  // default prefix plane
    const uint32_t hdr[8] = {0x434D4944u, 2u, dim, o.nlist, n, head, rotated ? 1u : 0u, 0u};
    write_or_throw(f, hdr, sizeof(hdr), name);
    if (rotated) write_or_throw(f, hrot.data(), hrot.size() * 4, name);
    close_or_throw(f, name);
  }
}

template <class F>
int guarded(F &&f) {
  try {
    f();
    return KATOANN_OK;
  } catch (const std::invalid_argument &e) {
    katoann_error_slot() = e.what();
    return KATOANN_INVALID_ARGUMENT;
  } catch (const std::exception &e) {
    katoann_error_slot() = e.what();
    return KATOANN_RUNTIME_ERROR;
  }
}

katoann_build_options options_or_default(const katoann_build_options *opt) {
  katoann_build_options o;
  katoann_default_build_options(&o);
  if (opt) o = *opt;
  return o;
}

}  // namespace

extern "C" int katoann_build(const float *data, uint32_t n, uint32_t dim, const katoann_build_options *opt,
                             const char *prefix) {
  return guarded([&] {
    if (!data || !prefix) throw std::invalid_argument("data and prefix are required");
    build_files(data, n, dim, options_or_default(opt), prefix);
  });
}

extern "C" int katoann_build_device(const float *data, uint32_t n, uint32_t dim, const katoann_build_options *opt,
                                    float *centroids, uint32_t *order, uint32_t *sizes, float *rotation,
                                    int32_t *rotated) {
  return guarded([&] {
    const katoann_build_options o = options_or_default(opt);
    if (!data || !centroids || !order || !sizes) throw std::invalid_argument("data and outputs are required");
    check_build_args(n, dim, o);
    CK(hipDeviceSynchronize());
    Stream st;
    Rows src;
    src.dev = data;
    src.n = n;
    src.d = dim;
    Centres C(o.nlist, dim);
    ListSort *lsp = nullptr;
    Dev<uint32_t> assign;
    bool rot = false;
    partition(src, o, C, lsp, assign, rotation, &rot, st.s);
    std::unique_ptr<ListSort> ls(lsp);
    hipLaunchKernelGGL(k_unpad, dim3(o.nlist), dim3(256), 0, st.s, C.c.p, C.ldc, o.nlist, dim, centroids);
    hipLaunchKernelGGL(k_sizes, dim3((o.nlist + 255) / 256), dim3(256), 0, st.s, ls->offs.p, o.nlist, sizes);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(order, ls->vals_out.p, n * 4ull, hipMemcpyDeviceToDevice, st.s));
    CK(hipStreamSynchronize(st.s));
    if (rotated) *rotated = rot ? 1 : 0;
  });
}

extern "C" int katoann_assign_lists(const float *x, uint32_t n, uint32_t dim, const float *centroids, uint32_t nlist,
                                    uint32_t *out) {
  return guarded([&] {
    if (dim == 0 || dim > 4096) throw std::invalid_argument("dimension must be 1 to 4096");
    if (nlist == 0) throw std::invalid_argument("nlist must be positive");
    if (!x || !centroids || !out) throw std::invalid_argument("data and outputs are required");
    CK(hipDeviceSynchronize());
    Stream st;
    Centres C(nlist, dim);
    hipLaunchKernelGGL(k_pad, dim3(nlist), dim3(256), 0, st.s, centroids, nlist, dim, C.c.p, C.ldc);
    CK(hipGetLastError());
    C.norms(st.s);
    launch_assign(x, dim, n, dim, C, out, st.s);
    CK(hipStreamSynchronize(st.s));
  });
}

extern "C" int katoann_kmeans_device(const float *train, uint32_t n, uint32_t dim, float *centroids, uint32_t nlist,
                                     uint32_t iters) {
  return guarded([&] {
    if (dim == 0 || dim > 4096) throw std::invalid_argument("dimension must be 1 to 4096");
    if (nlist == 0 || n == 0) throw std::invalid_argument("need at least one vector and one list");
    if (!train || !centroids) throw std::invalid_argument("data and outputs are required");
    CK(hipDeviceSynchronize());
    Stream st;
    Centres C(nlist, dim);
    hipLaunchKernelGGL(k_pad, dim3(nlist), dim3(256), 0, st.s, centroids, nlist, dim, C.c.p, C.ldc);
    CK(hipGetLastError());
    lloyd(train, n, dim, C, iters, st.s);
    hipLaunchKernelGGL(k_unpad, dim3(nlist), dim3(256), 0, st.s, C.c.p, C.ldc, nlist, dim, centroids);
    CK(hipGetLastError());
    CK(hipStreamSynchronize(st.s));
  });
}

extern "C" int katoann_pca_rotation_device(const float *sample, uint32_t m, uint32_t dim, float *rotation,
                                           double *eigenvalues, int32_t *rotated) {
  return guarded([&] {
    if (dim == 0 || dim > 4096) throw std::invalid_argument("dimension must be 1 to 4096");
    if (!sample || !rotation) throw std::invalid_argument("data and outputs are required");
    CK(hipDeviceSynchronize());
    Stream st;
    const bool r = pca_rotation(sample, m, dim, rotation, eigenvalues, st.s);
    if (rotated) *rotated = r ? 1 : 0;
  });
}
