// Learned nprobe: the kernels of a small MLP that reads each query's smallest
// sorted coarse distances and predicts how many lists the query needs. No
// counterpart in the reference (its search takes one nprobe per call,
// include/katoann/index.h:56-62).
//
// Arithmetic: GEMM operands in bf16 (round to nearest even), products summed
// in fp32 on gfx950's v_mfma_f32_16x16x32_bf16; activations, gradients and the
// Adam state in fp32, rounded to bf16 where they feed the next product.
#include <float.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "kernels.hpp"

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma_bf16(v8bf a, v8bf b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// 8 consecutive-k operand values (k0 .. k0 + 7) of row r: contiguous when the
// k stride is 1 and the address is 16-byte aligned, gathered otherwise; zero
// past K or past the row count.
__device__ __forceinline__ v8bf load8(const __bf16 *p, uint32_t r, uint32_t nr, uint32_t k0, uint32_t K, size_t sr,
                                      size_t sk) {
  v8bf v;
  if (r >= nr) {
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = static_cast<__bf16>(0.0f);
    return v;
  }
  const __bf16 *q = p + r * sr + k0 * sk;
  if (sk == 1 && k0 + 8 <= K && (reinterpret_cast<uintptr_t>(q) & 15) == 0) return *reinterpret_cast<const v8bf *>(q);
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = k0 + i < K ? q[i * sk] : static_cast<__bf16>(0.0f);
  return v;
}

// C[m][n] = act(sum_k A(m,k) B(n,k) + bias[n]), optionally gated by
// gate[m][n] > 0 (the ReLU derivative of the layer whose output it was).
// A(m,k) = A[m sam + k sak], B(n,k) = B[n sbn + k sbk]. Block of 4 waves:
// a 64 x 64 tile, each wave 32 x 32 as 2 x 2 MFMA tiles of 16 x 16 x 32.
//   operand lane l: row l % 16, k-elements 8 (l / 16) .. + 7;
//   result: C[4 (l / 16) + i][l % 16] in element i.
struct GemmArgs {
  uint32_t M, N, K;
  const __bf16 *A;
  size_t sam, sak;
  const __bf16 *B;
  size_t sbn, sbk;
  const float *bias;  // N or null
  const float *gate;  // M x ldg or null
  size_t ldg;
  int relu;
  float *C32;         // M x ldc or null
  __bf16 *C16;        // M x ldc or null
  size_t ldc;
};

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs a) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t m0 = blockIdx.y * 64 + 32 * (w >> 1), n0 = blockIdx.x * 64 + 32 * (w & 1);
  const uint32_t kk = 8 * (lane >> 4), rr = lane & 15;
  v4f acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
  for (uint32_t k0 = 0; k0 < a.K; k0 += 32) {
    v8bf fa[2], fb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) fa[i] = load8(a.A, m0 + 16 * i + rr, a.M, k0 + kk, a.K, a.sam, a.sak);
#pragma unroll
    for (int j = 0; j < 2; ++j) fb[j] = load8(a.B, n0 + 16 * j + rr, a.N, k0 + kk, a.K, a.sbn, a.sbk);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = mfma_bf16(fa[i], fb[j], acc[i][j]);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t m = m0 + 16 * i + 4 * (lane >> 4) + e, n = n0 + 16 * j + rr;
        if (m >= a.M || n >= a.N) continue;
        float v = acc[i][j][e];
        if (a.bias) v += a.bias[n];
        if (a.relu) v = v > 0.0f ? v : 0.0f;
        if (a.gate && !(a.gate[m * a.ldg + n] > 0.0f)) v = 0.0f;
        if (a.C32) a.C32[m * a.ldc + n] = v;
        if (a.C16) a.C16[m * a.ldc + n] = static_cast<__bf16>(v);
      }
}

__global__ void k_to_bf16(const float *x, __bf16 *y, uint64_t n) {
  const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i < n) y[i] = static_cast<__bf16>(x[i]);
}

// Mean squared error over n outputs: grad = 2 (p - y) / n; loss sum into *loss.
__global__ void k_mse(const float *p, const float *y, uint32_t n, float *grad, float *loss) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  float l = 0.0f;
  if (i < n) {
    const float d = p[i] - y[i];
    grad[i] = 2.0f * d / static_cast<float>(n);
    l = d * d / static_cast<float>(n);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, l);
}

// Pinball loss at quantile tau of y - p (the tau-quantile regression of y):
// rho(u) = u (tau - [u < 0]); grad = -(tau - [y < p]) / n; loss sum into *loss.
__global__ void k_pinball(const float *p, const float *y, uint32_t n, float tau, float *grad, float *loss) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  float l = 0.0f;
  if (i < n) {
    const float u = y[i] - p[i];
    const float w = u < 0.0f ? tau - 1.0f : tau;
    grad[i] = -w / static_cast<float>(n);
    l = u * w / static_cast<float>(n);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, l);
}

// Column sums of an R x C fp32 matrix (bias gradients): thread per column.
__global__ void k_colsum(const float *g, uint32_t R, uint32_t C, float *out) {
  const uint32_t c = blockIdx.x * 256 + threadIdx.x;
  if (c >= C) return;
  float s = 0.0f;
  for (uint32_t r = 0; r < R; ++r) s += g[static_cast<size_t>(r) * C + c];
  out[c] = s;
}

// Adam (fp32 master weights), and their bf16 copy for the products.
__global__ void k_adam(float *w, const float *g, float *m, float *v, uint32_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, __bf16 *w16) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (w16) w16[i] = static_cast<__bf16>(wi);
}

// Features of each query from its sorted coarse distances (B x np, ascending):
// x_0 = d_0, x_i = d_i - d_0 for 0 < i < F, standardised by (mean, inv_std).
__global__ void k_features(const float *pd, uint32_t B, uint32_t np, uint32_t F, const float *mean,
                           const float *inv_std, float *x32, __bf16 *x16) {
  const uint64_t t = blockIdx.x * 256ull + threadIdx.x;
  if (t >= static_cast<uint64_t>(B) * F) return;
  const uint32_t q = static_cast<uint32_t>(t / F), i = static_cast<uint32_t>(t % F);
  const float d0 = pd[static_cast<size_t>(q) * np];
  const float di = pd[static_cast<size_t>(q) * np + i];
  float x = i == 0 ? d0 : di - d0;
  x = (x - mean[i]) * inv_std[i];
  if (x32) x32[t] = x;
  if (x16) x16[t] = static_cast<__bf16>(x);
}

// Predicted nprobe: 2^(z * y_std + y_mean + margin) rounded up, in [nmin, nmax].
__global__ void k_nprobe_of(const float *z, uint32_t B, float y_mean, float y_std, float margin, uint32_t nmin,
                            uint32_t nmax, uint32_t *npq) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= B) return;
  const float l = z[q] * y_std + y_mean + margin;
  const float n = ceilf(exp2f(fminf(l, 30.0f)));
  npq[q] = n <= static_cast<float>(nmin) ? nmin : n >= static_cast<float>(nmax) ? nmax : static_cast<uint32_t>(n);
}

}  // namespace

namespace katoann_hip {

void launch_gemm_bf16(uint32_t M, uint32_t N, uint32_t K, const void *A, size_t sam, size_t sak, const void *B,
                      size_t sbn, size_t sbk, const float *bias, const float *gate, size_t ldg, int relu, float *C32,
                      void *C16, size_t ldc, hipStream_t s) {
  if (!M || !N) return;
  GemmArgs a{M, N, K, static_cast<const __bf16 *>(A), sam, sak, static_cast<const __bf16 *>(B), sbn, sbk, bias,
             gate, ldg, relu, C32, static_cast<__bf16 *>(C16), ldc};
  hipLaunchKernelGGL(k_gemm_bf16, dim3((N + 63) / 64, (M + 63) / 64), dim3(256), 0, s, a);
}

void launch_to_bf16(const float *x, void *y, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_to_bf16, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, s, x,
                     static_cast<__bf16 *>(y), n);
}

void launch_mse(const float *p, const float *y, uint32_t n, float *grad, float *loss, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_mse, dim3((n + 255) / 256), dim3(256), 0, s, p, y, n, grad, loss);
}

void launch_pinball(const float *p, const float *y, uint32_t n, float tau, float *grad, float *loss,
                    hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_pinball, dim3((n + 255) / 256), dim3(256), 0, s, p, y, n, tau, grad, loss);
}

void launch_colsum(const float *g, uint32_t R, uint32_t C, float *out, hipStream_t s) {
  if (!C) return;
  hipLaunchKernelGGL(k_colsum, dim3((C + 255) / 256), dim3(256), 0, s, g, R, C, out);
}

void launch_adam(float *w, const float *g, float *m, float *v, uint32_t n, float lr, float b1, float b2, float eps,
                 uint32_t step, void *w16, hipStream_t s) {
  if (!n) return;
  const float c1 = 1.0f - powf(b1, static_cast<float>(step)), c2 = 1.0f - powf(b2, static_cast<float>(step));
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, eps, c1, c2,
                     static_cast<__bf16 *>(w16));
}

void launch_features(const float *pd, uint32_t B, uint32_t np, uint32_t F, const float *mean, const float *inv_std,
                     float *x32, void *x16, hipStream_t s) {
  const uint64_t t = static_cast<uint64_t>(B) * F;
  if (!t) return;
  hipLaunchKernelGGL(k_features, dim3(static_cast<uint32_t>((t + 255) / 256)), dim3(256), 0, s, pd, B, np, F, mean,
                     inv_std, x32, static_cast<__bf16 *>(x16));
}

void launch_nprobe_of(const float *z, uint32_t B, float y_mean, float y_std, float margin, uint32_t nmin,
                      uint32_t nmax, uint32_t *npq, hipStream_t s) {
  if (!B) return;
  hipLaunchKernelGGL(k_nprobe_of, dim3((B + 255) / 256), dim3(256), 0, s, z, B, y_mean, y_std, margin, nmin, nmax,
                     npq);
}

}  // namespace katoann_hip
