// MI355X kernels of the warm-start network: a dense MLP (GELU hidden
// layers) trained with BF16 operands and FP32 accumulation on MFMA
// (v_mfma_f32_16x16x32_bf16) and evaluated in FP32 (v_mfma_f32_16x16x4f32).
//
// All matrices live in device memory as row-major FP32; the BF16 GEMM rounds
// its operands to BF16 (round to nearest even) as it loads them, so the
// products are exact in FP32 and only the operand rounding and the FP32
// accumulation order differ from an FP64 evaluation.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ps_mlp.h"

typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

namespace {

constexpr int TILE = 32;  // output tile of one workgroup (2 x 2 waves of 16 x 16)

__device__ __forceinline__ short to_bf16(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (short)(u >> 16);  // inf / nan: truncate
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// GELU, tanh form: 0.5 x (1 + tanh(s (x + 0.044715 x^3))), s = sqrt(2/pi)
__device__ __forceinline__ float gelu(float x) {
  const float s = 0.7978845608028654f;
  return 0.5f * x * (1.f + tanhf(s * (x + 0.044715f * x * x * x)));
}

// C[m][n] = sum_k A(m,k) B(k,n) (+ bias[n]); A(m,k) = TA ? A[k*lda+m] :
// A[m*lda+k], B(k,n) = TB ? B[n*ldb+k] : B[k*ldb+n]; with H, also H[m][n] =
// GELU(C[m][n]).  One wave per 16 x 16 output block: lane l holds row l%16
// of A and column l%16 of B over the k-values KP*(l/16) .. +KP-1 of each
// step, and D rows 4*(l/16)+r.
template <bool BF16, bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                   const float* __restrict__ B, int ldb, const float* __restrict__ bias,
                                                   float* __restrict__ C, int ldc, float* __restrict__ H) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m0 = blockIdx.y * TILE + (w >> 1) * 16, n0 = blockIdx.x * TILE + (w & 1) * 16;
  const int r16 = lane & 15, g = lane >> 4;
  const int am = m0 + r16, bn = n0 + r16;
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  auto ld_a = [&](int k) -> float {
    if (am >= M || k >= K) return 0.f;
    return TA ? A[(size_t)k * lda + am] : A[(size_t)am * lda + k];
  };
  auto ld_b = [&](int k) -> float {
    if (bn >= N || k >= K) return 0.f;
    return TB ? B[(size_t)bn * ldb + k] : B[(size_t)k * ldb + bn];
  };
  if constexpr (BF16) {
    for (int k0 = 0; k0 < K; k0 += 32) {
      v8s a, b;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        a[e] = to_bf16(ld_a(k0 + 8 * g + e));
        b[e] = to_bf16(ld_b(k0 + 8 * g + e));
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), acc, 0,
                                                     0, 0);
    }
  } else {
    for (int k0 = 0; k0 < K; k0 += 4) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ld_a(k0 + g), ld_b(k0 + g), acc, 0, 0, 0);
  }
  const int cn = n0 + r16;
  if (cn < N) {
    const float bv = bias ? bias[cn] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int cm = m0 + 4 * g + r;
      if (cm < M) {
        const float zv = acc[r] + bv;
        C[(size_t)cm * ldc + cn] = zv;
        if (H) H[(size_t)cm * ldc + cn] = gelu(zv);
      }
    }
  }
}

__device__ __forceinline__ float gelu_grad(float x) {
  const float s = 0.7978845608028654f;
  const float u = s * (x + 0.044715f * x * x * x);
  const float t = tanhf(u);
  return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * s * (1.f + 3.f * 0.044715f * x * x);
}

// dz may be dh (the backward pass updates the gradient in place)
__global__ void gelu_bwd_kernel(int n, const float* dh, const float* __restrict__ z, float* dz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dz[i] = dh[i] * gelu_grad(z[i]);
}

// dY = 2 (Y - T) / (rows * n_real) on the first n_real columns (0 on the
// padding); per-row squared errors into sq[rows].
__global__ void mse_kernel(int rows, int n, int n_real, const float* __restrict__ y, const float* __restrict__ t,
                           float* __restrict__ dy, float* __restrict__ sq) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  float s = 0.f;
  const float sc = 2.f / ((float)rows * (float)n_real);
  for (int j = 0; j < n; ++j) {
    const size_t e = (size_t)r * n + j;
    float d = 0.f;
    if (j < n_real) {
      d = y[e] - t[e];
      s += d * d;
    }
    dy[e] = sc * d;
  }
  sq[r] = s;
}

__global__ void colsum_kernel(int rows, int n, const float* __restrict__ x, float* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float s = 0.f;
  for (int r = 0; r < rows; ++r) s += x[(size_t)r * n + j];
  out[j] = s;
}

__global__ void adam_kernel(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// dst[r][:] = src[idx[r]][:]
__global__ void gather_rows_kernel(int rows, int n, const int* __restrict__ idx, const float* __restrict__ src,
                                   float* __restrict__ dst) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)rows * n) return;
  const int r = (int)(i / n), j = (int)(i - (size_t)r * n);
  dst[i] = src[(size_t)idx[r] * n + j];
}

inline dim3 grid1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

hipError_t psm_gemm(bool bf16, bool ta, bool tb, int M, int N, int K, const float* A, int lda, const float* B, int ldb,
                    const float* bias, float* C, int ldc, hipStream_t s, float* H) {
  const dim3 grid((N + TILE - 1) / TILE, (M + TILE - 1) / TILE);
#define PS_G(BF, TA_, TB_)                                                                                     \
  if (bf16 == BF && ta == TA_ && tb == TB_) {                                                                 \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(gemm_kernel<BF, TA_, TB_>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, \
                       bias, C, ldc, H);                                                                       \
    return hipGetLastError();                                                                                  \
  }
  PS_G(true, false, false)
  PS_G(true, true, false)
  PS_G(true, false, true)
  PS_G(false, false, false)
  PS_G(false, true, false)
  PS_G(false, false, true)
#undef PS_G
  return hipErrorInvalidValue;
}

hipError_t psm_gelu_bwd(int n, const float* dh, const float* z, float* dz, hipStream_t s) {
  hipLaunchKernelGGL(gelu_bwd_kernel, grid1(n), dim3(256), 0, s, n, dh, z, dz);
  return hipGetLastError();
}

hipError_t psm_mse(int rows, int n, int n_real, const float* y, const float* t, float* dy, float* sq, hipStream_t s) {
  hipLaunchKernelGGL(mse_kernel, grid1(rows), dim3(256), 0, s, rows, n, n_real, y, t, dy, sq);
  return hipGetLastError();
}

hipError_t psm_colsum(int rows, int n, const float* x, float* out, hipStream_t s) {
  hipLaunchKernelGGL(colsum_kernel, grid1(n), dim3(256), 0, s, rows, n, x, out);
  return hipGetLastError();
}

hipError_t psm_adam(int n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                    float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(adam_kernel, grid1(n), dim3(256), 0, s, n, w, g, m, v, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t psm_gather_rows(int rows, int n, const int* idx, const float* src, float* dst, hipStream_t s) {
  hipLaunchKernelGGL(gather_rows_kernel, grid1((size_t)rows * n), dim3(256), 0, s, rows, n, idx, src, dst);
  return hipGetLastError();
}
