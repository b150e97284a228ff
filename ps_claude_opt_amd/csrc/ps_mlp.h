// Launch wrappers of the warm-start network kernels (ps_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

// C = op(A) op(B) (+ bias): A(m,k) = ta ? A[k*lda+m] : A[m*lda+k],
// B(k,n) = tb ? B[n*ldb+k] : B[k*ldb+n]; bf16 rounds the operands to BF16
// (FP32 accumulation), otherwise FP32 throughout.  ta and tb are not both set.
// H (or NULL): also H = GELU(C), same layout, from the stored values of C (a
// hidden layer's activation in the same launch).
hipError_t psm_gemm(bool bf16, bool ta, bool tb, int M, int N, int K, const float* A, int lda, const float* B, int ldb,
                    const float* bias, float* C, int ldc, hipStream_t s, float* H = nullptr);
hipError_t psm_gelu_bwd(int n, const float* dh, const float* z, float* dz, hipStream_t s);
hipError_t psm_mse(int rows, int n, int n_real, const float* y, const float* t, float* dy, float* sq, hipStream_t s);
hipError_t psm_colsum(int rows, int n, const float* x, float* out, hipStream_t s);
hipError_t psm_adam(int n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                    float c1, float c2, hipStream_t s);
hipError_t psm_gather_rows(int rows, int n, const int* idx, const float* src, float* dst, hipStream_t s);
