// Reductions and Krylov-basis kernels: batched dot products, the classical
// Gram-Schmidt update, Ritz-vector reconstruction, density mixing.
//
// Reference: grid.py:57-73 (midpoint-rule integrals and inner products),
// lanczos.py:140-194 (alpha/beta, full reorthogonalisation, Ritz vectors),
// scf.py:61-63 and :422-423 (density and linear mixing).
//
// All sums accumulate in double with per-thread Neumaier compensation, a
// 64-lane wavefront shuffle tree, a block tree, and a deterministic
// last-block-finishes pass over the block partials (fixed order, so results
// are bitwise reproducible run to run).
#include "common.hpp"

namespace gato {
namespace {

constexpr int kThr = 256;
constexpr int kKB = 16;  // vectors per block group in the batched dot

// out[k] = scale * sum_i A[k*lda+i] * B[k*ldb+i] * C[i] for k in the block's group.
template <typename T>
__global__ __launch_bounds__(kThr) void dot_batched_kernel(long n, int nk, const T* __restrict__ A,
                                                           long lda, const T* __restrict__ B,
                                                           long ldb, const T* __restrict__ C,
                                                           const T* __restrict__ S,
                                                           const double* __restrict__ s_coef,
                                                           double scale, double* __restrict__ out,
                                                           unsigned int* counter,
                                                           double* partials) {
  __shared__ double red[16];
  __shared__ int flag;
  const int nblk = gridDim.x;
  const int g = blockIdx.y;
  const int k0 = g * kKB;
  const int kn = min(kKB, nk - k0);
  KSum acc[kKB];
#pragma unroll
  for (int k = 0; k < kKB; ++k) acc[k].init();
  const double sc = S ? s_coef[0] : 0.0;
  const long stride = (long)nblk * kThr;
  for (long i = (long)blockIdx.x * kThr + threadIdx.x; i < n; i += stride) {
    const double c = C ? (double)C[i] : 1.0;
    if (ldb == 0) {
      double bv = (double)B[i];
      if (S) bv -= sc * (double)S[i];
      const double bc = bv * c;
#pragma unroll
      for (int k = 0; k < kKB; ++k)
        if (k < kn) acc[k].add((double)A[(long)(k0 + k) * lda + i] * bc);
    } else {
#pragma unroll
      for (int k = 0; k < kKB; ++k)
        if (k < kn)
          acc[k].add((double)A[(long)(k0 + k) * lda + i] * (double)B[(long)(k0 + k) * ldb + i] * c);
    }
  }
  for (int k = 0; k < kn; ++k) {
    const double v = block_sum(acc[k].value(), red);
    if (threadIdx.x == 0) partials[(long)(k0 + k) * nblk + blockIdx.x] = v;
  }
  if (grid_finish(counter, (unsigned)(gridDim.x * gridDim.y), &flag)) {
    for (int k = threadIdx.x; k < nk; k += kThr) {
      KSum s;
      s.init();
      const volatile double* p = partials + (long)k * nblk;
      for (int b = 0; b < nblk; ++b) s.add(p[b]);
      out[k] = s.value() * scale;
    }
    if (threadIdx.x == 0) *counter = 0u;
  }
}

// w_out = w - sum_k coef[k] V[k]; beta = sqrt(scale * ||w_out||^2)
template <typename T>
__global__ __launch_bounds__(kThr) void cgs_update_kernel(long n, int nk, const T* __restrict__ V,
                                                          long ldv, const double* __restrict__ coef,
                                                          const T* __restrict__ S,
                                                          const double* __restrict__ s_coef,
                                                          const T* w, T* w_out, double scale,
                                                          double* __restrict__ beta,
                                                          unsigned int* counter, double* partials) {
  extern __shared__ double cs[];  // nk coefficients
  __shared__ double red[16];
  __shared__ int flag;
  for (int k = threadIdx.x; k < nk; k += kThr) cs[k] = coef[k];
  __syncthreads();
  KSum nrm;
  nrm.init();
  const double sc = S ? s_coef[0] : 0.0;
  const long stride = (long)gridDim.x * kThr;
  for (long i = (long)blockIdx.x * kThr + threadIdx.x; i < n; i += stride) {
    double a = (double)w[i];
    if (S) a -= sc * (double)S[i];
    int k = 0;
    for (; k + 4 <= nk; k += 4) {
      const double v0 = (double)V[(long)k * ldv + i];
      const double v1 = (double)V[(long)(k + 1) * ldv + i];
      const double v2 = (double)V[(long)(k + 2) * ldv + i];
      const double v3 = (double)V[(long)(k + 3) * ldv + i];
      a -= cs[k] * v0;
      a -= cs[k + 1] * v1;
      a -= cs[k + 2] * v2;
      a -= cs[k + 3] * v3;
    }
    for (; k < nk; ++k) a -= cs[k] * (double)V[(long)k * ldv + i];
    const T r = (T)a;
    w_out[i] = r;
    nrm.add((double)r * (double)r);
  }
  const double v = block_sum(nrm.value(), red);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
  if (grid_finish(counter, gridDim.x, &flag)) {
    if (threadIdx.x == 0) {
      KSum s;
      s.init();
      const volatile double* p = partials;
      for (unsigned b = 0; b < gridDim.x; ++b) s.add(p[b]);
      beta[0] = sqrt(fmax(s.value() * scale, 0.0));
      *counter = 0u;
    }
  }
}

template <typename T>
__global__ void scale_inv_kernel(long n, const T* __restrict__ w, const double* __restrict__ beta,
                                 T* __restrict__ v) {
  const double b = beta[0];
  const double inv = b > 0.0 ? 1.0 / b : 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    v[i] = (T)((double)w[i] * inv);
}

constexpr int kRitzMax = 16;

template <typename T>
__global__ __launch_bounds__(kThr) void ritz_kernel(long n, int m, int K, const T* __restrict__ V,
                                                    long ldv, const double* __restrict__ S,
                                                    T* __restrict__ out, long ldo) {
  extern __shared__ double sS[];  // m*K
  for (int i = threadIdx.x; i < m * K; i += kThr) sS[i] = S[i];
  __syncthreads();
  const long stride = (long)gridDim.x * kThr;
  for (long i = (long)blockIdx.x * kThr + threadIdx.x; i < n; i += stride) {
    double acc[kRitzMax];
#pragma unroll
    for (int k = 0; k < kRitzMax; ++k) acc[k] = 0.0;
    for (int j = 0; j < m; ++j) {
      const double v = (double)V[(long)j * ldv + i];
#pragma unroll
      for (int k = 0; k < kRitzMax; ++k)
        if (k < K) acc[k] += sS[j * K + k] * v;
    }
#pragma unroll
    for (int k = 0; k < kRitzMax; ++k)
      if (k < K) out[(long)k * ldo + i] = (T)acc[k];
  }
}

template <typename T>
__global__ void density_mix_kernel(long n, int n_orb, const T* __restrict__ orb, long ld,
                                   double mixing, const T* __restrict__ rho_old,
                                   T* __restrict__ rho) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double s = 0.0;
    for (int k = 0; k < n_orb; ++k) {
      const double v = (double)orb[(long)k * ld + i];
      s += v * v;
    }
    double r = mixing * 2.0 * s;
    if (rho_old) r += (1.0 - mixing) * (double)rho_old[i];
    rho[i] = (T)r;
  }
}

template <typename T>
__global__ void axpby_kernel(long n, double a, const T* __restrict__ x, double b, T* __restrict__ y) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    y[i] = (T)(a * (double)x[i] + b * (double)y[i]);
}

template <typename TI, typename TO>
__global__ void convert_kernel(long n, const TI* __restrict__ x, TO* __restrict__ y) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] = (TO)x[i];
}

inline int grid_blocks(long n, int per_thread = 4) {
  long b = (n + (long)kThr * per_thread - 1) / ((long)kThr * per_thread);
  if (b < 1) b = 1;
  if (b > kReduceBlocks) b = kReduceBlocks;
  return (int)b;
}

inline int ew_blocks(long n) {
  long b = (n + kThr * 4 - 1) / (kThr * 4);
  if (b < 1) b = 1;
  if (b > 65535) b = 65535;
  return (int)b;
}

}  // namespace

int dot_batched(int dtype, long n, int nk, const void* A, long lda, const void* B, long ldb,
                const void* C, const void* S, const double* s_coef, double scale, double* out,
                void* ws, hipStream_t s) {
  GATO_REQUIRE(n >= 1 && nk >= 1, GATO_ERR_VALUE, "bad sizes n=%ld nk=%d", n, nk);
  GATO_REQUIRE(A && B && out && ws, GATO_ERR_VALUE, "null pointer");
  GATO_REQUIRE(S == nullptr || (s_coef != nullptr && ldb == 0), GATO_ERR_VALUE,
               "S needs s_coef and a broadcast B (ldb == 0)");
  int nblk = grid_blocks(n);
  while (nblk > 1 && (long)nblk * nk > kPartialsCap) nblk /= 2;
  GATO_REQUIRE((long)nblk * nk <= kPartialsCap, GATO_ERR_UNSUPPORTED, "too many vectors (%d)", nk);
  dim3 grid(nblk, (nk + kKB - 1) / kKB);
  if (dtype == GATO_F32)
    hipLaunchKernelGGL((dot_batched_kernel<float>), grid, dim3(kThr), 0, s, n, nk, (const float*)A,
                       lda, (const float*)B, ldb, (const float*)C, (const float*)S, s_coef, scale,
                       out, ws_counters(ws), ws_partials(ws));
  else
    hipLaunchKernelGGL((dot_batched_kernel<double>), grid, dim3(kThr), 0, s, n, nk,
                       (const double*)A, lda, (const double*)B, ldb, (const double*)C,
                       (const double*)S, s_coef, scale, out, ws_counters(ws), ws_partials(ws));
  GATO_LAUNCH_CHECK();
  return GATO_OK;
}

int cgs_update(int dtype, long n, int nk, const void* V, long ldv, const double* coef,
               const void* S, const double* s_coef, const void* w, void* w_out, double scale,
               double* beta, void* ws, hipStream_t s) {
  GATO_REQUIRE(n >= 1 && nk >= 0 && nk <= 8192, GATO_ERR_VALUE, "bad sizes n=%ld nk=%d", n, nk);
  GATO_REQUIRE(S == nullptr || s_coef != nullptr, GATO_ERR_VALUE, "S needs s_coef");
  GATO_REQUIRE(w && w_out && beta && ws && (nk == 0 || (V && coef)), GATO_ERR_VALUE,
               "null pointer");
  const int nblk = grid_blocks(n, 2);
  const size_t sh = (size_t)(nk > 0 ? nk : 1) * sizeof(double);
  if (dtype == GATO_F32)
    hipLaunchKernelGGL((cgs_update_kernel<float>), dim3(nblk), dim3(kThr), sh, s, n, nk,
                       (const float*)V, ldv, coef, (const float*)S, s_coef, (const float*)w,
                       (float*)w_out, scale, beta, ws_counters(ws) + 1, ws_partials(ws));
  else
    hipLaunchKernelGGL((cgs_update_kernel<double>), dim3(nblk), dim3(kThr), sh, s, n, nk,
                       (const double*)V, ldv, coef, (const double*)S, s_coef, (const double*)w,
                       (double*)w_out, scale, beta, ws_counters(ws) + 1, ws_partials(ws));
  GATO_LAUNCH_CHECK();
  return GATO_OK;
}

int scale_by_inverse(int dtype, long n, const void* w, const double* beta, void* v, hipStream_t s) {
  GATO_REQUIRE(n >= 1 && w && beta && v, GATO_ERR_VALUE, "bad arguments");
  if (dtype == GATO_F32)
    hipLaunchKernelGGL((scale_inv_kernel<float>), dim3(ew_blocks(n)), dim3(kThr), 0, s, n,
                       (const float*)w, beta, (float*)v);
  else
    hipLaunchKernelGGL((scale_inv_kernel<double>), dim3(ew_blocks(n)), dim3(kThr), 0, s, n,
                       (const double*)w, beta, (double*)v);
  GATO_LAUNCH_CHECK();
  return GATO_OK;
}

int ritz_combine(int dtype, long n, int m, int K, const void* V, long ldv, const double* S,
                 void* out, long ldo, hipStream_t s) {
  GATO_REQUIRE(n >= 1 && m >= 1 && K >= 1 && K <= kRitzMax, GATO_ERR_VALUE,
               "bad sizes n=%ld m=%d K=%d (K <= %d)", n, m, K, kRitzMax);
  GATO_REQUIRE((size_t)m * K * sizeof(double) <= 60 * 1024, GATO_ERR_UNSUPPORTED,
               "Ritz coefficient block m*K=%d too large", m * K);
  const size_t sh = (size_t)m * K * sizeof(double);
  const int nb = ew_blocks(n);
  if (dtype == GATO_F32)
    hipLaunchKernelGGL((ritz_kernel<float>), dim3(nb), dim3(kThr), sh, s, n, m, K,
                       (const float*)V, ldv, S, (float*)out, ldo);
  else
    hipLaunchKernelGGL((ritz_kernel<double>), dim3(nb), dim3(kThr), sh, s, n, m, K,
                       (const double*)V, ldv, S, (double*)out, ldo);
  GATO_LAUNCH_CHECK();
  return GATO_OK;
}

}  // namespace gato

using namespace gato;

extern "C" size_t gato_reduce_workspace_bytes(void) { return kReduceWorkspaceBytes; }

extern "C" int gato_dot_batched(int dtype, long n, int nk, const void* A, long lda, const void* B,
                                long ldb, const void* C, const void* S, const double* s_coef,
                                double scale, double* out, void* ws, void* stream) {
  clear_ok();
  return dot_batched(dtype, n, nk, A, lda, B, ldb, C, S, s_coef, scale, out, ws,
                     as_stream(stream));
}

extern "C" int gato_cgs_update(int dtype, long n, int nk, const void* V, long ldv,
                               const double* coef, const void* S, const double* s_coef,
                               const void* w, void* w_out, double scale, double* beta_out,
                               void* ws, void* stream) {
  clear_ok();
  return cgs_update(dtype, n, nk, V, ldv, coef, S, s_coef, w, w_out, scale, beta_out, ws,
                    as_stream(stream));
}

extern "C" int gato_scale_by_inverse(int dtype, long n, const void* w, const double* beta, void* v,
                                     void* stream) {
  clear_ok();
  return scale_by_inverse(dtype, n, w, beta, v, as_stream(stream));
}

extern "C" int gato_ritz_combine(int dtype, long n, int m, int K, const void* V, long ldv,
                                 const double* S, void* out, long ldo, void* stream) {
  clear_ok();
  return ritz_combine(dtype, n, m, K, V, ldv, S, out, ldo, as_stream(stream));
}

extern "C" int gato_density_mix(int dtype, long n, int n_orb, const void* orbitals, long ld,
                                double mixing, const void* rho_old, void* rho_out, void* stream) {
  clear_ok();
  GATO_REQUIRE(n >= 1 && n_orb >= 1 && orbitals && rho_out, GATO_ERR_VALUE, "bad arguments");
  GATO_REQUIRE(rho_old != nullptr || mixing == 1.0, GATO_ERR_VALUE,
               "rho_old is required when mixing != 1");
  hipStream_t s = as_stream(stream);
  if (dtype == GATO_F32)
    hipLaunchKernelGGL((density_mix_kernel<float>), dim3(ew_blocks(n)), dim3(kThr), 0, s, n, n_orb,
                       (const float*)orbitals, ld, mixing, (const float*)rho_old, (float*)rho_out);
  else
    hipLaunchKernelGGL((density_mix_kernel<double>), dim3(ew_blocks(n)), dim3(kThr), 0, s, n,
                       n_orb, (const double*)orbitals, ld, mixing, (const double*)rho_old,
                       (double*)rho_out);
  GATO_LAUNCH_CHECK();
  return GATO_OK;
}

extern "C" int gato_axpby(int dtype, long n, double a, const void* x, double b, void* y,
                          void* stream) {
  clear_ok();
  GATO_REQUIRE(n >= 1 && x && y, GATO_ERR_VALUE, "bad arguments");
  hipStream_t s = as_stream(stream);
  if (dtype == GATO_F32)
    hipLaunchKernelGGL((axpby_kernel<float>), dim3(ew_blocks(n)), dim3(kThr), 0, s, n, a,
                       (const float*)x, b, (float*)y);
  else
    hipLaunchKernelGGL((axpby_kernel<double>), dim3(ew_blocks(n)), dim3(kThr), 0, s, n, a,
                       (const double*)x, b, (double*)y);
  GATO_LAUNCH_CHECK();
  return GATO_OK;
}

extern "C" int gato_convert(int in_dtype, const void* x, int out_dtype, void* y, long n,
                            void* stream) {
  clear_ok();
  GATO_REQUIRE(n >= 1 && x && y, GATO_ERR_VALUE, "bad arguments");
  hipStream_t s = as_stream(stream);
  const dim3 g(ew_blocks(n)), b(kThr);
  if (in_dtype == GATO_F32 && out_dtype == GATO_F64)
    hipLaunchKernelGGL((convert_kernel<float, double>), g, b, 0, s, n, (const float*)x, (double*)y);
  else if (in_dtype == GATO_F64 && out_dtype == GATO_F32)
    hipLaunchKernelGGL((convert_kernel<double, float>), g, b, 0, s, n, (const double*)x, (float*)y);
  else if (in_dtype == GATO_F32 && out_dtype == GATO_F32)
    hipLaunchKernelGGL((convert_kernel<float, float>), g, b, 0, s, n, (const float*)x, (float*)y);
  else
    hipLaunchKernelGGL((convert_kernel<double, double>), g, b, 0, s, n, (const double*)x,
                       (double*)y);
  GATO_LAUNCH_CHECK();
  return GATO_OK;
}
