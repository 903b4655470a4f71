// Hockney free-space Poisson solver: batched, pruned, LDS-staged Stockham FFTs.
//
// Reference: solvers/poisson.py:29-92 (softened kernel on the (2N)^3
// wrap-around grid, zero-padded density, rfftn * rfftn(G) -> irfftn, crop,
// times dV) and its callers exchange_apply (scf.py:66-85) and the energy
// decomposition (scf.py:223-229).
//
// Design (gfx950, HBM-bound):
//  * G is real and even, so its transform Gh is real and even.  Two real
//    densities rho_a, rho_b are therefore solved by ONE complex transform of
//    z = rho_a + i rho_b: ifft(Gh * fft(z)) = J[rho_a] + i J[rho_b].  A "pair"
//    below is one such complex transform.
//  * Gh is built once per (N, h, eps) in double on the device, folded to its
//    (M/2+1)^3 octant and pre-scaled by dV / M^3 (the inverse-FFT norm).
//  * The 3D transform is five pencil passes that skip the zero padding on the
//    way in and the cropped half on the way out:
//      z-fwd  (fields -> A[p][x][y][kz], x,y < N)      pair product fused in
//      y-fwd  (A -> B[p][x][ky][kz], x < N)
//      x      (B in place: fft, * Gh, ifft, keep x < N)
//      y-inv  (B -> A, keep y < N)
//      z-inv  (A -> epilogue, keep z < N)             K-accumulate / J store /
//                                                     energy reduction fused in
//  * When M = A*B has a four-step shape (M = 64, 128, 160, 192, 256, 320)
//    every pass keeps its pencils in registers: an A-point DFT, a twiddle,
//    one exchange through LDS and a B-point DFT per pencil (see the
//    four-step section below).  Other M take the generic passes: P pencils
//    per 256-thread block loaded into LDS interleaved ([element][pencil],
//    padded), mixed-radix (8,4,2,3,5,7) Stockham stages, written back.
//    Strided-axis passes take P pencils adjacent in z so every global access
//    is a run of P contiguous complex values.
#include <atomic>
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "common.hpp"

namespace gato {

namespace {

constexpr int kThreads = 256;
constexpr int kBatch = 8;  // global loads issued together per thread
constexpr int kMaxRad = 16;
constexpr int kMaxPairs = 8;

struct Radix {
  int n;
  int r[kMaxRad];
};

// ---------------------------------------------------------------------------
// In-register butterflies. dir = -1 forward (exp(-2 pi i/n)), +1 inverse.
// ---------------------------------------------------------------------------
template <typename C> __device__ __forceinline__ void bfly2(C* v) {
  C a = v[0];
  v[0] = cadd(a, v[1]);
  v[1] = csub(a, v[1]);
}

template <typename C> __device__ __forceinline__ void bfly4(C& a0, C& a1, C& a2, C& a3, int dir) {
  C t0 = cadd(a0, a2), t1 = csub(a0, a2);
  C t2 = cadd(a1, a3), t3 = cmul_i(csub(a1, a3), dir);
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <typename C> __device__ __forceinline__ void bfly8(C* v, int dir) {
  using R = decltype(C{}.x);
  const R s = (R)0.70710678118654752440;
  C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  bfly4(e0, e1, e2, e3, dir);
  bfly4(o0, o1, o2, o3, dir);
  o1 = cmul(o1, cmk<C>(s, dir * s));
  o2 = cmul_i(o2, dir);
  o3 = cmul(o3, cmk<C>(-s, dir * s));
  v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

template <typename C> __device__ __forceinline__ void bfly3(C* v, int dir) {
  using R = decltype(C{}.x);
  const R h3 = (R)0.86602540378443864676;
  C s = cadd(v[1], v[2]);
  C d = cscale(cmul_i(csub(v[1], v[2]), dir), h3);
  C m = csub(v[0], cscale(s, (R)0.5));
  v[0] = cadd(v[0], s);
  v[1] = cadd(m, d);
  v[2] = csub(m, d);
}

template <typename C> __device__ __forceinline__ void bfly5(C* v, int dir) {
  using R = decltype(C{}.x);
  const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
  const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
  C b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
  C d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
  C t1 = cadd(v[0], cadd(cscale(b1, c1), cscale(b2, c2)));
  C t2 = cadd(v[0], cadd(cscale(b1, c2), cscale(b2, c1)));
  C u1 = cmul_i(cadd(cscale(d1, s1), cscale(d2, s2)), dir);
  C u2 = cmul_i(csub(cscale(d1, s2), cscale(d2, s1)), dir);
  v[0] = cadd(v[0], cadd(b1, b2));
  v[1] = cadd(t1, u1);
  v[4] = csub(t1, u1);
  v[2] = cadd(t2, u2);
  v[3] = csub(t2, u2);
}

// Direct DFT for radix 7 using the plan's twiddle table (M divisible by 7).
template <typename C> __device__ __forceinline__ void bfly7(C* v, int dir, const C* __restrict__ tw, int M) {
  C y[7];
  const int st = M / 7;
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    C acc = v[0];
#pragma unroll
    for (int r = 1; r < 7; ++r) {
      C w = tw[((r * s) % 7) * st];
      if (dir > 0) w.y = -w.y;
      acc = cadd(acc, cmul(v[r], w));
    }
    y[s] = acc;
  }
#pragma unroll
  for (int s = 0; s < 7; ++s) v[s] = y[s];
}

template <typename C, int R>
__device__ __forceinline__ void stockham_pass(const C* __restrict__ src, C* __restrict__ dst, int M,
                                              int P, int ldp, int Ns, const C* __restrict__ tw,
                                              int dir) {
  const int MR = M / R;
  const int twstep = M / (Ns * R);
  const int nwork = MR * P;
  for (int w = threadIdx.x; w < nwork; w += blockDim.x) {
    const int t = w % P;
    const int j = w / P;
    const int k = j % Ns;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[(j + r * MR) * ldp + t];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        C wv = tw[r * k * twstep];
        if (dir > 0) wv.y = -wv.y;
        v[r] = cmul(v[r], wv);
      }
    }
    if (R == 2) bfly2(v);
    else if (R == 3) bfly3(v, dir);
    else if (R == 4) bfly4(v[0], v[1], v[2], v[3], dir);
    else if (R == 5) bfly5(v, dir);
    else if (R == 7) bfly7(v, dir, tw, M);
    else bfly8(v, dir);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[(base + r * Ns) * ldp + t] = v[r];
  }
}

// Runs the full transform on P interleaved pencils held in buf0; returns the
// buffer that holds the result.  Caller must __syncthreads() before.
template <typename C>
__device__ C* lds_fft(C* buf0, C* buf1, int M, int P, int ldp, const Radix& rad,
                      const C* __restrict__ tw, int dir) {
  C* src = buf0;
  C* dst = buf1;
  int Ns = 1;
  for (int s = 0; s < rad.n; ++s) {
    const int R = rad.r[s];
    switch (R) {
      case 8: stockham_pass<C, 8>(src, dst, M, P, ldp, Ns, tw, dir); break;
      case 4: stockham_pass<C, 4>(src, dst, M, P, ldp, Ns, tw, dir); break;
      case 2: stockham_pass<C, 2>(src, dst, M, P, ldp, Ns, tw, dir); break;
      case 3: stockham_pass<C, 3>(src, dst, M, P, ldp, Ns, tw, dir); break;
      case 5: stockham_pass<C, 5>(src, dst, M, P, ldp, Ns, tw, dir); break;
      default: stockham_pass<C, 7>(src, dst, M, P, ldp, Ns, tw, dir); break;
    }
    __syncthreads();
    C* t = src; src = dst; dst = t;
    Ns *= R;
  }
  return src;
}

// Compile-time instances of the same Stockham transform for the sizes the
// workloads use (M = 2N for N = 64, 80, 96, 160).  Identical arithmetic in
// the same radix order as factor_radix; with M, P and every stage's Ns known
// to the compiler, the index divisions become shifts and multiplications and
// the work loops unroll.
constexpr int ct_radix(int M, int stage) {  // radix of stage `stage`, 0 past the last
  const int order[6] = {8, 4, 2, 3, 5, 7};
  int m = M, k = 0;
  for (int q = 0; q < 6; ++q) {
    while (m % order[q] == 0) {
      if (k == stage) return order[q];
      m /= order[q];
      ++k;
    }
  }
  return 0;
}

// Pencils per block: the largest power of two <= 16 whose two LDS buffers
// (plus the exchange accumulator of N reals per pencil) fit in 80 KB.
constexpr int pencils_per_block(int M, int N, size_t csize, size_t rsize) {
  for (int P = 16; P >= 1; P /= 2) {
    const size_t bytes = 2 * (size_t)M * (P + 1) * csize + (size_t)N * P * rsize;
    if (bytes <= 80 * 1024) return P;
  }
  return 1;
}

template <typename T>
constexpr int ct_pencils(int M) {
  return pencils_per_block(M, M / 2, sizeof(cplx<T>), sizeof(T));
}

// One stage in place: every item of the block is read into registers, the
// block synchronises, then every result is written back over the same buffer.
// Reads all precede writes, so the result equals the ping-pong stage, with
// half the LDS of the two-buffer form (twice the blocks per CU).
template <typename C, int R, int M, int P, int Ns>
__device__ __forceinline__ void stockham_stage_ip(C* buf, const C* __restrict__ tw, int dir) {
  constexpr int MR = M / R, twstep = M / (Ns * R), nwork = MR * P, ldp = P + 1;
  constexpr int ITEMS = (nwork + kThreads - 1) / kThreads;
  C v[ITEMS][R];
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int w = threadIdx.x + it * kThreads;
    if (nwork % kThreads == 0 || w < nwork) {
      const int t = w % P;
      const int j = w / P;
      const int k = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = buf[(j + r * MR) * ldp + t];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          C wv = tw[r * k * twstep];
          if (dir > 0) wv.y = -wv.y;
          v[it][r] = cmul(v[it][r], wv);
        }
      }
      if constexpr (R == 2) bfly2(v[it]);
      else if constexpr (R == 3) bfly3(v[it], dir);
      else if constexpr (R == 4) bfly4(v[it][0], v[it][1], v[it][2], v[it][3], dir);
      else if constexpr (R == 5) bfly5(v[it], dir);
      else if constexpr (R == 7) bfly7(v[it], dir, tw, M);
      else bfly8(v[it], dir);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int w = threadIdx.x + it * kThreads;
    if (nwork % kThreads == 0 || w < nwork) {
      const int t = w % P;
      const int j = w / P;
      const int k = j % Ns;
      const int base = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(base + r * Ns) * ldp + t] = v[it][r];
    }
  }
  __syncthreads();
}

template <typename C, int M, int P, int S = 0, int Ns = 1>
__device__ __forceinline__ void lds_fft_ct(C* buf, const C* __restrict__ tw, int dir) {
  constexpr int R = ct_radix(M, S);
  if constexpr (R != 0) {
    stockham_stage_ip<C, R, M, P, Ns>(buf, tw, dir);
    lds_fft_ct<C, M, P, S + 1, Ns * R>(buf, tw, dir);
  }
}

// The compiled in-place transform when MC != 0 (then M = MC, P = PC and the
// result stays in buf0), else the runtime ping-pong one.
template <typename C, int MC, int PC>
__device__ __forceinline__ C* run_lds_fft(C* buf0, C* buf1, int M, int P, int ldp,
                                          const Radix& rad, const C* __restrict__ tw, int dir) {
  if constexpr (MC != 0) {
    lds_fft_ct<C, MC, PC>(buf0, tw, dir);
    return buf0;
  } else {
    return lds_fft(buf0, buf1, M, P, ldp, rad, tw, dir);
  }
}

// Dynamic LDS of the compiled kernels: one buffer of M x (P+1) complex plus
// the exchange accumulator of N x P reals.
template <typename T, int MC>
constexpr size_t ct_lds_bytes() {
  return (size_t)MC * (ct_pencils<T>(MC) + 1) * sizeof(cplx<T>) +
         (size_t)(MC / 2) * ct_pencils<T>(MC) * sizeof(T);
}

__device__ __forceinline__ int fold(int k, int M) { return k <= M - k ? k : M - k; }

// ---------------------------------------------------------------------------
// Generic strided pencil pass (complex -> complex), optionally fused with
// "* Gh then inverse" for the x pass.
//   element e of pencil (batch b, outer o, inner i) at
//       b*bs + o*os + i*is + e*es
//   a block takes P consecutive inner indices.
// ---------------------------------------------------------------------------
struct PassArgs {
  const void* src;
  void* dst;
  long bs_in, bs_out, os_in, os_out, es_in, es_out, is_in, is_out;
  int n_in, n_out, n_inner, M, dir, P, ldp;
  int nbatch;  // four-step x pass: batches looped inside each block
  const void* green;  // folded real table (or null): index (fold(e), fold(outer), fold(inner))
};

template <typename T, int MC = 0>
__global__ __launch_bounds__(kThreads) void fft_pass_kernel(PassArgs a, Radix rad,
                                                            const cplx<T>* __restrict__ tw) {
  using C = cplx<T>;
  constexpr int PC = MC ? ct_pencils<T>(MC) : 0;
  extern __shared__ char smem_raw[];
  const int P = PC ? PC : a.P, ldp = PC ? PC + 1 : a.ldp, M = MC ? MC : a.M;
  C* buf0 = reinterpret_cast<C*>(smem_raw);
  C* buf1 = MC ? buf0 : buf0 + (size_t)M * ldp;  // the compiled transform runs in place
  const int i0 = blockIdx.x * P;
  const long o = blockIdx.y;
  const long b = blockIdx.z;
  const C* src = reinterpret_cast<const C*>(a.src) + b * a.bs_in + o * a.os_in;
  C* dst = reinterpret_cast<C*>(a.dst) + b * a.bs_out + o * a.os_out;
  const int np = min(P, a.n_inner - i0);

  // load: each thread issues a batch of kBatch loads before writing any of
  // them to LDS, so the batch is in flight together
  const int nl = a.n_in * P;
  const bool contig = a.es_in == 1;
  for (int w0 = 0; w0 < nl; w0 += kThreads * kBatch) {
    C tmp[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int w = w0 + threadIdx.x + u * kThreads;
      if (w < nl) {
        int e, t;
        if (contig) { e = w % a.n_in; t = w / a.n_in; } else { t = w % P; e = w / P; }
        tmp[u] = (t < np) ? src[(long)(i0 + t) * a.is_in + (long)e * a.es_in] : cmk<C>(0, 0);
      }
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int w = w0 + threadIdx.x + u * kThreads;
      if (w < nl) {
        int e, t;
        if (contig) { e = w % a.n_in; t = w / a.n_in; } else { t = w % P; e = w / P; }
        buf0[e * ldp + t] = tmp[u];
      }
    }
  }
  for (int w = threadIdx.x; w < (M - a.n_in) * P; w += blockDim.x) {
    const int t = w % P, e = a.n_in + w / P;
    buf0[e * ldp + t] = cmk<C>(0, 0);
  }
  __syncthreads();
  C* res = run_lds_fft<C, MC, PC>(buf0, buf1, M, P, ldp, rad, tw, a.dir);
  if (a.green != nullptr) {
    const T* g = reinterpret_cast<const T*>(a.green);
    const int H = M / 2 + 1;
    const int fo = fold((int)o, M);
    for (int w = threadIdx.x; w < M * P; w += blockDim.x) {
      const int t = w % P, e = w / P;
      const int fi = fold(min(i0 + t, M - 1), M);
      const T gv = g[((long)fold(e, M) * H + fo) * H + fi];
      res[e * ldp + t] = cscale(res[e * ldp + t], gv);
    }
    __syncthreads();
    C* other = (res == buf0) ? buf1 : buf0;
    res = run_lds_fft<C, MC, PC>(res, other, M, P, ldp, rad, tw, +1);
  }
  // store
  const int ns = a.n_out * P;
  if (a.es_out == 1) {
    for (int w = threadIdx.x; w < ns; w += blockDim.x) {
      const int e = w % a.n_out, t = w / a.n_out;
      if (t < np) dst[(long)(i0 + t) * a.is_out + e] = res[e * ldp + t];
    }
  } else {
    for (int w = threadIdx.x; w < ns; w += blockDim.x) {
      const int t = w % P, e = w / P;
      if (t < np) dst[(long)(i0 + t) * a.is_out + (long)e * a.es_out] = res[e * ldp + t];
    }
  }
}

// ---------------------------------------------------------------------------
// z-forward with the source product fused in:
//   z(q, e) = f0[q*N+e] * g0[q*N+e] + i f1[q*N+e] * g1[q*N+e]  (e < N)
// (g null -> 1, f1 null -> imaginary part 0).  Output A[p][q][kz].
// ---------------------------------------------------------------------------
template <typename T> struct PairSrc {
  const T* f0[kMaxPairs];
  const T* g0[kMaxPairs];
  const T* f1[kMaxPairs];
  const T* g1[kMaxPairs];
};

template <typename T, int MC = 0>
__global__ __launch_bounds__(kThreads) void zfwd_kernel(PairSrc<T> s, cplx<T>* __restrict__ A,
                                                        int N_, int M_, int Mz, int P_, int ldp_,
                                                        Radix rad,
                                                        const cplx<T>* __restrict__ tw) {
  using C = cplx<T>;
  constexpr int PC = MC ? ct_pencils<T>(MC) : 0;
  const int N = MC ? MC / 2 : N_, M = MC ? MC : M_, P = PC ? PC : P_, ldp = PC ? PC + 1 : ldp_;
  extern __shared__ char smem_raw[];
  C* buf0 = reinterpret_cast<C*>(smem_raw);
  C* buf1 = MC ? buf0 : buf0 + (size_t)M * ldp;  // the compiled transform runs in place
  const int p = blockIdx.y;
  const long NN = (long)N * N;
  const long q0 = (long)blockIdx.x * P;
  const int np = (int)min((long)P, NN - q0);
  const T* f0 = s.f0[p];
  const T* g0 = s.g0[p];
  const T* f1 = s.f1[p];
  const T* g1 = s.g1[p];
  const int nl = N * P;
  for (int w0 = 0; w0 < nl; w0 += kThreads * kBatch) {
    C tmp[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int w = w0 + threadIdx.x + u * kThreads;
      C v = cmk<C>(0, 0);
      if (w < nl) {
        const int e = w % N, t = w / N;
        if (t < np) {
          const long idx = (q0 + t) * N + e;
          T re = f0[idx];
          if (g0) re *= g0[idx];
          T im = T(0);
          if (f1) {
            im = f1[idx];
            if (g1) im *= g1[idx];
          }
          v = cmk<C>(re, im);
        }
      }
      tmp[u] = v;
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int w = w0 + threadIdx.x + u * kThreads;
      if (w < nl) {
        const int e = w % N, t = w / N;
        buf0[e * ldp + t] = tmp[u];
      }
    }
  }
  for (int w = threadIdx.x; w < (M - N) * P; w += blockDim.x) {
    const int t = w % P, e = N + w / P;
    buf0[e * ldp + t] = cmk<C>(0, 0);
  }
  __syncthreads();
  C* res = run_lds_fft<C, MC, PC>(buf0, buf1, M, P, ldp, rad, tw, -1);
  C* out = A + (long)p * NN * Mz;
  for (int w = threadIdx.x; w < M * P; w += blockDim.x) {
    const int e = w % M, t = w / M;
    if (t < np) out[(q0 + t) * Mz + e] = res[e * ldp + t];
  }
}

// ---------------------------------------------------------------------------
// z-inverse with fused epilogues, looping over the pairs of the launch.
// ---------------------------------------------------------------------------
enum EpiMode { EPI_STORE = 0, EPI_EXCHANGE = 1, EPI_ENERGY = 2 };

template <typename T> struct EpiArgs {
  T* out0[kMaxPairs];          // EPI_STORE: J of the real part
  T* out1[kMaxPairs];          // EPI_STORE: J of the imaginary part
  const T* h0[kMaxPairs];      // EPI_EXCHANGE: multiplier of the real part
  const T* h1[kMaxPairs];      // EPI_EXCHANGE: multiplier of the imaginary part
  // EPI_EXCHANGE output field of each pair: consecutive pairs with the same
  // output are summed, and the sum is written (coef * sum, plus the old
  // value when bit p of accmask is set) after the last pair p of the run
  T* xout[kMaxPairs];
  double coef;                 // EPI_EXCHANGE scale
  unsigned accmask;            // EPI_EXCHANGE: run ending at pair p accumulates into its output
  double* partials;            // EPI_ENERGY block partials [2*npairs][nblk]
};

template <typename T>
__device__ __forceinline__ bool run_ends(const EpiArgs<T>& ea, int p, int npairs) {
  return p + 1 == npairs || ea.xout[p + 1] != ea.xout[p];
}

template <typename T, int MODE, int MC = 0>
__global__ __launch_bounds__(kThreads) void zinv_kernel(PairSrc<T> s, EpiArgs<T> ea,
                                                        const cplx<T>* __restrict__ A, int npairs,
                                                        int N_, int M_, int Mz, int P_, int ldp_,
                                                        Radix rad,
                                                        const cplx<T>* __restrict__ tw) {
  using C = cplx<T>;
  constexpr int PC = MC ? ct_pencils<T>(MC) : 0;
  const int N = MC ? MC / 2 : N_, M = MC ? MC : M_, P = PC ? PC : P_, ldp = PC ? PC + 1 : ldp_;
  extern __shared__ char smem_raw[];
  C* buf0 = reinterpret_cast<C*>(smem_raw);
  C* buf1 = MC ? buf0 : buf0 + (size_t)M * ldp;  // the compiled transform runs in place
  T* acc = reinterpret_cast<T*>(buf1 + (size_t)M * ldp);  // N*P (EPI_EXCHANGE)
  __shared__ double red[16];
  const long NN = (long)N * N;
  const long q0 = (long)blockIdx.x * P;
  const int np = (int)min((long)P, NN - q0);
  if (MODE == EPI_EXCHANGE) {
    for (int w = threadIdx.x; w < N * P; w += blockDim.x) acc[w] = T(0);
  }
  for (int p = 0; p < npairs; ++p) {
    const C* in = A + (long)p * NN * Mz;
    const int nl = M * P;
    for (int w0 = 0; w0 < nl; w0 += kThreads * kBatch) {
      C tmp[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int w = w0 + threadIdx.x + u * kThreads;
        if (w < nl) {
          const int e = w % M, t = w / M;
          tmp[u] = (t < np) ? in[(q0 + t) * Mz + e] : cmk<C>(0, 0);
        }
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int w = w0 + threadIdx.x + u * kThreads;
        if (w < nl) {
          const int e = w % M, t = w / M;
          buf0[e * ldp + t] = tmp[u];
        }
      }
    }
    __syncthreads();
    C* res = run_lds_fft<C, MC, PC>(buf0, buf1, M, P, ldp, rad, tw, +1);
    if (MODE == EPI_STORE) {
      T* o0 = ea.out0[p];
      T* o1 = ea.out1[p];
      for (int w = threadIdx.x; w < N * P; w += blockDim.x) {
        const int e = w % N, t = w / N;
        if (t < np) {
          const C v = res[e * ldp + t];
          const long idx = (q0 + t) * N + e;
          if (o0) o0[idx] = v.x;
          if (o1) o1[idx] = v.y;
        }
      }
    } else if (MODE == EPI_EXCHANGE) {
      const T* m0 = ea.h0[p];
      const T* m1 = ea.h1[p];
      const bool flush = run_ends(ea, p, npairs);
      const bool accum = (ea.accmask >> p) & 1u;
      T* xo = ea.xout[p];
      for (int w = threadIdx.x; w < N * P; w += blockDim.x) {
        const int e = w % N, t = w / N;
        if (t < np) {
          const C v = res[e * ldp + t];
          const long idx = (q0 + t) * N + e;
          T a = m0[idx] * v.x;
          if (m1) a += m1[idx] * v.y;
          acc[w] += a;
          if (flush) {  // the last pair of this output: write it and restart the sum
            T r = (T)ea.coef * acc[w];
            if (accum) r += xo[idx];
            xo[idx] = r;
            acc[w] = T(0);
          }
        }
      }
    } else {  // EPI_ENERGY: sum u0*Re and u1*Im with u = f*g
      const T* f0 = s.f0[p];
      const T* g0 = s.g0[p];
      const T* f1 = s.f1[p];
      const T* g1 = s.g1[p];
      KSum k0, k1;
      k0.init();
      k1.init();
      for (int w = threadIdx.x; w < N * P; w += blockDim.x) {
        const int e = w % N, t = w / N;
        if (t < np) {
          const C v = res[e * ldp + t];
          const long idx = (q0 + t) * N + e;
          double u0 = (double)f0[idx];
          if (g0) u0 *= (double)g0[idx];
          k0.add(u0 * (double)v.x);
          if (f1) {
            double u1 = (double)f1[idx];
            if (g1) u1 *= (double)g1[idx];
            k1.add(u1 * (double)v.y);
          }
        }
      }
      const double b0 = block_sum(k0.value(), red);
      const double b1 = block_sum(k1.value(), red);
      if (threadIdx.x == 0) {
        ea.partials[(long)(2 * p) * gridDim.x + blockIdx.x] = b0;
        ea.partials[(long)(2 * p + 1) * gridDim.x + blockIdx.x] = b1;
      }
    }
    __syncthreads();  // buffers are reused by the next pair
  }
}

// ---------------------------------------------------------------------------
// Four-step fast path for M = A*B: (A, B) = (8, 8), (16, 8), (16, 10),
// (16, 12), (16, 16), (16, 20), i.e. M = 64, 128, 160, 192, 256, 320.
//
// A block holds P pencils with T = max(A, B) threads each; thread (t = tid %
// P, s = tid / P) works on pencil t, in one of two register layouts:
//   B-layout: thread s < B holds the A elements s + B*j (j < A);
//   A-layout: thread s < A holds the B elements s + A*j (j < B).
// The forward transform takes a pencil in B-layout to its spectrum in
// A-layout.  With n = B*n1 + n2 and k = k1 + A*k2 (thread s = n2, register
// j = n1): an A-point DFT over j, the twiddle W_M^(n2*k1), one exchange
// through LDS (thread s becomes k1) and a B-point DFT over n2.  The inverse
// mirrors it (n = n_b + B*n_a, k = k_a + A*k_b): a B-point DFT over k_b in
// A-layout, the twiddle W_M^(-n_b*k_a), the exchange and an A-point DFT,
// ending in B-layout.  Forward transform, pointwise multiply and inverse
// therefore chain with no reordering.  The zero padding of a forward pass
// (n >= M/2, i.e. j >= A/2 in B-layout) is pruned from its first DFT, and
// the cropped half of an inverse pass is again j >= A/2.  For A = B the two
// layouts coincide.
//
// Each pencil owns an LDS row of pitch LD.  The forward exchange slot
// (k1, n2) sits at k1*(B+1) + n2, the inverse's (n_b, k_a) at n_b*(A+1) + k_a;
// A+1 and B+1 are odd.  8-byte LDS accesses are served a half-wave (32 lanes,
// 256 B = all 64 banks) at a time, so its 32 lanes must hit distinct bank
// pairs.  P = 16: a half-wave is 16 pencils x 2 consecutive s, at slots
// t*LD + s*(odd) + const, distinct mod 32 for LD = 2 (mod 32).  P = 32: 32
// pencils at one s, distinct for odd LD.
// ---------------------------------------------------------------------------
constexpr int fs_max(int a, int b) { return a > b ? a : b; }

constexpr int fs_pitch(int need, int residue) {
  int x = need;
  while (x % 32 != residue) ++x;
  return x;
}

template <int A_, int B_, int PP> struct FS {
  static constexpr int A = A_, B = B_, M = A * B, N = M / 2, P = PP;
  static constexpr int T = fs_max(A, B);  // threads per pencil
  static constexpr int THREADS = T * PP;
  static constexpr int LD =
      fs_pitch(fs_max(fs_max(A * (B + 1), B * (A + 1)), M), PP == 32 ? 1 : 2);
  static_assert(N % T == 0, "every thread takes the same share of a pencil row");
  static_assert(THREADS <= 1024, "block too large");
};

// cos(2 pi m / 16)
__host__ __device__ constexpr double cos16(int m) {
  constexpr double a = 0.92387953251128675613, b = 0.70710678118654752440,
                   c = 0.38268343236508977173;
  m &= 15;
  return m == 0 ? 1.0 : m == 1 ? a : m == 2 ? b : m == 3 ? c : m == 4 ? 0.0 : m == 5 ? -c
       : m == 6 ? -b : m == 7 ? -a : m == 8 ? -1.0 : m == 9 ? -a : m == 10 ? -b : m == 11 ? -c
       : m == 12 ? 0.0 : m == 13 ? c : m == 14 ? b : a;
}

// a * W_R^m with W_R = exp(DIR * 2 pi i / R); m is a constant after unrolling.
template <typename C, int R, int DIR>
__device__ __forceinline__ C twR(C a, int m) {
  using Re = decltype(C{}.x);
  const int k = (m * (16 / R)) & 15;
  if (k == 0) return a;
  if (k == 4) return cmul_i(a, DIR);
  if (k == 8) return cmk<C>(-a.x, -a.y);
  if (k == 12) return cmul_i(a, -DIR);
  return cmul(a, cmk<C>((Re)cos16(k), (Re)(DIR * cos16(k + 12))));
}

// In-register 16-point DFT, natural order in and out (n = a + 4b, k = c + 4d).
// HALF: inputs 8..15 are zero and are not read.
template <typename C, int DIR, bool HALF>
__device__ __forceinline__ void dft16(C* v) {
  C u[16];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    C y0, y1, y2, y3;
    if (HALF) {
      const C x0 = v[a], x1 = v[a + 4];
      const C ix1 = cmul_i(x1, DIR);
      y0 = cadd(x0, x1);
      y1 = cadd(x0, ix1);
      y2 = csub(x0, x1);
      y3 = csub(x0, ix1);
    } else {
      y0 = v[a]; y1 = v[a + 4]; y2 = v[a + 8]; y3 = v[a + 12];
      bfly4(y0, y1, y2, y3, DIR);
    }
    u[a * 4 + 0] = y0;
    u[a * 4 + 1] = twR<C, 16, DIR>(y1, a);
    u[a * 4 + 2] = twR<C, 16, DIR>(y2, 2 * a);
    u[a * 4 + 3] = twR<C, 16, DIR>(y3, 3 * a);
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    C z0 = u[c], z1 = u[4 + c], z2 = u[8 + c], z3 = u[12 + c];
    bfly4(z0, z1, z2, z3, DIR);
    v[c] = z0; v[c + 4] = z1; v[c + 8] = z2; v[c + 12] = z3;
  }
}

// In-register 8-point DFT (n = a + 4b, k = c + 2d).  HALF: inputs 4..7 zero.
template <typename C, int DIR, bool HALF>
__device__ __forceinline__ void dft8(C* v) {
  C u[8];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    C y0, y1;
    if (HALF) {
      y0 = v[a];
      y1 = v[a];
    } else {
      y0 = cadd(v[a], v[a + 4]);
      y1 = csub(v[a], v[a + 4]);
    }
    u[a * 2 + 0] = y0;
    u[a * 2 + 1] = twR<C, 8, DIR>(y1, a);
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    C z0 = u[c], z1 = u[2 + c], z2 = u[4 + c], z3 = u[6 + c];
    bfly4(z0, z1, z2, z3, DIR);
    v[c] = z0; v[c + 2] = z1; v[c + 4] = z2; v[c + 6] = z3;
  }
}

template <typename C, int P, int DIR>
__device__ __forceinline__ void dft_small(C* t) {
  if constexpr (P == 2) bfly2(t);
  else if constexpr (P == 3) bfly3(t, DIR);
  else if constexpr (P == 4) bfly4(t[0], t[1], t[2], t[3], DIR);
  else bfly5(t, DIR);
}

constexpr int inv_mod(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 1;
}

// Prime-factor (Good-Thomas) DFT of length L = P1*P2, gcd(P1, P2) = 1, in
// natural order in and out, with no twiddles: input n = (P2*n1 + P1*n2) mod L,
// a P1-point DFT over n1, a P2-point DFT over n2, and output k the residue
// with k = k1 (mod P1), k = k2 (mod P2), i.e. k = (k1*E1 + k2*E2) mod L with
// E1 = 1 (mod P1), 0 (mod P2) and E2 = 0 (mod P1), 1 (mod P2).
template <typename C, int P1, int P2, int DIR>
__device__ __forceinline__ void dft_pfa(C* v) {
  constexpr int L = P1 * P2;
  constexpr int E1 = P2 * inv_mod(P2 % P1, P1), E2 = P1 * inv_mod(P1 % P2, P2);
  static_assert(E1 % P1 == 1 && E1 % P2 == 0 && E2 % P1 == 0 && E2 % P2 == 1, "coprime factors");
  C u[L];
#pragma unroll
  for (int n2 = 0; n2 < P2; ++n2) {
    C t[P1];
#pragma unroll
    for (int n1 = 0; n1 < P1; ++n1) t[n1] = v[(P2 * n1 + P1 * n2) % L];
    dft_small<C, P1, DIR>(t);
#pragma unroll
    for (int k1 = 0; k1 < P1; ++k1) u[n2 * P1 + k1] = t[k1];
  }
#pragma unroll
  for (int k1 = 0; k1 < P1; ++k1) {
    C t[P2];
#pragma unroll
    for (int n2 = 0; n2 < P2; ++n2) t[n2] = u[n2 * P1 + k1];
    dft_small<C, P2, DIR>(t);
#pragma unroll
    for (int k2 = 0; k2 < P2; ++k2) v[(k1 * E1 + k2 * E2) % L] = t[k2];
  }
}

// L-point DFT of v[0..L).  HALF (L = 8, 16): the upper half is zero.
template <typename C, int L, int DIR, bool HALF>
__device__ __forceinline__ void dftL(C* v) {
  if constexpr (L == 16) {
    dft16<C, DIR, HALF>(v);
  } else if constexpr (L == 8) {
    dft8<C, DIR, HALF>(v);
  } else {
    static_assert(!HALF, "pruned DFTs are 8 or 16 points");
    if constexpr (L == 10) dft_pfa<C, 2, 5, DIR>(v);
    else if constexpr (L == 12) dft_pfa<C, 4, 3, DIR>(v);
    else {
      static_assert(L == 20, "four-step DFT lengths are 8, 10, 12, 16, 20");
      dft_pfa<C, 4, 5, DIR>(v);
    }
  }
}

// Forward length-M transform: B-layout in, A-layout out (see above).  row =
// this pencil's LDS row; twl = exp(-2 pi i k / M) in LDS.  HALF: elements
// n >= M/2 are zero.  Synchronises the block.
template <typename C, int A, int B, bool HALF>
__device__ __forceinline__ void fs_forward(C* v, int s, C* row, const C* twl) {
  const bool bl = A == B || s < B, al = A == B || s < A;
  if (bl) {
    dftL<C, A, -1, HALF>(v);
#pragma unroll
    for (int r = 1; r < A; ++r) v[r] = cmul(v[r], twl[s * r]);
  }
  __syncthreads();  // earlier readers of the LDS rows are done
  if (bl) {
#pragma unroll
    for (int r = 0; r < A; ++r) row[r * (B + 1) + s] = v[r];
  }
  __syncthreads();
  if (al) {
#pragma unroll
    for (int r = 0; r < B; ++r) v[r] = row[s * (B + 1) + r];
    dftL<C, B, -1, false>(v);
  }
}

// Inverse (unnormalised) length-M transform: A-layout in, B-layout out.
template <typename C, int A, int B>
__device__ __forceinline__ void fs_inverse(C* v, int s, C* row, const C* twl) {
  const bool bl = A == B || s < B, al = A == B || s < A;
  if (al) {
    dftL<C, B, +1, false>(v);
#pragma unroll
    for (int r = 1; r < B; ++r) {
      C w = twl[s * r];
      w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
  }
  __syncthreads();  // earlier readers of the LDS rows are done
  if (al) {
#pragma unroll
    for (int r = 0; r < B; ++r) row[r * (A + 1) + s] = v[r];
  }
  __syncthreads();
  if (bl) {
#pragma unroll
    for (int r = 0; r < A; ++r) v[r] = row[s * (A + 1) + r];
    dftL<C, A, +1, false>(v);
  }
}

template <typename T, int A, int B, int P>
constexpr size_t fs_lds_bytes() {
  return ((size_t)P * FS<A, B, P>::LD + FS<A, B, P>::M) * sizeof(cplx<T>);
}

template <typename C, int M>
__device__ __forceinline__ void load_twiddles(C* twl, const C* __restrict__ tw) {
  for (int i = threadIdx.x; i < M; i += blockDim.x) twl[i] = tw[i];
}

// Global accesses of W consecutive reals: 16 bytes where the row share of a
// thread allows it, else 8 or 4.
template <typename T, int W> struct VecOf;
template <> struct VecOf<float, 1> { using type = float; };
template <> struct VecOf<float, 2> { using type = float2; };
template <> struct VecOf<float, 4> { using type = float4; };
template <> struct VecOf<double, 1> { using type = double; };
template <> struct VecOf<double, 2> { using type = double2; };

template <typename T>
constexpr int max_vec() { return 16 / (int)sizeof(T); }

// Widest W <= 16 bytes with n reals split evenly over th threads in W-runs.
template <typename T>
constexpr int vec_width(int n, int th) {
  for (int w = max_vec<T>(); w > 1; w /= 2)
    if (n % (w * th) == 0) return w;
  return 1;
}

template <typename T, int W>
__device__ __forceinline__ void ldv(const T* __restrict__ p, T* o) {
  const typename VecOf<T, W>::type v = *reinterpret_cast<const typename VecOf<T, W>::type*>(p);
  const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
  for (int k = 0; k < W; ++k) o[k] = e[k];
}

template <typename T, int W>
__device__ __forceinline__ void stv(T* __restrict__ p, const T* a) {
  typename VecOf<T, W>::type v;
  T* e = reinterpret_cast<T*>(&v);
#pragma unroll
  for (int k = 0; k < W; ++k) e[k] = a[k];
  *reinterpret_cast<typename VecOf<T, W>::type*>(p) = v;
}

// Strided middle passes.  Pencils are adjacent along the inner index (stride
// 1); element e of pencil (batch b, outer o, inner i) is at
// b*bs + o*os + i + e*es.
enum FsMode { FS_FWD = 0, FS_X = 1, FS_INV = 2 };

// Float middle passes ask for 4 waves per SIMD (at most 128 VGPRs): the
// 16 x 20 x pass otherwise takes 130 and falls to 2 blocks of 5 waves per CU.
// (It then spills two registers; measured on a 160^3 exchange of 4 orbitals,
// 474 us against 477 us: the pass is bound by its VALU work, not occupancy.)
template <typename T>
constexpr int fs_pass_min_waves() {
  return sizeof(T) == 4 ? 4 : 1;
}

template <typename T, int A, int B, int P, int MODE>
__global__ __launch_bounds__((FS<A, B, P>::THREADS), (fs_pass_min_waves<T>())) void fs_pass_kernel(
    PassArgs a, const cplx<T>* __restrict__ tw) {
  using C = cplx<T>;
  using F = FS<A, B, P>;
  constexpr int M = F::M, LD = F::LD;
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  C* twl = lds + P * LD;
  const int t = threadIdx.x % P, s = threadIdx.x / P;
  const bool bl = A == B || s < B, al = A == B || s < A;
  load_twiddles<C, M>(twl, tw);
  const int i0 = blockIdx.x * P;
  C* row = lds + t * LD;
  // the x pass loops over the batch inside the block so the kernel values,
  // loaded once into registers (A-layout elements s + A*j), serve every batch
  T gv[B];
  if (MODE == FS_X && al) {
    const T* g = reinterpret_cast<const T*>(a.green);
    constexpr int H = M / 2 + 1;
    const long gb = (long)fold((int)blockIdx.y, M) * H + fold(i0 + t, M);
#pragma unroll
    for (int j = 0; j < B; ++j) gv[j] = g[(long)fold(s + A * j, M) * H * H + gb];
  }
  const int nb = MODE == FS_X ? a.nbatch : 1;
  for (int b = 0; b < nb; ++b) {
    const long bz = MODE == FS_X ? b : (long)blockIdx.z;
    const C* src = reinterpret_cast<const C*>(a.src) + bz * a.bs_in + blockIdx.y * a.os_in +
                   (i0 + t);
    C* dst = reinterpret_cast<C*>(a.dst) + bz * a.bs_out + blockIdx.y * a.os_out + (i0 + t);
    C v[F::T];
    if (MODE == FS_INV) {  // the whole spectrum, A-layout
      if (al) {
#pragma unroll
        for (int j = 0; j < B; ++j) v[j] = src[(long)(s + A * j) * a.es_in];
      }
    } else if (bl) {  // the first half of the pencil, B-layout
#pragma unroll
      for (int j = 0; j < A; ++j) v[j] = j < A / 2 ? src[(long)(s + B * j) * a.es_in] : cmk<C>(0, 0);
    }
    if (b == 0) __syncthreads();  // twiddle table
    if (MODE == FS_INV) fs_inverse<C, A, B>(v, s, row, twl);
    else fs_forward<C, A, B, true>(v, s, row, twl);
    if (MODE == FS_X) {
      if (al) {
#pragma unroll
        for (int j = 0; j < B; ++j) v[j] = cscale(v[j], gv[j]);
      }
      fs_inverse<C, A, B>(v, s, row, twl);
    }
    if (MODE == FS_FWD) {  // the whole spectrum, A-layout
      if (al) {
#pragma unroll
        for (int j = 0; j < B; ++j) dst[(long)(s + A * j) * a.es_out] = v[j];
      }
    } else if (bl) {  // the first half, B-layout
#pragma unroll
      for (int j = 0; j < A / 2; ++j) dst[(long)(s + B * j) * a.es_out] = v[j];
    }
  }
}

// z-forward with the source product fused in (see zfwd_kernel); P pencils of
// N contiguous values are staged through LDS for coalesced loads and stores.
template <typename T, int A, int B, int P>
__global__ __launch_bounds__((FS<A, B, P>::THREADS)) void fs_zfwd_kernel(
    PairSrc<T> src, cplx<T>* __restrict__ out_base, int Mz, const cplx<T>* __restrict__ tw) {
  using C = cplx<T>;
  using F = FS<A, B, P>;
  constexpr int M = F::M, N = F::N, LD = F::LD, TH = F::THREADS;
  constexpr int RW = vec_width<T>(P * N, TH), NL = P * N / TH;  // reals in
  constexpr int CV = vec_width<T>(2 * P * M, TH) / 2;  // complex per access
  constexpr int NS = P * M / TH;  // complex out
  static_assert(CV >= 1 && NL % RW == 0 && NS % CV == 0, "even shares");
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  C* twl = lds + P * LD;
  const int t = threadIdx.x % P, s = threadIdx.x / P;
  const bool bl = A == B || s < B, al = A == B || s < A;
  load_twiddles<C, M>(twl, tw);
  const int p = blockIdx.y;
  const long NN = (long)N * N;
  const long q0 = (long)blockIdx.x * P;
  const T* f0 = src.f0[p];
  const T* g0 = src.g0[p];
  const T* f1 = src.f1[p];
  const T* g1 = src.g1[p];
  const long base = q0 * N;
  T re[NL], im[NL];
#pragma unroll
  for (int u = 0; u < NL / RW; ++u) {
    const long idx = base + (long)(threadIdx.x + u * TH) * RW;
    T* r = re + u * RW;
    T* i = im + u * RW;
    ldv<T, RW>(f0 + idx, r);
    if (g0) {
      T g[RW];
      ldv<T, RW>(g0 + idx, g);
#pragma unroll
      for (int k = 0; k < RW; ++k) r[k] *= g[k];
    }
    if (f1) {
      ldv<T, RW>(f1 + idx, i);
      if (g1) {
        T g[RW];
        ldv<T, RW>(g1 + idx, g);
#pragma unroll
        for (int k = 0; k < RW; ++k) i[k] *= g[k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < RW; ++k) i[k] = T(0);
    }
  }
#pragma unroll
  for (int u = 0; u < NL / RW; ++u) {
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int w = (threadIdx.x + u * TH) * RW + k;
      lds[(w / N) * LD + (w % N)] = cmk<C>(re[u * RW + k], im[u * RW + k]);
    }
  }
  __syncthreads();
  C* row = lds + t * LD;
  C v[F::T];
  if (bl) {
#pragma unroll
    for (int j = 0; j < A; ++j) v[j] = j < A / 2 ? row[s + B * j] : cmk<C>(0, 0);
  }
  fs_forward<C, A, B, true>(v, s, row, twl);
  __syncthreads();
  if (al) {
#pragma unroll
    for (int j = 0; j < B; ++j) row[s + A * j] = v[j];
  }
  __syncthreads();
  C* out = out_base + (long)p * NN * Mz + q0 * Mz;
#pragma unroll
  for (int u = 0; u < NS / CV; ++u) {
    const int w = (threadIdx.x + u * TH) * CV;  // CV consecutive kz of one pencil
    C c[CV];
#pragma unroll
    for (int k = 0; k < CV; ++k) c[k] = lds[(w / M) * LD + (w % M) + k];
    stv<T, 2 * CV>(reinterpret_cast<T*>(out + (long)(w / M) * Mz + (w % M)),
                   reinterpret_cast<const T*>(c));
  }
}

// z-inverse with the fused epilogues of zinv_kernel, looping over the pairs.
template <typename T, int A, int B, int MODE, int P>
__global__ __launch_bounds__((FS<A, B, P>::THREADS)) void fs_zinv_kernel(
    PairSrc<T> src, EpiArgs<T> ea, const cplx<T>* __restrict__ in_base, int npairs, int Mz,
    const cplx<T>* __restrict__ tw) {
  using C = cplx<T>;
  using F = FS<A, B, P>;
  constexpr int M = F::M, N = F::N, LD = F::LD, TH = F::THREADS;
  constexpr int CV = vec_width<T>(2 * P * M, TH) / 2;  // complex per access
  constexpr int NL = P * M / TH;  // complex in
  constexpr int RW = vec_width<T>(P * N, TH), NE = P * N / TH;  // reals out
  static_assert(CV >= 1 && NL % CV == 0 && NE % RW == 0, "even shares");
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  C* twl = lds + P * LD;
  __shared__ double red[16];
  const int t = threadIdx.x % P, s = threadIdx.x / P;
  const bool bl = A == B || s < B, al = A == B || s < A;
  load_twiddles<C, M>(twl, tw);
  const long NN = (long)N * N;
  const long q0 = (long)blockIdx.x * P;
  C* row = lds + t * LD;
  T acc[NE];
#pragma unroll
  for (int u = 0; u < NE; ++u) acc[u] = T(0);
  for (int p = 0; p < npairs; ++p) {
    const C* in = in_base + (long)p * NN * Mz + q0 * Mz;
    C tmp[NL];
#pragma unroll
    for (int u = 0; u < NL / CV; ++u) {
      const int w = (threadIdx.x + u * TH) * CV;  // CV consecutive kz of one pencil
      ldv<T, 2 * CV>(reinterpret_cast<const T*>(in + (long)(w / M) * Mz + (w % M)),
                     reinterpret_cast<T*>(tmp + u * CV));
    }
    __syncthreads();  // the previous pair's epilogue has read the LDS rows
#pragma unroll
    for (int u = 0; u < NL / CV; ++u) {
      const int w = (threadIdx.x + u * TH) * CV;
#pragma unroll
      for (int k = 0; k < CV; ++k) lds[(w / M) * LD + (w % M) + k] = tmp[u * CV + k];
    }
    __syncthreads();
    C v[F::T];
    if (al) {
#pragma unroll
      for (int j = 0; j < B; ++j) v[j] = row[s + A * j];
    }
    fs_inverse<C, A, B>(v, s, row, twl);
    __syncthreads();
    if (bl) {
#pragma unroll
      for (int j = 0; j < A / 2; ++j) row[s + B * j] = v[j];
    }
    __syncthreads();
    // epilogue: RW consecutive z of one pencil per global access
    if (MODE == EPI_STORE) {
      T* o0 = ea.out0[p];
      T* o1 = ea.out1[p];
#pragma unroll
      for (int u = 0; u < NE / RW; ++u) {
        const int w = (threadIdx.x + u * TH) * RW;
        T re[RW], im[RW];
#pragma unroll
        for (int k = 0; k < RW; ++k) {
          const C val = lds[(w / N) * LD + (w % N) + k];
          re[k] = val.x;
          im[k] = val.y;
        }
        const long idx = q0 * N + w;
        if (o0) stv<T, RW>(o0 + idx, re);
        if (o1) stv<T, RW>(o1 + idx, im);
      }
    } else if (MODE == EPI_EXCHANGE) {
      const T* m0 = ea.h0[p];
      const T* m1 = ea.h1[p];
#pragma unroll
      for (int u = 0; u < NE / RW; ++u) {
        const int w = (threadIdx.x + u * TH) * RW;
        const long idx = q0 * N + w;
        T a[RW], b[RW];
        ldv<T, RW>(m0 + idx, a);
        if (m1) ldv<T, RW>(m1 + idx, b);
#pragma unroll
        for (int k = 0; k < RW; ++k) {
          const C val = lds[(w / N) * LD + (w % N) + k];
          T x = a[k] * val.x;
          if (m1) x += b[k] * val.y;
          acc[u * RW + k] += x;
        }
      }
      if (run_ends(ea, p, npairs)) {  // the last pair of this output: write it, restart
        const bool accum = (ea.accmask >> p) & 1u;
        T* xo = ea.xout[p];
#pragma unroll
        for (int u = 0; u < NE / RW; ++u) {
          const long idx = q0 * N + (long)(threadIdx.x + u * TH) * RW;
          T r[RW], o[RW];
          if (accum) ldv<T, RW>(xo + idx, o);
#pragma unroll
          for (int k = 0; k < RW; ++k) {
            r[k] = (T)ea.coef * acc[u * RW + k];
            if (accum) r[k] += o[k];
            acc[u * RW + k] = T(0);
          }
          stv<T, RW>(xo + idx, r);
        }
      }
    } else {  // EPI_ENERGY
      const T* f0 = src.f0[p];
      const T* g0 = src.g0[p];
      const T* f1 = src.f1[p];
      const T* g1 = src.g1[p];
      KSum k0, k1;
      k0.init();
      k1.init();
#pragma unroll
      for (int u = 0; u < NE / RW; ++u) {
        const int w = (threadIdx.x + u * TH) * RW;
        const long idx = q0 * N + w;
        T a[RW], b[RW], c[RW], d[RW];
        ldv<T, RW>(f0 + idx, a);
        if (g0) ldv<T, RW>(g0 + idx, b);
        if (f1) ldv<T, RW>(f1 + idx, c);
        if (f1 && g1) ldv<T, RW>(g1 + idx, d);
#pragma unroll
        for (int k = 0; k < RW; ++k) {
          const C val = lds[(w / N) * LD + (w % N) + k];
          double u0 = (double)a[k];
          if (g0) u0 *= (double)b[k];
          k0.add(u0 * (double)val.x);
          if (f1) {
            double u1 = (double)c[k];
            if (g1) u1 *= (double)d[k];
            k1.add(u1 * (double)val.y);
          }
        }
      }
      const double b0 = block_sum(k0.value(), red);
      const double b1 = block_sum(k1.value(), red);
      if (threadIdx.x == 0) {
        ea.partials[(long)(2 * p) * gridDim.x + blockIdx.x] = b0;
        ea.partials[(long)(2 * p + 1) * gridDim.x + blockIdx.x] = b1;
      }
    }
  }
}

// Green's function fill (double): C[x][y][z] = 1/sqrt(d(x)^2+d(y)^2+d(z)^2+eps^2)
__global__ void fill_green_kernel(double2* __restrict__ C, int N, int M, double h, double eps) {
  const long total = (long)M * M * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const int z = (int)(i % M);
    const int y = (int)((i / M) % M);
    const int x = (int)(i / ((long)M * M));
    const double dx = (x < N ? x : x - M) * h;
    const double dy = (y < N ? y : y - M) * h;
    const double dz = (z < N ? z : z - M) * h;
    C[i] = make_double2(1.0 / sqrt(dx * dx + dy * dy + dz * dz + eps * eps), 0.0);
  }
}

template <typename T>
__global__ void extract_green_kernel(const double2* __restrict__ C, T* __restrict__ g, int M,
                                     double scale) {
  const int H = M / 2 + 1;
  const long total = (long)H * H * H;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const int z = (int)(i % H);
    const int y = (int)((i / H) % H);
    const int x = (int)(i / ((long)H * H));
    g[i] = (T)(C[((long)x * M + y) * M + z].x * scale);
  }
}

bool factor_radix(int M, Radix& rad) {
  rad.n = 0;
  int m = M;
  while (m % 8 == 0 && rad.n < kMaxRad) { rad.r[rad.n++] = 8; m /= 8; }
  while (m % 4 == 0 && rad.n < kMaxRad) { rad.r[rad.n++] = 4; m /= 4; }
  while (m % 2 == 0 && rad.n < kMaxRad) { rad.r[rad.n++] = 2; m /= 2; }
  while (m % 3 == 0 && rad.n < kMaxRad) { rad.r[rad.n++] = 3; m /= 3; }
  while (m % 5 == 0 && rad.n < kMaxRad) { rad.r[rad.n++] = 5; m /= 5; }
  while (m % 7 == 0 && rad.n < kMaxRad) { rad.r[rad.n++] = 7; m /= 7; }
  return m == 1;
}

int choose_P(int M, int N, size_t csize, size_t rsize) {
  return pencils_per_block(M, N, csize, rsize);
}

// Calls f(std::integral_constant<int, MC>) with the compiled size MC that
// matches (M, P), or with MC = 0 (the runtime kernels).
template <typename T, typename F>
void dispatch_ct(int M, int P, F&& f) {
  if (M == 128 && P == ct_pencils<T>(128)) return f(std::integral_constant<int, 128>{});
  if (M == 160 && P == ct_pencils<T>(160)) return f(std::integral_constant<int, 160>{});
  if (M == 192 && P == ct_pencils<T>(192)) return f(std::integral_constant<int, 192>{});
  if (M == 320 && P == ct_pencils<T>(320)) return f(std::integral_constant<int, 320>{});
  return f(std::integral_constant<int, 0>{});
}

size_t lds_bytes(int M, int N, int P, size_t csize, size_t rsize) {
  return 2 * (size_t)M * (P + 1) * csize + (size_t)N * P * rsize;
}

}  // namespace
}  // namespace gato

struct gato_poisson_plan {
  int dtype, N, M, Mz, max_pairs, P, ldp;
  int fsA, fsB;  // four-step shape (M = fsA * fsB) or 0 for the generic Stockham passes
  int fsP;       // pencils per block of the four-step middle passes (float, A = 16: 16 or 32)
  double h, eps, dV;
  gato::Radix rad;
  void* tw;      // dtype twiddles exp(-2 pi i k / M)
  void* ghat;    // dtype folded transform (M/2+1)^3, scaled by dV / M^3
  void* A;       // max_pairs * N*N*Mz complex dtype ([p][x][y][kz], row pitch Mz)
  void* B;       // max_pairs * N*M*Mz complex dtype ([p][x][ky][kz], row pitch Mz)
  size_t lds;
};

namespace gato {
namespace {

template <typename T>
int run_fft_pass(const gato_poisson_plan* pl, const PassArgs& a0, int n_outer, int nbatch,
                 hipStream_t s, const void* tw, bool probe = false) {
  PassArgs a = a0;
  a.P = pl->P;
  a.ldp = pl->ldp;
  a.M = pl->M;
  dim3 grid((a.n_inner + a.P - 1) / a.P, n_outer, nbatch);
  const bool probed = probe && probe_on(PROBE_FFT_MIDDLE, (int)sizeof(T));
  if (probed) probe_begin(s);
  dispatch_ct<T>(a.M, a.P, [&](auto mc) {
    constexpr int MC = decltype(mc)::value;
    const size_t lds = MC ? ct_lds_bytes<T, MC>() : pl->lds;
    hipLaunchKernelGGL((fft_pass_kernel<T, MC>), grid, dim3(kThreads), lds, s, a, pl->rad,
                       (const cplx<T>*)tw);
  });
  GATO_LAUNCH_CHECK();
  if (probed) {
    // algorithmic bytes: every pencil element read (n_in of M) and written (n_out of M) once,
    // plus the folded kernel transform when fused
    const double pencils = (double)n_outer * a.n_inner * nbatch;
    const int H = pl->M / 2 + 1;
    double b = pencils * (a.n_in + a.n_out) * sizeof(cplx<T>);
    if (a.green) b += (double)H * H * H * sizeof(T);
    probe_end(s, b);
  }
  return GATO_OK;
}

template <typename T, int A, int B, int MODE, int P>
int run_fs_pass(const gato_poisson_plan* pl, const PassArgs& a0, int n_outer, int nbatch,
                hipStream_t s) {
  using F = FS<A, B, P>;
  PassArgs a = a0;
  a.nbatch = nbatch;
  dim3 grid(F::M / P, n_outer, MODE == FS_X ? 1 : nbatch);
  const bool probed = probe_on(PROBE_FFT_MIDDLE, (int)sizeof(T));
  if (probed) probe_begin(s);
  const size_t lds = fs_lds_bytes<T, A, B, P>();
  hipLaunchKernelGGL((fs_pass_kernel<T, A, B, P, MODE>), grid, dim3(F::THREADS), lds, s, a,
                     (const cplx<T>*)pl->tw);
  GATO_LAUNCH_CHECK();
  if (probed) {
    const double pencils = (double)n_outer * F::M * nbatch;
    const int H = F::M / 2 + 1;
    double b = pencils * (a.n_in + a.n_out) * sizeof(cplx<T>);
    if (a.green) b += (double)H * H * H * sizeof(T);
    probe_end(s, b);
  }
  return GATO_OK;
}

// Four-step shape of a transform length: M = A * B with an A-point first DFT
// (pruned to its nonzero half) and a B-point second one; (0, 0) when M has
// none, and the generic Stockham passes run.
void fs_shape(int M, int& A, int& B) {
  A = B = 0;
  switch (M) {
    case 64: A = 8; B = 8; break;
    case 128: A = 16; B = 8; break;
    case 160: A = 16; B = 10; break;
    case 192: A = 16; B = 12; break;
    case 256: A = 16; B = 16; break;
    case 320: A = 16; B = 20; break;
    default: break;
  }
}

// Calls f(integral_constant A, integral_constant B) for the plan's shape.
template <typename F>
int with_fs_shape(const gato_poisson_plan* pl, F&& f) {
  using std::integral_constant;
  const int A = pl->fsA, B = pl->fsB;
  if (A == 8 && B == 8) return f(integral_constant<int, 8>{}, integral_constant<int, 8>{});
  if (A == 16 && B == 8) return f(integral_constant<int, 16>{}, integral_constant<int, 8>{});
  if (A == 16 && B == 10) return f(integral_constant<int, 16>{}, integral_constant<int, 10>{});
  if (A == 16 && B == 12) return f(integral_constant<int, 16>{}, integral_constant<int, 12>{});
  if (A == 16 && B == 16) return f(integral_constant<int, 16>{}, integral_constant<int, 16>{});
  if (A == 16 && B == 20) return f(integral_constant<int, 16>{}, integral_constant<int, 20>{});
  set_error(GATO_ERR_UNSUPPORTED, "no four-step kernels for %d x %d", A, B);
  return GATO_ERR_UNSUPPORTED;
}

// Calls f for every four-step shape.
template <typename F>
void for_each_fs_shape(F&& f) {
  using std::integral_constant;
  f(integral_constant<int, 8>{}, integral_constant<int, 8>{});
  f(integral_constant<int, 16>{}, integral_constant<int, 8>{});
  f(integral_constant<int, 16>{}, integral_constant<int, 10>{});
  f(integral_constant<int, 16>{}, integral_constant<int, 12>{});
  f(integral_constant<int, 16>{}, integral_constant<int, 16>{});
  f(integral_constant<int, 16>{}, integral_constant<int, 20>{});
}

// Pencils per block of the four-step passes.  Middle passes: 32 for A = 8
// (256 threads); float A = 16 takes 16 or 32, whichever the plan measured
// faster (both give the same bits); double 16.  z-forward: 32 for A = 8 and
// for float 16 x 16 (the conflict-free odd LDS pitch, 256-byte runs), else 16.
// z-inverse: 32 for A = 8, else 16 (it loops over the pairs of a launch, and
// larger blocks would leave too few blocks in flight).
template <typename T, int A, int B>
constexpr int zfwd_pencils() {
  return A == 8 ? 32 : (A == 16 && B == 16 && sizeof(T) == 4) ? 32 : 16;
}
template <int A>
constexpr int zinv_pencils() {
  return A == 8 ? 32 : 16;
}

std::atomic<int> g_forced_fs_pencils{0};

template <typename T, int A, int B, int P>
int run_fs_mode(const gato_poisson_plan* pl, const PassArgs& a, int mode, int n_outer, int nbatch,
                hipStream_t s) {
  if (mode == FS_FWD) return run_fs_pass<T, A, B, FS_FWD, P>(pl, a, n_outer, nbatch, s);
  if (mode == FS_X) return run_fs_pass<T, A, B, FS_X, P>(pl, a, n_outer, nbatch, s);
  return run_fs_pass<T, A, B, FS_INV, P>(pl, a, n_outer, nbatch, s);
}

// One middle pass: the four-step kernel when M has a four-step shape, else
// the generic Stockham kernel.  All middle passes have n_inner = M and unit
// inner stride, which the four-step kernel assumes.
template <typename T>
int middle_pass(const gato_poisson_plan* pl, const PassArgs& a, int mode, int n_outer, int nbatch,
                hipStream_t s) {
  if (pl->fsA == 0) return run_fft_pass<T>(pl, a, n_outer, nbatch, s, pl->tw, true);
  return with_fs_shape(pl, [&](auto ca, auto cb) {
    constexpr int A = decltype(ca)::value, B = decltype(cb)::value;
    if constexpr (A == 8) {
      return run_fs_mode<T, A, B, 32>(pl, a, mode, n_outer, nbatch, s);
    } else if constexpr (sizeof(T) == 4) {
      const int forced = g_forced_fs_pencils.load();
      if ((forced ? forced : pl->fsP) == 16)
        return run_fs_mode<T, A, B, 16>(pl, a, mode, n_outer, nbatch, s);
      return run_fs_mode<T, A, B, 32>(pl, a, mode, n_outer, nbatch, s);
    } else {
      return run_fs_mode<T, A, B, 16>(pl, a, mode, n_outer, nbatch, s);
    }
  });
}

// The three middle passes (y-fwd, x with Gh, y-inv) over npairs transforms.
template <typename T>
int middle_passes(const gato_poisson_plan* pl, int npairs, hipStream_t s) {
  const int N = pl->N, M = pl->M, Mz = pl->Mz;
  // rows along kz have pitch Mz > M so that the y and x strides are not powers of two
  const long NM = (long)N * Mz, MM = (long)M * Mz;
  PassArgs a{};
  // y-forward: A[p][x][y][kz] -> B[p][x][ky][kz]
  a.src = pl->A; a.dst = pl->B;
  a.bs_in = (long)N * NM; a.bs_out = (long)N * MM;
  a.os_in = NM; a.os_out = MM;
  a.es_in = Mz; a.es_out = Mz;
  a.is_in = 1; a.is_out = 1;
  a.n_in = N; a.n_out = M; a.n_inner = M; a.dir = -1; a.green = nullptr;
  int st = middle_pass<T>(pl, a, FS_FWD, N, npairs, s);
  if (st) return st;
  // x: B[p][x][ky][kz], pencils along x; forward, * Gh, inverse; keep x < N
  a.src = pl->B; a.dst = pl->B;
  a.bs_in = (long)N * MM; a.bs_out = (long)N * MM;
  a.os_in = Mz; a.os_out = Mz;
  a.es_in = MM; a.es_out = MM;
  a.is_in = 1; a.is_out = 1;
  a.n_in = N; a.n_out = N; a.n_inner = M; a.dir = -1; a.green = pl->ghat;
  st = middle_pass<T>(pl, a, FS_X, M, npairs, s);
  if (st) return st;
  // y-inverse: B -> A, keep y < N
  a.src = pl->B; a.dst = pl->A;
  a.bs_in = (long)N * MM; a.bs_out = (long)N * NM;
  a.os_in = MM; a.os_out = NM;
  a.es_in = Mz; a.es_out = Mz;
  a.is_in = 1; a.is_out = 1;
  a.n_in = M; a.n_out = N; a.n_inner = M; a.dir = +1; a.green = nullptr;
  return middle_pass<T>(pl, a, FS_INV, N, npairs, s);
}

// Whether the four-step z kernels may run: 2N has a four-step shape and every field pointer
// 16-byte aligned (their 16-byte global accesses); otherwise the generic
// mixed-radix z kernels run (same intermediate layout).
inline bool a16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

template <typename T>
bool fast_z(const gato_poisson_plan* pl, const PairSrc<T>& src, const EpiArgs<T>& ea, int np) {
  if (pl->fsA == 0) return false;
  bool ok = true;
  for (int p = 0; p < np; ++p)
    ok = ok && a16(src.f0[p]) && a16(src.g0[p]) && a16(src.f1[p]) && a16(src.g1[p]) &&
         a16(ea.out0[p]) && a16(ea.out1[p]) && a16(ea.h0[p]) && a16(ea.h1[p]) &&
         a16(ea.xout[p]);
  return ok;
}

// Pencils per block of the z-inverse pass (blocks = N*N / this).
int z_pencils(const gato_poisson_plan* pl, bool fast) {
  if (!fast) return pl->P;
  return pl->fsA == 8 ? zinv_pencils<8>() : zinv_pencils<16>();
}

// Distinct field pointers among the sources of a launch (algorithmic reads).
template <typename T>
int distinct_fields(const PairSrc<T>& src, int npairs) {
  std::vector<const T*> seen;
  for (int p = 0; p < npairs; ++p)
    for (const T* f : {src.f0[p], src.g0[p], src.f1[p], src.g1[p]})
      if (f && std::find(seen.begin(), seen.end(), f) == seen.end()) seen.push_back(f);
  return (int)seen.size();
}

template <typename T>
int zfwd(const gato_poisson_plan* pl, const PairSrc<T>& src, int npairs, bool fast,
         hipStream_t s) {
  const long NN = (long)pl->N * pl->N;
  // every four-step block takes P whole pencils (N*N is a multiple of P for every shape)
  if (fast) {
    const int st = with_fs_shape(pl, [&](auto ca, auto cb) {
      constexpr int P = zfwd_pencils<T, decltype(ca)::value, decltype(cb)::value>();
      GATO_REQUIRE(NN % P == 0, GATO_ERR_UNSUPPORTED, "N*N=%ld not a multiple of %d", NN, P);
      return GATO_OK;
    });
    if (st) return st;
  }
  const bool probed = probe_on(PROBE_FFT_Z, (int)sizeof(T));
  if (probed) probe_begin(s);
  if (fast) {
    const int st = with_fs_shape(pl, [&](auto ca, auto cb) {
      constexpr int A = decltype(ca)::value, B = decltype(cb)::value;
      constexpr int P = zfwd_pencils<T, A, B>();
      using F = FS<A, B, P>;
      const size_t lds = fs_lds_bytes<T, A, B, P>();
      hipLaunchKernelGGL((fs_zfwd_kernel<T, A, B, P>), dim3((unsigned)(NN / P), npairs, 1),
                         dim3(F::THREADS), lds, s, src, (cplx<T>*)pl->A, pl->Mz,
                         (const cplx<T>*)pl->tw);
      return GATO_OK;
    });
    if (st) return st;
  } else {
    dim3 grid((unsigned)((NN + pl->P - 1) / pl->P), npairs, 1);
    dispatch_ct<T>(pl->M, pl->P, [&](auto mc) {
      constexpr int MC = decltype(mc)::value;
      const size_t lds = MC ? ct_lds_bytes<T, MC>() : pl->lds;
      hipLaunchKernelGGL((zfwd_kernel<T, MC>), grid, dim3(kThreads), lds, s, src,
                         (cplx<T>*)pl->A, pl->N, pl->M, pl->Mz, pl->P, pl->ldp, pl->rad,
                         (const cplx<T>*)pl->tw);
    });
  }
  GATO_LAUNCH_CHECK();
  if (probed) {
    // read each distinct source field once, write N*N*M complex per pair
    probe_end(s, (double)distinct_fields(src, npairs) * NN * pl->N * sizeof(T) +
                     (double)npairs * NN * pl->M * sizeof(cplx<T>));
  }
  return GATO_OK;
}

template <typename T, int MODE>
int zinv(const gato_poisson_plan* pl, const PairSrc<T>& src, const EpiArgs<T>& ea, int npairs,
         bool fast, hipStream_t s) {
  const long NN = (long)pl->N * pl->N;
  if (fast) {
    const int st = with_fs_shape(pl, [&](auto ca, auto cb) {
      constexpr int P = zinv_pencils<decltype(ca)::value>();
      (void)cb;
      GATO_REQUIRE(NN % P == 0, GATO_ERR_UNSUPPORTED, "N*N=%ld not a multiple of %d", NN, P);
      return GATO_OK;
    });
    if (st) return st;
  }
  const bool probed = probe_on(PROBE_FFT_Z, (int)sizeof(T));
  if (probed) probe_begin(s);
  if (fast) {
    const int st = with_fs_shape(pl, [&](auto ca, auto cb) {
      constexpr int A = decltype(ca)::value, B = decltype(cb)::value;
      constexpr int P = zinv_pencils<A>();
      using F = FS<A, B, P>;
      const size_t lds = fs_lds_bytes<T, A, B, P>();
      hipLaunchKernelGGL((fs_zinv_kernel<T, A, B, MODE, P>), dim3((unsigned)(NN / P), 1, 1),
                         dim3(F::THREADS), lds, s, src, ea, (const cplx<T>*)pl->A, npairs,
                         pl->Mz, (const cplx<T>*)pl->tw);
      return GATO_OK;
    });
    if (st) return st;
  } else {
    dim3 grid((unsigned)((NN + pl->P - 1) / pl->P), 1, 1);
    dispatch_ct<T>(pl->M, pl->P, [&](auto mc) {
      constexpr int MC = decltype(mc)::value;
      const size_t lds = MC ? ct_lds_bytes<T, MC>() : pl->lds;
      hipLaunchKernelGGL((zinv_kernel<T, MODE, MC>), grid, dim3(kThreads), lds, s, src, ea,
                         (const cplx<T>*)pl->A, npairs, pl->N, pl->M, pl->Mz, pl->P, pl->ldp,
                         pl->rad, (const cplx<T>*)pl->tw);
    });
  }
  GATO_LAUNCH_CHECK();
  if (probed) {
    // read N*N*M complex per pair, plus the epilogue's fields
    const double field = (double)NN * pl->N * sizeof(T);
    double b = (double)npairs * NN * pl->M * sizeof(cplx<T>);
    if (MODE == EPI_STORE) {
      for (int p = 0; p < npairs; ++p) b += field * ((ea.out0[p] != nullptr) + (ea.out1[p] != nullptr));
    } else if (MODE == EPI_EXCHANGE) {
      for (int p = 0; p < npairs; ++p) {
        b += field * ((ea.h0[p] != nullptr) + (ea.h1[p] != nullptr));
        if (p + 1 == npairs || ea.xout[p + 1] != ea.xout[p])
          b += field * (1 + ((ea.accmask >> p) & 1u));  // each output written once
      }
    } else {
      b += field * distinct_fields(src, npairs);
    }
    probe_end(s, b);
  }
  return GATO_OK;
}

// Device scratch of build_green, released on every exit path: the stream is
// synchronised first, so no queued kernel or copy still uses the scratch or
// the host twiddle table (declared before the guard, destroyed after it).
struct GreenScratch {
  hipStream_t s;
  double2* C = nullptr;
  double2* dtw = nullptr;
  explicit GreenScratch(hipStream_t st) : s(st) {}
  GreenScratch(const GreenScratch&) = delete;
  GreenScratch& operator=(const GreenScratch&) = delete;
  ~GreenScratch() {
    (void)hipStreamSynchronize(s);
    if (C) (void)hipFree(C);
    if (dtw) (void)hipFree(dtw);
  }
};

template <typename T>
int build_green(gato_poisson_plan* pl, hipStream_t s) {
  const int N = pl->N, M = pl->M;
  const long M3 = (long)M * M * M;
  // double twiddles
  std::vector<double2> tw64(M);
  for (int k = 0; k < M; ++k) {
    const double ang = -2.0 * M_PI * (double)k / (double)M;
    tw64[k] = make_double2(std::cos(ang), std::sin(ang));
  }
  GreenScratch g(s);
  GATO_HIP_CHECK(hipMalloc(&g.C, M3 * sizeof(double2)));
  double2* const& C = g.C;
  GATO_HIP_CHECK(hipMalloc(&g.dtw, M * sizeof(double2)));
  double2* const& dtw = g.dtw;
  GATO_HIP_CHECK(hipMemcpyAsync(dtw, tw64.data(), M * sizeof(double2), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(fill_green_kernel, dim3(2048), dim3(256), 0, s, C, N, M, pl->h, pl->eps);
  GATO_LAUNCH_CHECK();
  // unpruned 3D forward transform in double, in place
  gato_poisson_plan tmp = *pl;
  tmp.P = choose_P(M, 0, sizeof(double2), 0);
  tmp.ldp = tmp.P + 1;
  tmp.lds = lds_bytes(M, 0, tmp.P, sizeof(double2), 0);
  PassArgs a{};
  a.src = C; a.dst = C; a.dir = -1; a.green = nullptr;
  a.n_in = M; a.n_out = M;
  // z: pencils contiguous
  a.bs_in = a.bs_out = 0; a.os_in = a.os_out = 0;
  a.es_in = a.es_out = 1; a.is_in = a.is_out = M; a.n_inner = M * M;
  int st = run_fft_pass<double>(&tmp, a, 1, 1, s, dtw);
  if (st) return st;
  // y
  a.os_in = a.os_out = (long)M * M; a.es_in = a.es_out = M; a.is_in = a.is_out = 1; a.n_inner = M;
  st = run_fft_pass<double>(&tmp, a, M, 1, s, dtw);
  if (st) return st;
  // x
  a.os_in = a.os_out = M; a.es_in = a.es_out = (long)M * M; a.is_in = a.is_out = 1; a.n_inner = M;
  st = run_fft_pass<double>(&tmp, a, M, 1, s, dtw);
  if (st) return st;
  const int H = M / 2 + 1;
  GATO_HIP_CHECK(hipMalloc(&pl->ghat, (size_t)H * H * H * sizeof(T)));
  const double scale = pl->dV / ((double)M * M * M);
  hipLaunchKernelGGL((extract_green_kernel<T>), dim3(1024), dim3(256), 0, s, C, (T*)pl->ghat, M,
                     scale);
  GATO_LAUNCH_CHECK();
  GATO_HIP_CHECK(hipStreamSynchronize(s));
  return GATO_OK;  // g frees C and dtw
}

template <typename T>
int plan_create(gato_poisson_plan* pl, hipStream_t s) {
  const int M = pl->M, N = pl->N;
  std::vector<cplx<T>> tw(M);
  for (int k = 0; k < M; ++k) {
    const double ang = -2.0 * M_PI * (double)k / (double)M;
    tw[k].x = (T)std::cos(ang);
    tw[k].y = (T)std::sin(ang);
  }
  GATO_HIP_CHECK(hipMalloc(&pl->tw, M * sizeof(cplx<T>)));
  // synchronous: the host table is a local that ends with this call, on every path
  GATO_HIP_CHECK(hipMemcpyAsync(pl->tw, tw.data(), M * sizeof(cplx<T>), hipMemcpyHostToDevice, s));
  GATO_HIP_CHECK(hipStreamSynchronize(s));
  pl->P = choose_P(M, N, sizeof(cplx<T>), sizeof(T));
  pl->ldp = pl->P + 1;
  pl->lds = lds_bytes(M, N, pl->P, sizeof(cplx<T>), sizeof(T));
  pl->Mz = M + 16;
  fs_shape(M, pl->fsA, pl->fsB);
  const size_t abytes = (size_t)pl->max_pairs * N * N * pl->Mz * sizeof(cplx<T>);
  const size_t bbytes = (size_t)pl->max_pairs * N * (size_t)M * pl->Mz * sizeof(cplx<T>);
  GATO_HIP_CHECK(hipMalloc(&pl->A, abytes));
  GATO_HIP_CHECK(hipMalloc(&pl->B, bbytes));
  int st = build_green<T>(pl, s);
  if (st) return st;
  pl->fsP = (pl->fsA == 16 && sizeof(T) == 4) ? 32 : 0;
  if (pl->fsP && autotune_on()) {
    // time the middle passes of a two-pair transform (one Lanczos matvec of
    // four occupied orbitals) with 32 and with 16 pencils per block on the
    // plan's own (zeroed) scratch: one warm-up and three timed runs each
    GATO_HIP_CHECK(hipMemsetAsync(pl->A, 0, abytes, s));
    GATO_HIP_CHECK(hipMemsetAsync(pl->B, 0, bbytes, s));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    GATO_HIP_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
      (void)hipEventDestroy(e0);
      GATO_HIP_CHECK(hipGetLastError());
    }
    const int np = pl->max_pairs < 2 ? pl->max_pairs : 2;
    float best_ms = 0.f;
    int best = 32;
    for (int cand : {32, 16}) {
      pl->fsP = cand;
      st = middle_passes<T>(pl, np, s);
      for (int r = 0; !st && r < 3; ++r) {
        if (r == 0 && hipEventRecord(e0, s) != hipSuccess) st = GATO_ERR_HIP;
        if (!st) st = middle_passes<T>(pl, np, s);
      }
      float ms = 0.f;
      if (!st && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                  hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        st = GATO_ERR_HIP;
      if (st) break;
      if (cand == 32 || ms < best_ms) {
        best = cand;
        best_ms = ms;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    pl->fsP = best;
  }
  return GATO_OK;
}

// Hartree potential of n_rho densities, two per transform.
template <typename T>
int hartree_potential_t(gato_poisson_plan* pl, int n_rho, const T* rho, long ld, T* J, long ldj,
                        hipStream_t s) {
  int k = 0;
  while (k < n_rho) {
    PairSrc<T> src{};
    EpiArgs<T> ea{};
    int np = 0;
    for (; np < pl->max_pairs && k < n_rho; ++np) {
      src.f0[np] = rho + (long)k * ld;
      ea.out0[np] = J + (long)k * ldj;
      ++k;
      if (k < n_rho) {
        src.f1[np] = rho + (long)k * ld;
        ea.out1[np] = J + (long)k * ldj;
        ++k;
      }
    }
    const bool fast = fast_z(pl, src, ea, np);
    int st = zfwd<T>(pl, src, np, fast, s);
    if (!st) st = middle_passes<T>(pl, np, s);
    if (!st) st = zinv<T, EPI_STORE>(pl, src, ea, np, fast, s);
    if (st) return st;
  }
  return GATO_OK;
}

template <typename T>
int energy_pairs_t(gato_poisson_plan* pl, int nsol, const T* const* f, const T* const* g,
                   double scale, double* out, void* ws, hipStream_t s) {
  // nsol independent solves of densities f[k]*g[k]; out[k] = scale * int f g J[f g]
  int k = 0;
  while (k < nsol) {
    PairSrc<T> src{};
    int np = 0;
    const int k0 = k;
    for (; np < pl->max_pairs && k < nsol; ++np) {
      src.f0[np] = f[k];
      src.g0[np] = g[k];
      ++k;
      if (k < nsol) {
        src.f1[np] = f[k];
        src.g1[np] = g[k];
        ++k;
      }
    }
    EpiArgs<T> ea{};
    ea.partials = ws_partials(ws);
    const long NN = (long)pl->N * pl->N;
    const bool fast = fast_z(pl, src, ea, np);
    const long nblk = (NN + z_pencils(pl, fast) - 1) / z_pencils(pl, fast);
    GATO_REQUIRE(nblk * 2 * np <= kPartialsCap, GATO_ERR_UNSUPPORTED,
                 "grid too large for the reduction workspace (N=%d)", pl->N);
    int st = zfwd<T>(pl, src, np, fast, s);
    if (!st) st = middle_passes<T>(pl, np, s);
    if (!st) st = zinv<T, EPI_ENERGY>(pl, src, ea, np, fast, s);
    // slot 2p (2p+1) holds the real (imaginary) solve of pair p, i.e. solve
    // k0 + slot; the empty imaginary slot of an odd tail is not finalized
    if (!st) st = finalize_partials(ws_partials(ws), nblk, k - k0, scale, false, out + k0, s);
    if (st) return st;
  }
  return GATO_OK;
}

// K psi_c for nrhs right-hand sides psi_c = psi + c*ldpsi, into out + c*ldo:
// out_c = (accumulate ? out_c : 0) + coef * sum_j phi_j J[phi_j psi_c].  The
// n_orb solves of every right-hand side go two orbitals per complex pair,
// and the pairs of all right-hand sides fill the launches (up to max_pairs
// pairs each), so the kernel transform is read once per launch.  The pairs of
// one output are consecutive; an output split across two launches is
// accumulated by the second.
template <typename T>
int exchange_apply_t(gato_poisson_plan* pl, int nrhs, const T* psi, long ldpsi, const T* orb,
                     long ld, int n_orb, double coef, int accumulate, T* out, long ldo,
                     hipStream_t s) {
  int c = 0, j = 0;
  bool written = false;  // output c already holds a partial sum
  while (c < nrhs) {
    PairSrc<T> src{};
    EpiArgs<T> ea{};
    int np = 0;
    for (; np < pl->max_pairs && c < nrhs; ++np) {
      const T* g = psi + (long)c * ldpsi;
      src.f0[np] = orb + (long)j * ld;
      src.g0[np] = g;
      ea.h0[np] = orb + (long)j * ld;
      ++j;
      if (j < n_orb) {
        src.f1[np] = orb + (long)j * ld;
        src.g1[np] = g;
        ea.h1[np] = orb + (long)j * ld;
        ++j;
      }
      ea.xout[np] = out + (long)c * ldo;
      // the run of output c ends at this pair if its orbitals are done or the
      // launch is full; it accumulates if the caller asked or an earlier
      // launch wrote part of it
      if ((written ? 1 : accumulate) != 0) ea.accmask |= 1u << np;
      if (j >= n_orb) {
        ++c;
        j = 0;
        written = false;
      } else if (np + 1 == pl->max_pairs) {
        written = true;
      }
    }
    ea.coef = coef;
    const bool fast = fast_z(pl, src, ea, np);
    int st = zfwd<T>(pl, src, np, fast, s);
    if (!st) st = middle_passes<T>(pl, np, s);
    if (!st) st = zinv<T, EPI_EXCHANGE>(pl, src, ea, np, fast, s);
    if (st) return st;
  }
  return GATO_OK;
}

// Opt every FFT kernel in to dynamic LDS beyond 64 KB (all launches stay at
// or below 80 KB; gfx950 has 160 KB per CU).  Failures are ignored: the
// launch itself reports a real problem.
template <typename T, int MC>
void allow_ct_lds(int bytes) {
  (void)hipFuncSetAttribute((const void*)fft_pass_kernel<T, MC>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)zfwd_kernel<T, MC>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)zinv_kernel<T, EPI_STORE, MC>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)zinv_kernel<T, EPI_EXCHANGE, MC>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)zinv_kernel<T, EPI_ENERGY, MC>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

template <typename T, int A, int B, int P>
void allow_fs_mid_lds(int bytes) {
  (void)hipFuncSetAttribute((const void*)fs_pass_kernel<T, A, B, P, FS_FWD>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)fs_pass_kernel<T, A, B, P, FS_X>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)fs_pass_kernel<T, A, B, P, FS_INV>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// The four-step kernels that middle_pass, zfwd and zinv launch for shape A x B.
template <typename T, int A, int B>
void allow_fs_lds(int bytes) {
  if constexpr (A == 8) {
    allow_fs_mid_lds<T, A, B, 32>(bytes);
  } else if constexpr (sizeof(T) == 4) {
    allow_fs_mid_lds<T, A, B, 16>(bytes);
    allow_fs_mid_lds<T, A, B, 32>(bytes);
  } else {
    allow_fs_mid_lds<T, A, B, 16>(bytes);
  }
  constexpr int PF = zfwd_pencils<T, A, B>(), PI = zinv_pencils<A>();
  (void)hipFuncSetAttribute((const void*)fs_zfwd_kernel<T, A, B, PF>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)fs_zinv_kernel<T, A, B, EPI_STORE, PI>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)fs_zinv_kernel<T, A, B, EPI_EXCHANGE, PI>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)fs_zinv_kernel<T, A, B, EPI_ENERGY, PI>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

bool set_large_lds_attributes() {
  const int bytes = 96 * 1024;
  allow_ct_lds<float, 128>(bytes);
  allow_ct_lds<float, 160>(bytes);
  allow_ct_lds<float, 192>(bytes);
  allow_ct_lds<float, 320>(bytes);
  allow_ct_lds<double, 128>(bytes);
  allow_ct_lds<double, 160>(bytes);
  allow_ct_lds<double, 192>(bytes);
  allow_ct_lds<double, 320>(bytes);
  (void)hipFuncSetAttribute((const void*)fft_pass_kernel<float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)fft_pass_kernel<double>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)zfwd_kernel<float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)zfwd_kernel<double>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)zinv_kernel<float, EPI_STORE>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)zinv_kernel<double, EPI_STORE>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)zinv_kernel<float, EPI_EXCHANGE>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)zinv_kernel<double, EPI_EXCHANGE>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)zinv_kernel<float, EPI_ENERGY>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)zinv_kernel<double, EPI_ENERGY>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  // every four-step kernel (the largest, double 16 x 20, takes 94 KB)
  for_each_fs_shape([&](auto ca, auto cb) {
    constexpr int A = decltype(ca)::value, B = decltype(cb)::value;
    allow_fs_lds<float, A, B>(bytes);
    allow_fs_lds<double, A, B>(bytes);
  });
  (void)hipGetLastError();
  return true;
}

// Once per process; a thread that arrives during the first call waits for it
// (function-local static initialisation is thread-safe).
void allow_large_lds() {
  static const bool done = set_large_lds_attributes();
  (void)done;
}

}  // namespace

// internal entry used by the Fock operator
int poisson_exchange(gato_poisson_plan* pl, int nrhs, const void* psi, long ldpsi,
                     const void* orb, long ld, int n_orb, double coef, int accumulate, void* out,
                     long ldo, hipStream_t s) {
  if (pl->dtype == GATO_F32)
    return exchange_apply_t<float>(pl, nrhs, (const float*)psi, ldpsi, (const float*)orb, ld,
                                   n_orb, coef, accumulate, (float*)out, ldo, s);
  return exchange_apply_t<double>(pl, nrhs, (const double*)psi, ldpsi, (const double*)orb, ld,
                                  n_orb, coef, accumulate, (double*)out, ldo, s);
}

int poisson_plan_dtype(const gato_poisson_plan* pl) { return pl->dtype; }
int poisson_plan_N(const gato_poisson_plan* pl) { return pl->N; }

}  // namespace gato

using namespace gato;

extern "C" int gato_poisson_plan_create(int dtype, int N, double h, double epsilon, int max_pairs,
                                        gato_poisson_plan** plan, void* stream) {
  clear_ok();
  GATO_REQUIRE(plan != nullptr, GATO_ERR_VALUE, "null plan pointer");
  *plan = nullptr;
  GATO_REQUIRE(dtype == GATO_F32 || dtype == GATO_F64, GATO_ERR_VALUE, "unknown dtype %d", dtype);
  GATO_REQUIRE(N >= 1 && N <= 512, GATO_ERR_UNSUPPORTED, "grid N=%d outside [1, 512]", N);
  GATO_REQUIRE(h > 0 && epsilon >= 0, GATO_ERR_VALUE, "bad spacing %g or softening %g", h, epsilon);
  GATO_REQUIRE(max_pairs >= 1 && max_pairs <= kMaxPairs, GATO_ERR_VALUE,
               "max_pairs=%d outside [1, %d]", max_pairs, kMaxPairs);
  auto* pl = new gato_poisson_plan();
  pl->dtype = dtype;
  pl->N = N;
  pl->M = 2 * N;
  pl->h = h;
  pl->eps = epsilon;
  pl->dV = h * h * h;
  pl->max_pairs = max_pairs;
  if (!factor_radix(pl->M, pl->rad)) {
    delete pl;
    set_error(GATO_ERR_UNSUPPORTED,
              "Poisson grid 2N=%d has a prime factor other than 2, 3, 5, 7", 2 * N);
    return GATO_ERR_UNSUPPORTED;
  }
  hipStream_t s = as_stream(stream);
  allow_large_lds();
  int st = (dtype == GATO_F32) ? plan_create<float>(pl, s) : plan_create<double>(pl, s);
  if (st) {
    gato_poisson_plan_destroy(pl);
    return st;
  }
  *plan = pl;
  return GATO_OK;
}

extern "C" void gato_poisson_force_pencils(int pencils) {
  gato::g_forced_fs_pencils.store(pencils == 16 || pencils == 32 ? pencils : 0);
}

extern "C" int gato_poisson_plan_destroy(gato_poisson_plan* pl) {
  if (!pl) return GATO_OK;
  (void)hipDeviceSynchronize();
  if (pl->tw) (void)hipFree(pl->tw);
  if (pl->ghat) (void)hipFree(pl->ghat);
  if (pl->A) (void)hipFree(pl->A);
  if (pl->B) (void)hipFree(pl->B);
  delete pl;
  return GATO_OK;
}

extern "C" int gato_hartree_potential(gato_poisson_plan* pl, int n_rho, const void* rho, long ld,
                                      void* J, long ldj, void* stream) {
  clear_ok();
  GATO_REQUIRE(pl && rho && J && n_rho >= 1, GATO_ERR_VALUE, "bad arguments");
  hipStream_t s = as_stream(stream);
  if (pl->dtype == GATO_F32)
    return hartree_potential_t<float>(pl, n_rho, (const float*)rho, ld, (float*)J, ldj, s);
  return hartree_potential_t<double>(pl, n_rho, (const double*)rho, ld, (double*)J, ldj, s);
}

extern "C" int gato_hartree_energy(gato_poisson_plan* pl, int n_rho, const void* rho, long ld,
                                   double* E, void* ws, void* stream) {
  clear_ok();
  GATO_REQUIRE(pl && rho && E && ws && n_rho >= 1 && n_rho <= 64, GATO_ERR_VALUE,
               "bad arguments");
  hipStream_t s = as_stream(stream);
  const size_t es = pl->dtype == GATO_F32 ? sizeof(float) : sizeof(double);
  std::vector<const void*> f(n_rho), g(n_rho, nullptr);
  for (int k = 0; k < n_rho; ++k) f[k] = (const char*)rho + (size_t)k * ld * es;
  if (pl->dtype == GATO_F32)
    return energy_pairs_t<float>(pl, n_rho, (const float* const*)f.data(),
                                 (const float* const*)g.data(), 0.5 * pl->dV, E, ws, s);
  return energy_pairs_t<double>(pl, n_rho, (const double* const*)f.data(),
                                (const double* const*)g.data(), 0.5 * pl->dV, E, ws, s);
}

extern "C" int gato_exchange_apply(gato_poisson_plan* pl, const void* psi, const void* orbitals,
                                   long ld, int n_orb, double coef, int accumulate, void* out,
                                   void* stream) {
  clear_ok();
  GATO_REQUIRE(pl && psi && orbitals && out && n_orb >= 1, GATO_ERR_VALUE, "bad arguments");
  return poisson_exchange(pl, 1, psi, 0, orbitals, ld, n_orb, coef, accumulate, out, 0,
                          as_stream(stream));
}

extern "C" int gato_exchange_pair_integrals(gato_poisson_plan* pl, const void* orbitals, long ld,
                                            int n_orb, double* out, void* ws, void* stream) {
  clear_ok();
  GATO_REQUIRE(pl && orbitals && out && ws && n_orb >= 1 && n_orb <= 16, GATO_ERR_VALUE,
               "bad arguments");
  hipStream_t s = as_stream(stream);
  const size_t es = pl->dtype == GATO_F32 ? sizeof(float) : sizeof(double);
  std::vector<const void*> f, g;
  for (int i = 0; i < n_orb; ++i)
    for (int j = i; j < n_orb; ++j) {
      f.push_back((const char*)orbitals + (size_t)i * ld * es);
      g.push_back((const char*)orbitals + (size_t)j * ld * es);
    }
  const int nsol = (int)f.size();
  if (pl->dtype == GATO_F32)
    return energy_pairs_t<float>(pl, nsol, (const float* const*)f.data(),
                                 (const float* const*)g.data(), pl->dV, out, ws, s);
  return energy_pairs_t<double>(pl, nsol, (const double* const*)f.data(),
                                (const double* const*)g.data(), pl->dV, out, ws, s);
}
