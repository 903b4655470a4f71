// Shared helpers for the gato_amd HIP library (gfx950 / CDNA4).
//
// Conventions used by every kernel in this library:
//  * A field on an N^3 grid is C-ordered [x][y][z] with z contiguous; a stack
//    of fields is [k][x][y][z].
//  * T is the storage/compute type of fields (float or double); every
//    reduction (dot products, energies, Lanczos coefficients) accumulates in
//    double with compensated (Neumaier) summation.
//  * Entry points return 0 on success and a nonzero gato_status otherwise;
//    the message is kept per host thread and read with gato_last_error().
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/gato_amd.h"

namespace gato {

void set_error(int code, const char* fmt, ...);
int clear_ok();

#define GATO_HIP_CHECK(expr)                                                   \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      ::gato::set_error(GATO_ERR_HIP, "%s failed: %s (%s:%d)", #expr,          \
                        hipGetErrorString(_e), __FILE__, __LINE__);            \
      return GATO_ERR_HIP;                                                     \
    }                                                                          \
  } while (0)

#define GATO_LAUNCH_CHECK()                                                    \
  do {                                                                         \
    hipError_t _e = hipGetLastError();                                         \
    if (_e != hipSuccess) {                                                    \
      ::gato::set_error(GATO_ERR_HIP, "kernel launch failed: %s (%s:%d)",      \
                        hipGetErrorString(_e), __FILE__, __LINE__);            \
      return GATO_ERR_HIP;                                                     \
    }                                                                          \
  } while (0)

#define GATO_REQUIRE(cond, code, ...)                                          \
  do {                                                                         \
    if (!(cond)) {                                                             \
      ::gato::set_error((code), __VA_ARGS__);                                  \
      return (code);                                                           \
    }                                                                          \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------------------
// Complex helpers (interleaved re/im pairs).
// ---------------------------------------------------------------------------
template <typename T> struct cplx_t;
template <> struct cplx_t<float> { using type = float2; };
template <> struct cplx_t<double> { using type = double2; };
template <typename T> using cplx = typename cplx_t<T>::type;

template <typename C> __device__ __forceinline__ C cmk(decltype(C{}.x) r, decltype(C{}.x) i) {
  C c; c.x = r; c.y = i; return c;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by conj(b)
template <typename C> __device__ __forceinline__ C cmulc(C a, C b) {
  return cmk<C>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename C, typename R> __device__ __forceinline__ C cscale(C a, R s) { return cmk<C>(a.x * s, a.y * s); }
// multiply by -i (dir=-1) or +i (dir=+1)
template <typename C> __device__ __forceinline__ C cmul_i(C a, int dir) {
  return dir < 0 ? cmk<C>(a.y, -a.x) : cmk<C>(-a.y, a.x);
}

// ---------------------------------------------------------------------------
// Compensated double accumulation (Neumaier / TwoSum).
// ---------------------------------------------------------------------------
struct KSum {
  double s, c;
  __device__ __forceinline__ void init() { s = 0.0; c = 0.0; }
  __device__ __forceinline__ void add(double x) {
    double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ double value() const { return s + c; }
};

// s += p[first], p[first + step], ... below np, in that order, with the loads
// of 8 terms issued before their additions (the partials come from L2; one
// dependent load per term would serialise their latencies).
__device__ __forceinline__ void ksum_strided(KSum& s, const double* p, long np, long first,
                                             long step) {
  long b = first;
  for (; b + 7 * step < np; b += 8 * step) {
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = p[b + u * step];
#pragma unroll
    for (int u = 0; u < 8; ++u) s.add(t[u]);
  }
  for (; b < np; b += step) s.add(p[b]);
}

// Wavefront (64-lane) tree reduction of a double via cross-lane shuffles.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block reduction of one double (blockDim.x*blockDim.y*blockDim.z <= 1024).
// Result valid in thread 0. `sh` must hold >= 16 doubles.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int nthr = blockDim.x * blockDim.y * blockDim.z;
  v = wave_sum(v);
  const int lane = tid & 63, wid = tid >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (tid == 0) {
    const int nw = (nthr + 63) >> 6;
    for (int w = 0; w < nw; ++w) r += sh[w];
  }
  return r;
}

// ---------------------------------------------------------------------------
// Vector loads: VEC consecutive elements as one 16-byte (float4 / double2) or
// scalar access, widened to double.
// ---------------------------------------------------------------------------
template <typename T, int VEC> struct VecT;
template <> struct VecT<float, 4> { using type = float4; };
template <> struct VecT<double, 2> { using type = double2; };

template <typename T, int VEC>
__device__ __forceinline__ void vload(const T* __restrict__ p, long i, double* o) {
  if constexpr (VEC == 1) {
    o[0] = (double)p[i];
  } else if constexpr (VEC == 2) {
    const double2 v = *reinterpret_cast<const double2*>(p + i);
    o[0] = v.x; o[1] = v.y;
  } else {
    const float4 v = *reinterpret_cast<const float4*>(p + i);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  }
}

// vload with the non-temporal hint.
template <typename T, int VEC>
__device__ __forceinline__ void vload_nt(const T* __restrict__ p, long i, double* o) {
  if constexpr (VEC == 1) {
    o[0] = (double)__builtin_nontemporal_load(p + i);
  } else if constexpr (VEC == 2) {
    typedef double v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p + i));
    o[0] = v.x; o[1] = v.y;
  } else {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p + i));
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  }
}

template <typename T, int VEC>
__device__ __forceinline__ void vstore(T* __restrict__ p, long i, const double* a) {
  if constexpr (VEC == 1) {
    p[i] = (T)a[0];
  } else if constexpr (VEC == 2) {
    *reinterpret_cast<double2*>(p + i) = make_double2(a[0], a[1]);
  } else {
    *reinterpret_cast<float4*>(p + i) = make_float4((float)a[0], (float)a[1], (float)a[2], (float)a[3]);
  }
}

// Raw 16-byte vector (or scalar) load, kept unconverted until use, and its
// dot product with VEC doubles.
template <typename T, int VEC> struct RawVec { using type = T; };
template <> struct RawVec<float, 4> { using type = float4; };
template <> struct RawVec<double, 2> { using type = double2; };

template <typename T, int VEC>
__device__ __forceinline__ typename RawVec<T, VEC>::type rload(const T* __restrict__ p, long i) {
  return *reinterpret_cast<const typename RawVec<T, VEC>::type*>(p + i);
}

// The same load with the non-temporal hint (streamed data that this pass does
// not read again).
template <typename T, int VEC>
__device__ __forceinline__ typename RawVec<T, VEC>::type rload_nt(const T* __restrict__ p, long i) {
  if constexpr (VEC == 4) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 r = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p + i));
    return make_float4(r.x, r.y, r.z, r.w);
  } else if constexpr (VEC == 2) {
    typedef double v2 __attribute__((ext_vector_type(2)));
    const v2 r = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p + i));
    return make_double2(r.x, r.y);
  } else {
    return __builtin_nontemporal_load(p + i);
  }
}

template <typename T, int VEC>
__device__ __forceinline__ typename RawVec<T, VEC>::type rzero() {
  if constexpr (VEC == 4) return make_float4(0.f, 0.f, 0.f, 0.f);
  else if constexpr (VEC == 2) return make_double2(0.0, 0.0);
  else return T(0);
}

// element u (a compile-time index) of a raw vector, widened to double
template <typename T, int VEC>
__device__ __forceinline__ double rget(const typename RawVec<T, VEC>::type& a, int u) {
  if constexpr (VEC == 4) {
    return u == 0 ? a.x : u == 1 ? a.y : u == 2 ? a.z : a.w;
  } else if constexpr (VEC == 2) {
    return u == 0 ? a.x : a.y;
  } else {
    return (double)a;
  }
}

template <typename T, int VEC>
__device__ __forceinline__ double rdot(const typename RawVec<T, VEC>::type& a, const double* b) {
  if constexpr (VEC == 1) {
    return (double)a * b[0];
  } else if constexpr (VEC == 2) {
    return a.x * b[0] + a.y * b[1];
  } else {
    return (double)a.x * b[0] + (double)a.y * b[1] + (double)a.z * b[2] + (double)a.w * b[3];
  }
}

inline bool aligned16(const void* p) { return p == nullptr || ((uintptr_t)p & 15u) == 0; }

// Sum over the 64 lanes of a wave of NV values per lane (NV a power of two,
// at most 64), scattered: the first log2(NV) butterfly steps each keep half of
// the values and add the partner lane's copy of that half, the rest add the
// single remaining value.  On return every lane holds in v[0] the wave total of
// value index wave_scatter_index<NV>(lane), and the lanes that agree in their
// top log2(NV) bits hold the same index.  Costs NV - 1 + log2(64/NV) shuffles
// instead of NV * 6 for NV separate wave sums.
template <int NV>
__device__ __forceinline__ void wave_reduce_scatter(double (&v)[NV], int lane) {
#pragma unroll
  for (int s = 0, h = NV / 2; h >= 1; ++s, h >>= 1) {
    const int mask = 32 >> s;
    const bool up = (lane & mask) != 0;
#pragma unroll
    for (int j = 0; j < h; ++j) {
      const double send = up ? v[j] : v[j + h];
      const double keep = up ? v[j + h] : v[j];
      v[j] = keep + __shfl_xor(send, mask, 64);
    }
  }
#pragma unroll
  for (int mask = 32 / NV; mask >= 1; mask >>= 1) v[0] += __shfl_xor(v[0], mask, 64);
}

template <int NV>
__device__ __forceinline__ int wave_scatter_index(int lane) {
  int idx = 0;
#pragma unroll
  for (int s = 0, h = NV / 2; h >= 1; ++s, h >>= 1)
    if (lane & (32 >> s)) idx += h;
  return idx;
}

// Two-kernel deterministic grid reduction.  A reduction kernel stores one
// partial per block (or wave) into partials[slot*np + b] with plain stores;
// finalize_partials() then launches one block per slot on the same stream:
// out[slot] = scale * sum_b partials[slot*np + b] (sqrt(max(., 0)) when root),
// summed in a fixed order.  The kernel boundary makes the partials visible;
// no in-kernel device-scope fence is needed (on gfx950 every such fence
// writes back and invalidates the XCD's L2 under the streaming blocks).
int finalize_partials(const double* partials, long np, int nslots, double scale, bool root,
                      double* out, hipStream_t s);

// Workspace layout for reductions: a 256-byte header (unused, kept so the
// partials start 256-byte aligned), then kPartialsCap double partials, then
// kStageCap doubles of staging.  Callers pass at least
// gato_reduce_workspace_bytes() bytes; nothing in it has to start at zero.
constexpr size_t kPartialsOffset = 256;  // bytes
constexpr long kPartialsCap = 1L << 21;  // doubles (16 MiB)
constexpr long kStageCap = 4096;         // doubles
constexpr int kReduceBlocks = 1024;      // max blocks of a plain grid reduction
constexpr size_t kReduceWorkspaceBytes = kPartialsOffset + (kPartialsCap + kStageCap) * 8;

// ---------------------------------------------------------------------------
// Launch probe: when enabled for one kernel kind, HIP events bracket every
// launch of that kind on its own stream and the algorithmic bytes of each
// launch are summed; gato_probe_read() reports launches, GPU milliseconds
// and bytes.  Off by default (one branch per launch).  The reset argument is
// kind | (element_bytes << 8): a non-zero element size restricts the probe to
// launches on fields of that element size (4: float, 8: double).
// ---------------------------------------------------------------------------
enum ProbeKind {
  PROBE_NONE = 0,
  PROBE_CGS_UPDATE = 1,   // cgs_update_kernel (reorthogonalisation update)
  PROBE_CGS_DOT = 2,      // proj2_kernel / gemvt_kernel with nk > 2 (reorthogonalisation
                          // coefficients)
  PROBE_FFT_MIDDLE = 3,   // the three strided FFT passes of a Poisson solve
  PROBE_FFT_Z = 4,        // z-forward and z-inverse passes
  PROBE_BLOCK_DOT = 5,    // block projection D = V^T X on the matrix cores (block Lanczos)
  PROBE_BLOCK_UPDATE = 6, // block update Y = X - V C (block Lanczos)
};
bool probe_on(int kind, int elem_bytes);

// Measured variant choice (stencil variants, FFT pencils per block) is on
// unless gato_set_autotune(0) switched it off.
bool autotune_on();
void probe_begin(hipStream_t s);
void probe_end(hipStream_t s, double bytes);

inline double* ws_partials(void* ws) {
  return reinterpret_cast<double*>(reinterpret_cast<char*>(ws) + kPartialsOffset);
}
inline double* ws_stage(void* ws) { return ws_partials(ws) + kPartialsCap; }

}  // namespace gato
