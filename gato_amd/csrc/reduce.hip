// Reductions and Krylov-basis kernels: batched dot products, the classical
// Gram-Schmidt update, Ritz-vector reconstruction, density mixing.
//
// Reference: grid.py:57-73 (midpoint-rule integrals and inner products),
// lanczos.py:140-194 (alpha/beta, full reorthogonalisation, Ritz vectors),
// scf.py:61-63 and :422-423 (density and linear mixing).
//
// Every sum accumulates in double.  The streaming kernels (gemvt_kernel,
// proj2_kernel, cgs_update_kernel, the Ritz kernels) add their products per
// lane in plain double over a lane's few dozen elements, then reduce the 64
// lanes with a shuffle tree into one partial per wave; the finalize kernels
// sum those partials with per-thread Neumaier compensation and a block tree.
// dot_pairs_kernel and the scaling-pass partials compensate per thread as
// well.  Every reduction sums in a fixed order, so results are bitwise
// reproducible run to run.
#include <mutex>
#include "common.hpp"

namespace gato {
namespace {

constexpr int kThr = 256;

// Sum of `np` partials of one slot by the whole block (compensated per thread,
// then a tree).  Result valid in thread 0.
__device__ __forceinline__ double block_sum_partials(const double* p, long np, double* red) {
  KSum s;
  s.init();
  ksum_strided(s, p, np, threadIdx.x, blockDim.x);
  return block_sum(s.value(), red);
}

// ---------------------------------------------------------------------------
// Batched dot against one broadcast vector (the reorthogonalisation
// coefficients c = V^T b'):  part[k*nw + wave] = sum over the wave's chunk of
// A_k * b',  b' = (B - s*S) * C.  Each 64-lane wave owns a contiguous chunk of
// 64*VEC*EPT elements, keeps b' for it in registers, and streams the nk
// vectors through it two at a time with 16-byte loads; wave partials need no
// block synchronisation and are summed by finalize_partials.
// ---------------------------------------------------------------------------
template <typename T, int VEC, int EPT>
__global__ __launch_bounds__(kThr) void gemvt_kernel(long n, int nk, const T* __restrict__ A,
                                                     long lda, const T* __restrict__ B,
                                                     const T* __restrict__ C,
                                                     const T* __restrict__ S,
                                                     const double* __restrict__ s_coef,
                                                     double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const long nw = (long)gridDim.x * (kThr / 64);
  const long wave = (long)blockIdx.x * (kThr / 64) + (threadIdx.x >> 6);
  constexpr long CH = 64L * VEC * EPT;
  const long base = wave * CH;
  const double sc = S ? s_coef[0] : 0.0;
  double b[EPT][VEC];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const long i = base + ((long)e * 64 + lane) * VEC;
    if (i < n) {
      vload<T, VEC>(B, i, b[e]);
      if (S) {
        double t[VEC];
        vload<T, VEC>(S, i, t);
#pragma unroll
        for (int v = 0; v < VEC; ++v) b[e][v] -= sc * t[v];
      }
      if (C) {
        double t[VEC];
        vload<T, VEC>(C, i, t);
#pragma unroll
        for (int v = 0; v < VEC; ++v) b[e][v] *= t[v];
      }
    } else {
#pragma unroll
      for (int v = 0; v < VEC; ++v) b[e][v] = 0.0;
    }
  }
  // two basis vectors in flight: each wave streams 2 x 64*VEC*EPT contiguous
  // elements per step (long runs keep DRAM pages open).  Full chunks take an
  // unguarded path that keeps the loads as raw vectors until the FMAs.
  int k = 0;
  if (base + CH <= n) {
    using RV = typename RawVec<T, VEC>::type;
    for (; k + 2 <= nk; k += 2) {
      RV t0[EPT], t1[EPT];
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const long i = base + ((long)e * 64 + lane) * VEC;
        t0[e] = rload<T, VEC>(A + (long)k * lda, i);
        t1[e] = rload<T, VEC>(A + (long)(k + 1) * lda, i);
      }
      double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        acc0 += rdot<T, VEC>(t0[e], b[e]);
        acc1 += rdot<T, VEC>(t1[e], b[e]);
      }
      const double r0 = wave_sum(acc0);
      const double r1 = wave_sum(acc1);
      if (lane == 0) {
        part[(long)k * nw + wave] = r0;
        part[(long)(k + 1) * nw + wave] = r1;
      }
    }
  }
  for (; k + 2 <= nk; k += 2) {
    double a0[EPT][VEC], a1[EPT][VEC];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const long i = base + ((long)e * 64 + lane) * VEC;
      if (i < n) {
        vload<T, VEC>(A + (long)k * lda, i, a0[e]);
        vload<T, VEC>(A + (long)(k + 1) * lda, i, a1[e]);
      }
    }
    double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const long i = base + ((long)e * 64 + lane) * VEC;
      if (i < n) {
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          acc0 += a0[e][v] * b[e][v];
          acc1 += a1[e][v] * b[e][v];
        }
      }
    }
    const double r0 = wave_sum(acc0);
    const double r1 = wave_sum(acc1);
    if (lane == 0) {
      part[(long)k * nw + wave] = r0;
      part[(long)(k + 1) * nw + wave] = r1;
    }
  }
  if (k < nk) {
    double acc = 0.0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const long i = base + ((long)e * 64 + lane) * VEC;
      if (i < n) {
        double a[VEC];
        vload<T, VEC>(A + (long)k * lda, i, a);
#pragma unroll
        for (int v = 0; v < VEC; ++v) acc += a[v] * b[e][v];
      }
    }
    const double r = wave_sum(acc);
    if (lane == 0) part[(long)k * nw + wave] = r;
  }
}

// out[k] = scale * sum of the np partials of slot k (sqrt(max(., 0)) when
// root)   (grid: one 1024-thread block per slot; each thread issues four
// independent loads before accumulating them, in a fixed order)
constexpr int kFinThr = 1024;
__global__ __launch_bounds__(kFinThr) void finalize_kernel(const double* __restrict__ part,
                                                           long np, double scale, int root,
                                                           double* __restrict__ out) {
  __shared__ double red[16];
  const double* p = part + (long)blockIdx.x * np;
  KSum s;
  s.init();
  long b = threadIdx.x;
  for (; b + 3L * kFinThr < np; b += 4L * kFinThr) {
    const double v0 = p[b], v1 = p[b + kFinThr], v2 = p[b + 2L * kFinThr], v3 = p[b + 3L * kFinThr];
    s.add(v0);
    s.add(v1);
    s.add(v2);
    s.add(v3);
  }
  for (; b < np; b += kFinThr) s.add(p[b]);
  const double t = block_sum(s.value(), red);
  if (threadIdx.x == 0) out[blockIdx.x] = root ? sqrt(fmax(t * scale, 0.0)) : t * scale;
}

// Lanczos projections of step j in one pass over the basis: wave partials of
// d_k = <A_k, w> go to part[k*nw + wave] and of g_k = <A_k, v_j> to
// part[(nk + k)*nw + wave], k < nk.  A is the whole basis (v_j = A[nk-1])
// or, when the partials of the whole basis exceed the workspace, one slab of
// it, with v_j passed separately either way.  Layout and loads as
// gemvt_kernel with two broadcast vectors.  Grid row blockIdx.y takes basis
// vectors [y*kper, (y+1)*kper): on grids too small to give every wave a wide
// chunk, the vectors are split among more waves instead of the elements, so
// each wave still reduces once per 64*VEC*EPT elements of a vector.
template <typename T, int VEC, int EPT>
__global__ __launch_bounds__(kThr) void proj2_kernel(long n, int nk, int kper,
                                                     const T* __restrict__ A, long lda,
                                                     const T* __restrict__ vj,
                                                     const T* __restrict__ w,
                                                     double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int klo = blockIdx.y * kper;
  const int khi = klo + kper < nk ? klo + kper : nk;
  const long nw = (long)gridDim.x * (kThr / 64);
  const long wave = (long)blockIdx.x * (kThr / 64) + (threadIdx.x >> 6);
  constexpr long CH = 64L * VEC * EPT;
  const long base = wave * CH;
  double bw[EPT][VEC], bv[EPT][VEC];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const long i = base + ((long)e * 64 + lane) * VEC;
    if (i < n) {
      vload<T, VEC>(w, i, bw[e]);
      vload<T, VEC>(vj, i, bv[e]);
    } else {
#pragma unroll
      for (int v = 0; v < VEC; ++v) bw[e][v] = bv[e][v] = 0.0;
    }
  }
  int k = klo;
  if (base + CH <= n) {
    using RV = typename RawVec<T, VEC>::type;
    for (; k + 2 <= khi; k += 2) {
      RV t0[EPT], t1[EPT];
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const long i = base + ((long)e * 64 + lane) * VEC;
        t0[e] = rload_nt<T, VEC>(A + (long)k * lda, i);
        t1[e] = rload_nt<T, VEC>(A + (long)(k + 1) * lda, i);
      }
      double d0 = 0.0, d1 = 0.0, g0 = 0.0, g1 = 0.0;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        d0 += rdot<T, VEC>(t0[e], bw[e]);
        g0 += rdot<T, VEC>(t0[e], bv[e]);
        d1 += rdot<T, VEC>(t1[e], bw[e]);
        g1 += rdot<T, VEC>(t1[e], bv[e]);
      }
      d0 = wave_sum(d0);
      g0 = wave_sum(g0);
      d1 = wave_sum(d1);
      g1 = wave_sum(g1);
      if (lane == 0) {
        part[(long)k * nw + wave] = d0;
        part[(long)(nk + k) * nw + wave] = g0;
        part[(long)(k + 1) * nw + wave] = d1;
        part[(long)(nk + k + 1) * nw + wave] = g1;
      }
    }
  }
  for (; k < khi; ++k) {
    double d = 0.0, g = 0.0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const long i = base + ((long)e * 64 + lane) * VEC;
      if (i < n) {
        double a[VEC];
        vload<T, VEC>(A + (long)k * lda, i, a);
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          d += a[v] * bw[e][v];
          g += a[v] * bv[e][v];
        }
      }
    }
    d = wave_sum(d);
    g = wave_sum(g);
    if (lane == 0) {
      part[(long)k * nw + wave] = d;
      part[(long)(nk + k) * nw + wave] = g;
    }
  }
}

// Block k: D_k, G_k and D_j (j = nk - 1) from the proj2 partials (fixed
// order); alpha = scale*D_j, coef[k] = scale*D_k - alpha*scale*G_k
// (= <v_k, w - alpha v_j>); block j also writes alpha.  alpha_given: the
// slab holds no v_j, and alpha comes from alpha[0], written by the finalize
// of the slab that holds it (launched earlier on the same stream).
__global__ __launch_bounds__(kFinThr) void proj2_finalize_kernel(const double* __restrict__ part,
                                                                 long np, int nk, double scale,
                                                                 double* __restrict__ coef,
                                                                 double* __restrict__ alpha,
                                                                 int alpha_given) {
  __shared__ double red[16];
  const int k = blockIdx.x;
  const double* pd = part + (long)k * np;
  const double* pg = part + (long)(nk + k) * np;
  const double* pj = part + (long)(nk - 1) * np;
  KSum sd, sg, sj;
  sd.init();
  sg.init();
  sj.init();
  long b = threadIdx.x;
  for (; b + 3 * kFinThr < np; b += 4 * kFinThr) {  // 12 loads issued before the adds
    double a[4], c[4], e[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = pd[b + u * kFinThr];
      c[u] = pg[b + u * kFinThr];
      e[u] = pj[b + u * kFinThr];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      sd.add(a[u]);
      sg.add(c[u]);
      sj.add(e[u]);
    }
  }
  for (; b < np; b += kFinThr) {
    const double a = pd[b], c = pg[b], e = pj[b];
    sd.add(a);
    sg.add(c);
    sj.add(e);
  }
  const double D = block_sum(sd.value(), red);
  const double G = block_sum(sg.value(), red);
  const double Dj = block_sum(sj.value(), red);
  if (threadIdx.x == 0) {
    const double al = alpha_given ? alpha[0] : scale * Dj;
    coef[k] = scale * D - al * (scale * G);
    if (!alpha_given && k == nk - 1) alpha[0] = al;
  }
}

// Pairwise dots out[k] = scale * sum_i A_k[i] B_k[i] C[i]  (ldb != 0; small nk:
// energies).  grid (nblk, nk); block partials part[k*nblk + b].
template <typename T>
__global__ __launch_bounds__(kThr) void dot_pairs_kernel(long n, int nk, const T* __restrict__ A,
                                                         long lda, const T* __restrict__ B,
                                                         long ldb, const T* __restrict__ C,
                                                         double* part) {
  __shared__ double red[16];
  const int nblk = gridDim.x;
  const int k = blockIdx.y;
  const T* a = A + (long)k * lda;
  const T* b = B + (long)k * ldb;
  KSum acc;
  acc.init();
  const long stride = (long)nblk * kThr;
  for (long i = (long)blockIdx.x * kThr + threadIdx.x; i < n; i += stride) {
    double t = (double)a[i] * (double)b[i];
    if (C) t *= (double)C[i];
    acc.add(t);
  }
  const double v = block_sum(acc.value(), red);
  if (threadIdx.x == 0) part[(long)k * nblk + blockIdx.x] = v;
}

// w_out = w - s*S - sum_k coef[k] V[k];  part[wave] = the wave's share of
// ||w_out||^2 (beta = sqrt(scale * sum) is formed by finalize_partials or by
// scale_inv_partials_kernel).  Each 64-lane wave owns a contiguous chunk of
// 64*VEC*EPT elements, keeps w' for it in double registers, and streams the
// basis vectors through it two at a time with 16-byte loads, from V[nk-1]
// down to V[0]: the projection pass that precedes this update walks the
// basis upwards, so the vectors it read last are still cached when this pass
// starts (measured: 4% less time for the pair at 40-160 vectors of 128^3).
template <typename T, int VEC, int EPT>
__global__ __launch_bounds__(kThr) void cgs_update_kernel(long n, int nk, const T* __restrict__ V,
                                                          long ldv, const double* __restrict__ coef,
                                                          const T* __restrict__ S,
                                                          const double* __restrict__ s_coef,
                                                          const T* w, T* w_out, double* part) {
  extern __shared__ double cs[];  // nk coefficients
  for (int k = threadIdx.x; k < nk; k += kThr) cs[k] = coef[k];
  __syncthreads();
  const double sc = S ? s_coef[0] : 0.0;
  const int lane = threadIdx.x & 63;
  const long wave = (long)blockIdx.x * (kThr / 64) + (threadIdx.x >> 6);
  constexpr long CH = 64L * VEC * EPT;
  const long base = wave * CH;
  double a[EPT][VEC];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const long i = base + ((long)e * 64 + lane) * VEC;
    if (i < n) {
      vload<T, VEC>(w, i, a[e]);
      if (S) {
        double t[VEC];
        vload<T, VEC>(S, i, t);
#pragma unroll
        for (int v = 0; v < VEC; ++v) a[e][v] -= sc * t[v];
      }
    }
  }
  int q = 0;
  for (; q + 2 <= nk; q += 2) {
    const int k0 = nk - 1 - q, k1 = nk - 2 - q;
    // raw 16-byte loads, widened at the FMA (keeps 8-load chunks in registers)
    using RV = typename RawVec<T, VEC>::type;
    RV t0[EPT], t1[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const long i = base + ((long)e * 64 + lane) * VEC;
      if (i < n) {
        t0[e] = rload_nt<T, VEC>(V + (long)k0 * ldv, i);
        t1[e] = rload_nt<T, VEC>(V + (long)k1 * ldv, i);
      }
    }
    const double c0 = cs[k0], c1 = cs[k1];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const long i = base + ((long)e * 64 + lane) * VEC;
      if (i < n) {
        const T* e0 = reinterpret_cast<const T*>(&t0[e]);
        const T* e1 = reinterpret_cast<const T*>(&t1[e]);
#pragma unroll
        for (int v = 0; v < VEC; ++v) a[e][v] -= c0 * (double)e0[v] + c1 * (double)e1[v];
      }
    }
  }
  if (q < nk) {  // odd nk: V[0] is left
    const double c0 = cs[0];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const long i = base + ((long)e * 64 + lane) * VEC;
      if (i < n) {
        double t0[VEC];
        vload<T, VEC>(V, i, t0);
#pragma unroll
        for (int v = 0; v < VEC; ++v) a[e][v] -= c0 * t0[v];
      }
    }
  }
  double nrm = 0.0;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const long i = base + ((long)e * 64 + lane) * VEC;
    if (i < n) {
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const double r = (double)(T)a[e][v];
        nrm += r * r;
      }
      vstore<T, VEC>(w_out, i, a[e]);
    }
  }
  nrm = wave_sum(nrm);
  if (lane == 0) part[wave] = nrm;
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

// beta = sqrt(max(scale * sum of the np partials, 0)) and v = w / beta (0 when
// beta == 0).  Every block sums the partials in the same fixed order, so all
// blocks hold the same beta; block 0 stores it.  This takes the place of a
// finalize launch followed by scale_inv_kernel.
template <typename T, int VEC>
__global__ __launch_bounds__(kThr) void scale_inv_partials_kernel(long n, const T* w,
                                                                  const double* __restrict__ part,
                                                                  long np, double scale,
                                                                  double* __restrict__ beta_out,
                                                                  T* v) {
  __shared__ double red[16];
  __shared__ double s_inv;
  // the thread's first U vectors are loaded before the partials are summed,
  // so the two latencies overlap (w and v may be the same field: every
  // element is loaded and stored by one thread)
  constexpr int U = 4;
  const long nv = n / VEC;
  const long stride = (long)gridDim.x * kThr;
  const long q0 = (long)blockIdx.x * kThr + threadIdx.x;
  double a[U][VEC];
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (q0 + u * stride < nv) vload<T, VEC>(w, (q0 + u * stride) * VEC, a[u]);
  KSum acc;
  acc.init();
  ksum_strided(acc, part, np, threadIdx.x, kThr);
  const double t = block_sum(acc.value(), red);
  if (threadIdx.x == 0) {
    const double b = sqrt(fmax(t * scale, 0.0));
    s_inv = b > 0.0 ? 1.0 / b : 0.0;
    if (blockIdx.x == 0) beta_out[0] = b;
  }
  __syncthreads();
  const double inv = s_inv;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (q0 + u * stride < nv) {
#pragma unroll
      for (int e = 0; e < VEC; ++e) a[u][e] *= inv;
      vstore<T, VEC>(v, (q0 + u * stride) * VEC, a[u]);
    }
  }
  for (long q = q0 + U * stride; q < nv; q += stride) {
    double c[VEC];
    vload<T, VEC>(w, q * VEC, c);
#pragma unroll
    for (int e = 0; e < VEC; ++e) c[e] *= inv;
    vstore<T, VEC>(v, q * VEC, c);
  }
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

// The same combination for K <= 4 with VEC consecutive elements per thread:
// 16-byte non-temporal loads of the basis (each vector is read once), two
// vectors in flight, and the accumulation order of ritz_kernel (so the
// results are the same bit for bit).
template <typename T, int VEC>
__global__ __launch_bounds__(kThr) void ritz4_kernel(long n, int m, int K,
                                                     const T* __restrict__ V, long ldv,
                                                     const double* __restrict__ S,
                                                     T* __restrict__ out, long ldo) {
  extern __shared__ double sS[];  // m*K
  for (int i = threadIdx.x; i < m * K; i += kThr) sS[i] = S[i];
  __syncthreads();
  const long nv = n / VEC;
  const long stride = (long)gridDim.x * kThr;
  for (long q = (long)blockIdx.x * kThr + threadIdx.x; q < nv; q += stride) {
    const long i = q * VEC;
    double acc[4][VEC];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int v = 0; v < VEC; ++v) acc[k][v] = 0.0;
    int j = 0;
    for (; j + 2 <= m; j += 2) {
      double a[VEC], b[VEC];
      vload_nt<T, VEC>(V + (long)j * ldv, i, a);
      vload_nt<T, VEC>(V + (long)(j + 1) * ldv, i, b);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (k < K) {
          const double s0 = sS[j * K + k], s1 = sS[(j + 1) * K + k];
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            acc[k][v] += s0 * a[v];
            acc[k][v] += s1 * b[v];
          }
        }
      }
    }
    if (j < m) {
      double a[VEC];
      vload_nt<T, VEC>(V + (long)j * ldv, i, a);
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k < K) {
          const double s0 = sS[j * K + k];
#pragma unroll
          for (int v = 0; v < VEC; ++v) acc[k][v] += s0 * a[v];
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k < K) vstore<T, VEC>(out + (long)k * ldo, i, acc[k]);
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

int finalize_partials(const double* partials, long np, int nslots, double scale, bool root,
                      double* out, hipStream_t s) {
  if (nslots <= 0) return GATO_OK;
  hipLaunchKernelGGL(finalize_kernel, dim3(nslots), dim3(kFinThr), 0, s, partials, np, scale,
                     root ? 1 : 0, out);
  GATO_LAUNCH_CHECK();
  return GATO_OK;
}

// np_out != nullptr: leave the wave partials (slot k at k * *np_out) in the
// workspace instead of finalizing them into out.
template <typename T, int VEC, int EPT>
int launch_gemvt(long n, int nk, const void* A, long lda, const void* B, const void* C,
                 const void* S, const double* s_coef, double scale, double* out, void* ws,
                 hipStream_t s, long* np_out, long cap) {
  constexpr long CH = 64L * VEC * EPT;
  const long nw0 = (n + CH - 1) / CH;
  const long nblk = (nw0 + 3) / 4;
  const long nw = nblk * 4;
  GATO_REQUIRE(nw * nk <= cap, GATO_ERR_UNSUPPORTED,
               "%d vectors of %ld elements need %ld partials, more than the %ld of the reduction "
               "workspace", nk, n, nw * nk, cap);
  const bool probed = nk > 2 && probe_on(PROBE_CGS_DOT, (int)sizeof(T));
  if (probed) probe_begin(s);
  hipLaunchKernelGGL((gemvt_kernel<T, VEC, EPT>), dim3((unsigned)nblk), dim3(kThr), 0, s, n, nk,
                     (const T*)A, lda, (const T*)B, (const T*)C, (const T*)S, s_coef,
                     ws_partials(ws));
  GATO_LAUNCH_CHECK();
  // algorithmic bytes: nk basis vectors + the broadcast vector (+ S, C) read once
  if (probed)
    probe_end(s, (double)n * sizeof(T) * (nk + 1 + (S ? 1 : 0) + (C ? 1 : 0)));
  if (np_out) {
    *np_out = nw;
    return GATO_OK;
  }
  return finalize_partials(ws_partials(ws), nw, nk, scale, false, out, s);
}

// The widest chunk per wave that still gives >= 1024 waves (longer contiguous
// runs, fewer partials); small vectors use 2 and widen only when their
// partials would not fit the workspace.  The choice depends on the size of a
// full reduction workspace only; ``cap`` (the partials the caller's scratch
// actually holds) is checked after it, so a smaller scratch runs the same
// kernel in the same summation order or refuses the call.
template <typename T, int VEC>
int gemvt_ept(long n, int nk, const void* A, long lda, const void* B, const void* C, const void* S,
              const double* s_coef, double scale, double* out, void* ws, hipStream_t s,
              long* np_out = nullptr, long cap = kPartialsCap) {
  auto waves = [&](long ept) { return (n + 64L * VEC * ept - 1) / (64L * VEC * ept); };
  auto fits = [&](long ept) { return (waves(ept) + 4) * nk <= kPartialsCap; };
  if (!fits(8) && !np_out && cap == kPartialsCap) {
    // more vectors than one pass can hold partials for: slabs of vectors,
    // each reduced and finalized on its own (the outputs are independent)
    const long per = kPartialsCap / (waves(8) + 4);
    GATO_REQUIRE(per >= 1, GATO_ERR_UNSUPPORTED,
                 "a vector of %ld elements exceeds the reduction workspace", n);
    for (long k0 = 0; k0 < nk; k0 += per) {
      const int m = (int)(nk - k0 < per ? nk - k0 : per);
      const int st = gemvt_ept<T, VEC>(n, m, (const T*)A + k0 * lda, lda, B, C, S, s_coef, scale,
                                       out + k0, ws, s, nullptr, cap);
      if (st) return st;
    }
    return GATO_OK;
  }
  if (waves(8) >= 1024 && fits(8))
    return launch_gemvt<T, VEC, 8>(n, nk, A, lda, B, C, S, s_coef, scale, out, ws, s, np_out, cap);
  if (waves(4) >= 1024 && fits(4))
    return launch_gemvt<T, VEC, 4>(n, nk, A, lda, B, C, S, s_coef, scale, out, ws, s, np_out, cap);
  if (fits(2))
    return launch_gemvt<T, VEC, 2>(n, nk, A, lda, B, C, S, s_coef, scale, out, ws, s, np_out, cap);
  if (fits(4))
    return launch_gemvt<T, VEC, 4>(n, nk, A, lda, B, C, S, s_coef, scale, out, ws, s, np_out, cap);
  return launch_gemvt<T, VEC, 8>(n, nk, A, lda, B, C, S, s_coef, scale, out, ws, s, np_out, cap);
}

int dot_batched(int dtype, long n, int nk, const void* A, long lda, const void* B, long ldb,
                const void* C, const void* S, const double* s_coef, double scale, double* out,
                void* ws, hipStream_t s, long cap) {
  GATO_REQUIRE(n >= 1 && nk >= 1, GATO_ERR_VALUE, "bad sizes n=%ld nk=%d", n, nk);
  GATO_REQUIRE(A && B && out && ws, GATO_ERR_VALUE, "null pointer");
  GATO_REQUIRE(S == nullptr || (s_coef != nullptr && ldb == 0), GATO_ERR_VALUE,
               "S needs s_coef and a broadcast B (ldb == 0)");
  if (ldb != 0) {
    int nblk = grid_blocks(n);
    while (nblk > 1 && (long)nblk * nk > kPartialsCap) nblk /= 2;
    GATO_REQUIRE((long)nblk * nk <= cap, GATO_ERR_UNSUPPORTED,
                 "%d pairs of %ld elements need %ld partials, more than the %ld of the reduction "
                 "workspace", nk, n, (long)nblk * nk, cap);
    dim3 grid(nblk, nk);
    if (dtype == GATO_F32)
      hipLaunchKernelGGL((dot_pairs_kernel<float>), grid, dim3(kThr), 0, s, n, nk, (const float*)A,
                         lda, (const float*)B, ldb, (const float*)C, ws_partials(ws));
    else
      hipLaunchKernelGGL((dot_pairs_kernel<double>), grid, dim3(kThr), 0, s, n, nk,
                         (const double*)A, lda, (const double*)B, ldb, (const double*)C,
                         ws_partials(ws));
    GATO_LAUNCH_CHECK();
    return finalize_partials(ws_partials(ws), nblk, nk, scale, false, out, s);
  }
  const int vec = dtype == GATO_F32 ? 4 : 2;
  const bool vok = (n % vec == 0) && (nk == 1 || lda % vec == 0) && aligned16(A) && aligned16(B) &&
                   aligned16(C) && aligned16(S);
  if (dtype == GATO_F32)
    return vok ? gemvt_ept<float, 4>(n, nk, A, lda, B, C, S, s_coef, scale, out, ws, s, nullptr, cap)
               : gemvt_ept<float, 1>(n, nk, A, lda, B, C, S, s_coef, scale, out, ws, s, nullptr, cap);
  return vok ? gemvt_ept<double, 2>(n, nk, A, lda, B, C, S, s_coef, scale, out, ws, s, nullptr, cap)
             : gemvt_ept<double, 1>(n, nk, A, lda, B, C, S, s_coef, scale, out, ws, s, nullptr, cap);
}

// Enough vector groups for about 1024 waves in all, each group at least 8
// vectors long.  One slab of the basis, vectors [k0, k1): partials of the
// slab, then its finalize into coef[k0, k1).
template <typename T, int VEC, int EPT>
void proj2_slab(long n, int k0, int k1, long nblk, const T* V, long ldv, const T* vj, const T* w,
                double scale, double* coef, double* alpha, int alpha_given, void* ws,
                hipStream_t s, double probe_bytes) {
  const int nk = k1 - k0;
  const long nw = nblk * 4;
  long groups = (1024 + nw - 1) / nw;
  if (groups > nk / 8) groups = nk / 8;
  if (groups < 1) groups = 1;
  const int kper = (int)((nk + groups - 1) / groups);
  groups = (nk + kper - 1) / kper;
  if (probe_bytes > 0) probe_begin(s);
  hipLaunchKernelGGL((proj2_kernel<T, VEC, EPT>), dim3((unsigned)nblk, (unsigned)groups),
                     dim3(kThr), 0, s, n, nk, kper, V + (long)k0 * ldv, ldv, vj, w,
                     ws_partials(ws));
  if (probe_bytes > 0) probe_end(s, probe_bytes);
  hipLaunchKernelGGL(proj2_finalize_kernel, dim3(nk), dim3(kFinThr), 0, s, ws_partials(ws), nw,
                     nk, scale, coef + k0, alpha, alpha_given);
}

// The whole basis in one slab when its 2*nk wave partials fit the workspace.
// Otherwise slabs of as many vectors as fit, the slab holding v_j first: its
// finalize writes alpha, which the later slabs' finalizes read.  (The slab
// boundaries change which vectors the kernel pairs, so coefficients may then
// differ from a one-slab pass in the last bits.)
template <typename T, int VEC, int EPT>
int launch_proj2(long n, int nk, const void* V, long ldv, const void* w, double scale, double* coef,
                 double* alpha, void* ws, hipStream_t s) {
  constexpr long CH = 64L * VEC * EPT;
  const long nblk = ((n + CH - 1) / CH + 3) / 4;
  const long nw = nblk * 4;
  GATO_REQUIRE(2 * nw <= kPartialsCap && nblk <= 0x7fffffffL, GATO_ERR_UNSUPPORTED,
               "a vector of %ld elements exceeds the reduction workspace", n);
  const int slab = (int)(kPartialsCap / (2 * nw) < nk ? kPartialsCap / (2 * nw) : nk);
  const T* Vt = (const T*)V;
  const T* vj = Vt + (long)(nk - 1) * ldv;
  // the probe brackets the projection kernel of a one-slab pass, with its
  // algorithmic bytes: nk basis vectors (v_j among them) + w, each read once
  const bool probed = nk > 2 && slab == nk && probe_on(PROBE_CGS_DOT, (int)sizeof(T));
  proj2_slab<T, VEC, EPT>(n, nk - slab, nk, nblk, Vt, ldv, vj, (const T*)w, scale, coef, alpha, 0,
                          ws, s, probed ? (double)n * sizeof(T) * (nk + 1) : 0.0);
  for (int k1 = nk - slab; k1 > 0; k1 -= slab) {
    const int k0 = k1 - slab > 0 ? k1 - slab : 0;
    proj2_slab<T, VEC, EPT>(n, k0, k1, nblk, Vt, ldv, vj, (const T*)w, scale, coef, alpha, 1, ws,
                            s, 0.0);
  }
  GATO_LAUNCH_CHECK();
  return GATO_OK;
}

// Chunks of 8 vector loads per lane whenever they give at least 64 waves
// (launch_proj2 then splits the basis among more waves); single loads only on
// tiny fields.  Measured on float fields of 64^3 - 96^3 with 40-160 vectors:
// single-load chunks of all the elements ran at 2.0-3.2 TB/s, 8-load chunks
// at 5.7-5.9 TB/s on 128^3.
template <typename T, int VEC>
int proj2_ept(long n, int nk, const void* V, long ldv, const void* w, double scale, double* coef,
              double* alpha, void* ws, hipStream_t s) {
  if ((n + 64L * VEC * 8 - 1) / (64L * VEC * 8) >= 64)
    return launch_proj2<T, VEC, 8>(n, nk, V, ldv, w, scale, coef, alpha, ws, s);
  return launch_proj2<T, VEC, 1>(n, nk, V, ldv, w, scale, coef, alpha, ws, s);
}

// np_out == nullptr: finalize beta here; otherwise leave the wave partials in
// the workspace and return their count (scale_by_partials consumes them).
template <typename T, int VEC, int EPT>
int launch_cgs_ept(long n, int nk, const void* V, long ldv, const double* coef, const void* S,
                   const double* s_coef, const void* w, void* w_out, double scale, double* beta,
                   void* ws, hipStream_t s, long* np_out) {
  constexpr long CH = 64L * VEC * EPT;
  const long nb = ((n + CH - 1) / CH + 3) / 4;
  GATO_REQUIRE(nb * 4 <= kPartialsCap && nb <= 0x7fffffffL, GATO_ERR_UNSUPPORTED,
               "vector of %ld elements exceeds the reduction workspace", n);
  const size_t sh = (size_t)(nk > 0 ? nk : 1) * sizeof(double);
  const bool probed = probe_on(PROBE_CGS_UPDATE, (int)sizeof(T));
  if (probed) probe_begin(s);
  hipLaunchKernelGGL((cgs_update_kernel<T, VEC, EPT>), dim3((unsigned)nb), dim3(kThr), sh, s, n,
                     nk, (const T*)V, ldv, coef, (const T*)S, s_coef, (const T*)w, (T*)w_out,
                     ws_partials(ws));
  GATO_LAUNCH_CHECK();
  // algorithmic bytes: nk basis vectors, w (and S) read, w_out written
  if (probed) probe_end(s, (double)n * sizeof(T) * (nk + 2 + (S ? 1 : 0)));
  if (np_out) {
    *np_out = nb * (kThr / 64);
    return GATO_OK;
  }
  return finalize_partials(ws_partials(ws), nb * (kThr / 64), 1, scale, true, beta, s);
}

// Chunks of 4 vectors per lane once that still gives >= 1024 waves (8 per
// lane, 1024 waves at 128^3, measured 145 against 119 us).
template <typename T, int VEC>
int launch_cgs(long n, int nk, const void* V, long ldv, const double* coef, const void* S,
               const double* s_coef, const void* w, void* w_out, double scale, double* beta,
               void* ws, hipStream_t s, long* np_out) {
  if ((n + 64L * VEC * 4 - 1) / (64L * VEC * 4) >= 1024)
    return launch_cgs_ept<T, VEC, 4>(n, nk, V, ldv, coef, S, s_coef, w, w_out, scale, beta, ws, s,
                                     np_out);
  return launch_cgs_ept<T, VEC, 1>(n, nk, V, ldv, coef, S, s_coef, w, w_out, scale, beta, ws, s,
                                   np_out);
}

int cgs_coefficients(int dtype, long n, int nk, const void* V, long ldv, const void* w,
                     double scale, double* coef, double* alpha, void* ws, hipStream_t s) {
  GATO_REQUIRE(n >= 1 && nk >= 1, GATO_ERR_VALUE, "bad sizes n=%ld nk=%d", n, nk);
  GATO_REQUIRE(V && w && coef && alpha && ws, GATO_ERR_VALUE, "null pointer");
  const int vec = dtype == GATO_F32 ? 4 : 2;
  const bool vok = (n % vec == 0) && (nk == 1 || ldv % vec == 0) && aligned16(V) && aligned16(w);
  if (dtype == GATO_F32)
    return vok ? proj2_ept<float, 4>(n, nk, V, ldv, w, scale, coef, alpha, ws, s)
               : proj2_ept<float, 1>(n, nk, V, ldv, w, scale, coef, alpha, ws, s);
  return vok ? proj2_ept<double, 2>(n, nk, V, ldv, w, scale, coef, alpha, ws, s)
             : proj2_ept<double, 1>(n, nk, V, ldv, w, scale, coef, alpha, ws, s);
}

namespace {
int cgs_update_impl(int dtype, long n, int nk, const void* V, long ldv, const double* coef,
                    const void* S, const double* s_coef, const void* w, void* w_out, double scale,
                    double* beta, void* ws, hipStream_t s, long* np_out) {
  GATO_REQUIRE(n >= 1 && nk >= 0 && nk <= 8192, GATO_ERR_VALUE, "bad sizes n=%ld nk=%d", n, nk);
  GATO_REQUIRE(S == nullptr || s_coef != nullptr, GATO_ERR_VALUE, "S needs s_coef");
  GATO_REQUIRE(w && w_out && (beta || np_out) && ws && (nk == 0 || (V && coef)), GATO_ERR_VALUE,
               "null pointer");
  const int vec = dtype == GATO_F32 ? 4 : 2;
  const bool vok = (n % vec == 0) && (nk <= 1 || ldv % vec == 0) && aligned16(V) &&
                   aligned16(S) && aligned16(w) && aligned16(w_out);
  if (dtype == GATO_F32)
    return vok ? launch_cgs<float, 4>(n, nk, V, ldv, coef, S, s_coef, w, w_out, scale, beta, ws, s,
                                      np_out)
               : launch_cgs<float, 1>(n, nk, V, ldv, coef, S, s_coef, w, w_out, scale, beta, ws, s,
                                      np_out);
  return vok ? launch_cgs<double, 2>(n, nk, V, ldv, coef, S, s_coef, w, w_out, scale, beta, ws, s,
                                     np_out)
             : launch_cgs<double, 1>(n, nk, V, ldv, coef, S, s_coef, w, w_out, scale, beta, ws, s,
                                     np_out);
}
}  // namespace

int cgs_update(int dtype, long n, int nk, const void* V, long ldv, const double* coef,
               const void* S, const double* s_coef, const void* w, void* w_out, double scale,
               double* beta, void* ws, hipStream_t s) {
  return cgs_update_impl(dtype, n, nk, V, ldv, coef, S, s_coef, w, w_out, scale, beta, ws, s,
                         nullptr);
}

int cgs_update_partials(int dtype, long n, int nk, const void* V, long ldv, const double* coef,
                        const void* S, const double* s_coef, const void* w, void* w_out,
                        long* np, void* ws, hipStream_t s) {
  GATO_REQUIRE(np != nullptr, GATO_ERR_VALUE, "null np");
  return cgs_update_impl(dtype, n, nk, V, ldv, coef, S, s_coef, w, w_out, 0.0, nullptr, ws, s, np);
}

int scale_by_partials_at(int dtype, long n, const void* w, const double* part, long np,
                         double scale, double* beta, void* v, hipStream_t s);

int scale_by_partials(int dtype, long n, const void* w, const void* ws, long np, double scale,
                      double* beta, void* v, hipStream_t s) {
  GATO_REQUIRE(ws != nullptr, GATO_ERR_VALUE, "bad arguments");
  return scale_by_partials_at(dtype, n, w, ws_partials(const_cast<void*>(ws)), np, scale, beta, v,
                              s);
}

int scale_by_partials_at(int dtype, long n, const void* w, const double* part, long np,
                         double scale, double* beta, void* v, hipStream_t s) {
  GATO_REQUIRE(n >= 1 && np >= 1 && np <= kPartialsCap && w && part && beta && v, GATO_ERR_VALUE,
               "bad arguments");
  const int vec = dtype == GATO_F32 ? 4 : 2;
  const bool vok = (n % vec == 0) && aligned16(w) && aligned16(v);
  // a few vectors per thread keep the per-block partial sums (np doubles
  // from L2 each) small next to the field traffic
  auto blocks = [&](long per_thread) {
    long b = (n + kThr * per_thread - 1) / (kThr * per_thread);
    return (int)(b < 1 ? 1 : b > kReduceBlocks ? kReduceBlocks : b);
  };
  if (dtype == GATO_F32) {
    if (vok)
      hipLaunchKernelGGL((scale_inv_partials_kernel<float, 4>), dim3(blocks(16)), dim3(kThr), 0, s,
                         n, (const float*)w, part, np, scale, beta, (float*)v);
    else
      hipLaunchKernelGGL((scale_inv_partials_kernel<float, 1>), dim3(blocks(16)), dim3(kThr), 0, s,
                         n, (const float*)w, part, np, scale, beta, (float*)v);
  } else {
    if (vok)
      hipLaunchKernelGGL((scale_inv_partials_kernel<double, 2>), dim3(blocks(8)), dim3(kThr), 0, s,
                         n, (const double*)w, part, np, scale, beta, (double*)v);
    else
      hipLaunchKernelGGL((scale_inv_partials_kernel<double, 1>), dim3(blocks(8)), dim3(kThr), 0, s,
                         n, (const double*)w, part, np, scale, beta, (double*)v);
  }
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

int normalize(int dtype, long n, const void* w, double scale, double* norm_out, void* v, void* ws,
              hipStream_t s) {
  GATO_REQUIRE(n >= 1 && w && v && ws, GATO_ERR_VALUE, "bad arguments");
  const int vec = dtype == GATO_F32 ? 4 : 2;
  const bool vok = (n % vec == 0) && aligned16(w);
  long np = 0;
  int st;
  if (dtype == GATO_F32)
    st = vok ? gemvt_ept<float, 4>(n, 1, w, 0, w, nullptr, nullptr, nullptr, 1.0, nullptr, ws, s, &np)
             : gemvt_ept<float, 1>(n, 1, w, 0, w, nullptr, nullptr, nullptr, 1.0, nullptr, ws, s, &np);
  else
    st = vok ? gemvt_ept<double, 2>(n, 1, w, 0, w, nullptr, nullptr, nullptr, 1.0, nullptr, ws, s, &np)
             : gemvt_ept<double, 1>(n, 1, w, 0, w, nullptr, nullptr, nullptr, 1.0, nullptr, ws, s, &np);
  if (st) return st;
  return scale_by_partials(dtype, n, w, ws, np, scale, norm_out ? norm_out : ws_stage(ws), v, s);
}

int ritz_combine(int dtype, long n, int m, int K, const void* V, long ldv, const double* S,
                 void* out, long ldo, hipStream_t s) {
  GATO_REQUIRE(n >= 1 && m >= 1 && K >= 1 && K <= kRitzMax, GATO_ERR_VALUE,
               "bad sizes n=%ld m=%d K=%d (K <= %d)", n, m, K, kRitzMax);
  GATO_REQUIRE((size_t)m * K * sizeof(double) <= 60 * 1024, GATO_ERR_UNSUPPORTED,
               "Ritz coefficient block m*K=%d too large", m * K);
  const size_t sh = (size_t)m * K * sizeof(double);
  const int vec = dtype == GATO_F32 ? 4 : 2;
  if (K <= 4 && n % vec == 0 && ldv % vec == 0 && ldo % vec == 0 && aligned16(V) &&
      aligned16(out)) {
    const long nv = n / vec;  // one VEC-vector per thread, grid-stride beyond 65535 blocks
    const int nb = (int)((nv + kThr - 1) / kThr < 65535 ? (nv + kThr - 1) / kThr : 65535);
    if (dtype == GATO_F32)
      hipLaunchKernelGGL((ritz4_kernel<float, 4>), dim3(nb), dim3(kThr), sh, s, n, m, K,
                         (const float*)V, ldv, S, (float*)out, ldo);
    else
      hipLaunchKernelGGL((ritz4_kernel<double, 2>), dim3(nb), dim3(kThr), sh, s, n, m, K,
                         (const double*)V, ldv, S, (double*)out, ldo);
    GATO_LAUNCH_CHECK();
    return GATO_OK;
  }
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
                     as_stream(stream), kPartialsCap);
}

// ---------------------------------------------------------------------------
// gato_dot_batched without a workspace argument, for callers that cannot
// allocate device memory.  Its reduction partials live in ``scratch``, a
// file-scope device array: every GPU the library runs on holds one copy of
// it (2 MiB + 256 bytes) until the process ends.  The host never passes the
// array's name to a kernel (in host code it is only a stand-in symbol); the
// device address of each GPU's copy is looked up once per device, with that
// device current, and kept in g_scratch_addr.  Every call on one GPU uses
// that GPU's copy: calls that overlap on one GPU (two streams, two host
// threads) race on it.  The reduction kernels store every partial before the
// finalize kernel reads it, so nothing has to be cleared between calls.
// ---------------------------------------------------------------------------
namespace {
constexpr long kScratchPartials = 1L << 18;  // doubles of partials (2 MiB)
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxScratchDevices] = {};
}  // namespace

static __device__ __align__(256) double scratch[kPartialsOffset / sizeof(double) + kScratchPartials];

// The current device's copy of ``scratch``, looked up on the first call on
// that device.
static int scratch_address(void** out) {
  int dev = -1;
  GATO_HIP_CHECK(hipGetDevice(&dev));
  GATO_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, GATO_ERR_UNSUPPORTED,
               "device %d is beyond the %d devices the static scratch supports", dev,
               kMaxScratchDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_addr[dev] == nullptr) {
    void* p = nullptr;
    GATO_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    GATO_REQUIRE(p != nullptr, GATO_ERR_HIP, "no device address for the static scratch");
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return GATO_OK;
}

extern "C" int gato_dot_batched_noalloc(int dtype, long n, int nk, const void* A, long lda,
                                        const void* B, long ldb, const void* C, const void* S,
                                        const double* s_coef, double scale, double* out,
                                        void* stream) {
  clear_ok();
  void* ws = nullptr;
  int st = scratch_address(&ws);
  if (st) return st;
  return dot_batched(dtype, n, nk, A, lda, B, ldb, C, S, s_coef, scale, out, ws,
                     as_stream(stream), kScratchPartials);
}

extern "C" int gato_cgs_update(int dtype, long n, int nk, const void* V, long ldv,
                               const double* coef, const void* S, const double* s_coef,
                               const void* w, void* w_out, double scale, double* beta_out,
                               void* ws, void* stream) {
  clear_ok();
  return cgs_update(dtype, n, nk, V, ldv, coef, S, s_coef, w, w_out, scale, beta_out, ws,
                    as_stream(stream));
}

extern "C" int gato_cgs_coefficients(int dtype, long n, int nk, const void* V, long ldv,
                                     const void* w, double scale, double* coef, double* alpha,
                                     void* ws, void* stream) {
  clear_ok();
  return cgs_coefficients(dtype, n, nk, V, ldv, w, scale, coef, alpha, ws, as_stream(stream));
}

extern "C" int gato_scale_by_inverse(int dtype, long n, const void* w, const double* beta, void* v,
                                     void* stream) {
  clear_ok();
  return scale_by_inverse(dtype, n, w, beta, v, as_stream(stream));
}

extern "C" int gato_normalize(int dtype, long n, const void* w, double scale, double* norm_out,
                              void* v, void* ws, void* stream) {
  clear_ok();
  return normalize(dtype, n, w, scale, norm_out, v, ws, as_stream(stream));
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
