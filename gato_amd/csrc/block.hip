// Block Lanczos with full reorthogonalisation: kernels, CholQR2 block
// orthonormalisation, the native loop over a Fock operator, and the host
// dense symmetric eigensolver for the block-tridiagonal matrix.
//
// The reference has no block eigensolver; solvers/lanczos.py:83-107 names it
// as the fix for degenerate manifolds (a single start vector spans one
// direction per invariant subspace, so an exactly degenerate level is only
// split by round-off), README.md:431 lists it as open work.  This solver
// keeps the reference Lanczos structure (lanczos.py:140-194) with vectors
// replaced by blocks of b vectors:
//   W   = H Q_i - Q_{i-1} B_{i-1}^T           (lanczos.py:144-148)
//   A_i = Q_i^T W                              (lanczos.py:150)
//   W  -= Q_i A_i + V C,  C = V^T (W - Q_i A_i) (full reorthogonalisation,
//                                               lanczos.py:151-159)
//   W   = Q_{i+1} B_i                          (CholQR2; lanczos.py:163-167)
// and T block tridiagonal with diagonal blocks A_i and off-diagonal B_i.
// As in the scalar path, one projection pass reads the basis once for both
// D = V^T W and G = V^T Q_i, and C = D - G A_i folds the A_i term in exactly
// (the computed G carries the basis' loss of orthogonality), so one sweep
// over V does the update.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

struct gato_fock;

namespace gato {

int fock_apply(gato_fock* f, const void* psi, void* out, const void* sub, const double* sub_coef,
               void* ws, hipStream_t s);
int fock_apply_block(gato_fock* f, int nb, const void* psi, void* out, void* ws, hipStream_t s,
                     bool with_ace);
int fock_ace(const gato_fock* f, const void** xi);
int fock_dtype(const gato_fock* f);
long fock_n(const gato_fock* f);
int ritz_combine(int dtype, long n, int m, int K, const void* V, long ldv, const double* S,
                 void* out, long ldo, hipStream_t s);
void tridiag_inverse_iteration(int m, const double* a, const double* b, int K, const double* lam,
                               double* Z);
void tridiag_lowest_eigvals(int m, const double* a, const double* b, int K, double* out);

namespace {

constexpr int kThr = 256;
constexpr int kMaxBlock = 8;  // b <= 8 vectors per block

// The rows of a projection or an update: rows r < np from P (stride ldp),
// the others from V (stride ldv).  The block Lanczos step puts the fields of
// the compressed exchange ahead of the Krylov basis this way, so one pass
// over both applies the exchange and reorthogonalises.
template <typename T>
struct Rows {
  const T* P;
  long ldp;
  int np;
  const T* V;
  long ldv;
  __host__ __device__ __forceinline__ const T* operator()(int r) const {
    return r < np ? P + (long)r * ldp : V + (long)(r - np) * ldv;
  }
  // the rows from r0 on
  Rows shift(int r0) const {
    if (r0 < np) return {P + (long)r0 * ldp, ldp, np - r0, V, ldv};
    return {nullptr, 0, 0, V + (long)(r0 - np) * ldv, ldv};
  }
};

// ---------------------------------------------------------------------------
// C = V^T X:  for k < nk, c < nb:  sum_i V_k[i] X_c[i].
// Each 64-lane wave owns chunks of 64*VEC*EPT elements (NB*EPT = 16 keeps the
// NB broadcast columns of a chunk in registers), streams the basis vectors
// through them two at a time, and reduces the 2*NB per-lane sums with one
// butterfly reduce-scatter.  Every wave accumulates into its own LDS slice,
// so the loop has no block barrier; the slices are summed in a fixed order
// at the end and stored as block partials part[slot * gridDim.x + block],
// slot = k * nb + c.
// ---------------------------------------------------------------------------
template <typename T, int VEC, int NB>
__global__ __launch_bounds__(kThr) void block_dot_kernel(long n, int nk, Rows<T> V,
                                                         const T* __restrict__ X, long ldx, int nb,
                                                         double* __restrict__ part) {
  constexpr int EPT = 16 / NB;
  constexpr int NV = 2 * NB;
  using RV = typename RawVec<T, VEC>::type;
  extern __shared__ double acc[];  // [4 waves][nk * nb]
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int slots = nk * nb;
  for (int i = threadIdx.x; i < 4 * slots; i += kThr) acc[i] = 0.0;
  __syncthreads();
  double* mine = acc + wid * slots;
  constexpr long CH = 64L * VEC * EPT;
  const long nchunk = (n + CH - 1) / CH;
  const int idx = wave_scatter_index<NV>(lane);
  const bool writer = (lane & (64 / NV - 1)) == 0;
  const int kk = idx / NB, cc = idx % NB;
  for (long g = blockIdx.x; g * 4 < nchunk; g += gridDim.x) {
    const long base = (g * 4 + wid) * CH;
    double x[NB][EPT][VEC];
#pragma unroll
    for (int c = 0; c < NB; ++c)
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const long i = base + ((long)e * 64 + lane) * VEC;
        if (c < nb && i < n) {
          vload<T, VEC>(X + (long)c * ldx, i, x[c][e]);
        } else {
#pragma unroll
          for (int u = 0; u < VEC; ++u) x[c][e][u] = 0.0;
        }
      }
    for (int k = 0; k < nk; k += 2) {
      const bool two = k + 1 < nk;
      double a[2][EPT][VEC];
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const long i = base + ((long)e * 64 + lane) * VEC;
        if (i < n) {
          vload<T, VEC>(V(k), i, a[0][e]);
          if (two) {
            vload<T, VEC>(V(k + 1), i, a[1][e]);
          } else {
#pragma unroll
            for (int u = 0; u < VEC; ++u) a[1][e][u] = 0.0;
          }
        } else {
#pragma unroll
          for (int u = 0; u < VEC; ++u) a[0][e][u] = a[1][e][u] = 0.0;
        }
      }
      double v[NV];
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          double t = 0.0;
#pragma unroll
          for (int e = 0; e < EPT; ++e)
#pragma unroll
            for (int u = 0; u < VEC; ++u) t += a[q][e][u] * x[c][e][u];
          v[q * NB + c] = t;
        }
      wave_reduce_scatter<NV>(v, lane);
      if (writer && cc < nb && k + kk < nk) mine[(k + kk) * nb + cc] += v[0];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < slots; i += kThr)
    part[(long)i * gridDim.x + blockIdx.x] =
        ((acc[i] + acc[slots + i]) + acc[2 * slots + i]) + acc[3 * slots + i];
}

typedef double d4v __attribute__((ext_vector_type(4)));

// Eight consecutive elements held as raw 16-byte vectors, widened to double.
template <typename T>
__device__ __forceinline__ void widen8(const typename RawVec<T, 16 / sizeof(T)>::type* r,
                                       double (&o)[8]) {
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      o[4 * q] = r[q].x; o[4 * q + 1] = r[q].y; o[4 * q + 2] = r[q].z; o[4 * q + 3] = r[q].w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      o[2 * q] = r[q].x; o[2 * q + 1] = r[q].y;
    }
  }
}

// The same product C = V^T X on the matrix cores (n % 64 == 0, 16-byte
// aligned rows, nb <= 16, nk <= 16*NKT).  v_mfma_f64_16x16x4f64 takes 16
// basis vectors as the rows of A and the broadcast columns as the columns of
// B; its reduction index runs over the four 16-lane groups.  Lane l
// (r = l % 16, g = l / 16) covers 16 consecutive elements at 64 rg + 16 g of
// vector 16t + r and of column r, in two halves of 8: it widens them to
// double (products of floats are then exact and the sums double, as in the
// vector-ALU kernel) and issues one MFMA per element.  Each 16-vector tile
// keeps its 16 x 16 accumulator in registers over the wave's whole element
// range.  Accumulator layout of this instruction on gfx950 (measured): lane
// l, register q holds D[l/16 + 4q][l % 16].  The four waves of a block add
// their accumulators through LDS in a fixed order; block partials go to
// part[(k * nb + c) * gridDim.x + block].
template <typename T, int NKT>
__global__ __launch_bounds__(kThr) void block_dot_mfma_kernel(long n, int nk, Rows<T> V,
                                                              const T* __restrict__ X, long ldx,
                                                              int nb, double* __restrict__ part) {
  extern __shared__ double red[];  // [4 waves][nk * nb]
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int r16 = lane & 15, g = lane >> 4;
  const long nw = (long)gridDim.x * (kThr / 64);
  const long wave = (long)blockIdx.x * (kThr / 64) + wid;
  const int nkt = (nk + 15) >> 4;
  const int slots = nk * nb;
  const bool colok = r16 < nb;
  d4v acc[NKT];
#pragma unroll
  for (int t = 0; t < NKT; ++t) acc[t] = (d4v){0.0, 0.0, 0.0, 0.0};
  // The wave's work is a sequence of half ranges (8 elements per lane).  Two
  // register buffers alternate: the 16-byte loads of the next half range are
  // issued before the MFMAs of the current one, so the stream from HBM and
  // the matrix cores overlap (NKT * 32 bytes per lane and buffer).
  using RV = typename RawVec<T, 16 / sizeof(T)>::type;
  constexpr int NQ = 8 * sizeof(T) / 16;  // raw vectors per 8 elements
  const long nr = n >> 6;
  const long nj = wave < nr ? 2 * ((nr - wave + nw - 1) / nw) : 0;
  // Element order: the K position a lane group g takes in an MFMA only has to
  // agree between V and X, so piece q of half range j is elements
  // 32 (j & 1) + 4 E q + E g (E elements per 16-byte load): the 4 lanes that
  // share a vector read 64 contiguous bytes per load instead of four 16-byte
  // pieces 64 bytes apart.
  constexpr int E = 16 / sizeof(T);
  auto issue = [&](long j, RV(&xr)[NQ], RV(&vr)[NKT][NQ]) {
    const long base = (wave + (j >> 1) * nw) * 64 + (j & 1) * 32 + E * g;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      xr[q] = colok ? *reinterpret_cast<const RV*>(X + (long)r16 * ldx + base + 4 * E * q) : RV{};
#pragma unroll
    for (int t = 0; t < NKT; ++t) {
      const int k = t * 16 + r16;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        vr[t][q] = (t < nkt && k < nk)
                       ? *reinterpret_cast<const RV*>(V(k) + base + 4 * E * q)
                       : RV{};
    }
  };
  auto consume = [&](const RV(&xr)[NQ], const RV(&vr)[NKT][NQ]) {
    double xv[8];
    widen8<T>(xr, xv);
#pragma unroll
    for (int t = 0; t < NKT; ++t) {
      if (t < nkt) {
        double vv[8];
        widen8<T>(vr[t], vv);
#pragma unroll
        for (int e = 0; e < 8; ++e)
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(vv[e], xv[e], acc[t], 0, 0, 0);
      }
    }
  };
  RV xa[NQ], va[NKT][NQ], xb[NQ], vb[NKT][NQ];
  if (nj > 0) issue(0, xa, va);
  for (long j = 0; j < nj; j += 2) {  // nj is even: two halves per range
    issue(j + 1, xb, vb);
    consume(xa, va);
    if (j + 2 < nj) issue(j + 2, xa, va);
    consume(xb, vb);
  }
  double* mine = red + wid * slots;
  if (colok) {
#pragma unroll
    for (int t = 0; t < NKT; ++t) {
      if (t < nkt) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int k = t * 16 + g + 4 * q;
          if (k < nk) mine[k * nb + r16] = acc[t][q];
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < slots; i += kThr)
    part[(long)i * gridDim.x + blockIdx.x] =
        ((red[i] + red[slots + i]) + red[2 * slots + i]) + red[3 * slots + i];
}

// The float projection with staged loads.  The operand layout above makes
// every load instruction touch 16 vectors (64 bytes of each), and that
// pattern streams at about 3.2 TB/s on its own; loads of whole 128-byte lines
// of one vector per 8 lanes stream at the copy rate.  So each wave loads a
// 32-element half of every row (nk vectors, then nb columns at row 64) with
// 8 lanes per row, stages it in its own LDS tile (pitch 36 floats: the
// 16-byte reads of 16 rows hit distinct banks), and reads the operands from
// there: lane (r, g) takes elements 8g .. 8g+7 of vector 16t + r and of
// column r, and MFMA k of the half consumes element 8g + k (the K position
// only has to agree between the operands).  The next half's loads are in
// flight during the MFMAs of the current one.  With at most 8 columns (a
// block of b = 4 projects on 2b = 8) the 4 x 4 x 4 four-block shape covers 8
// vectors x 8 columns with no idle products.  Measured at 128^3, 8 columns:
// 64 / 128 / 160 vectors 121.9 / 251.8 / 314.2 us (4.9-5.0 TB/s) against
// 156.2 / 318.4 / 387.3 us for the unstaged kernel; the 16 x 16 x 4 shape
// staged 124.2 / 269.4 / 324.2 us.  The probe that found the load pattern: a
// kernel with the same loads and no MFMAs ran 190.7 us for 64 vectors + 8
// columns (3.2 TB/s), one reading 1 KB of one vector per instruction 94.5 us.
constexpr int kStageRows = 80;   // 64 vectors + 16 columns
constexpr int kStagePitch = 36;  // floats per staged row
// The 4 x 4 x 4 path stages 64 vectors + 8 columns at pitch 32 with the
// 16-byte piece p of row r at position p ^ (r & 7): conflict-free for its
// writes and reads, and 36 KB per block (4 blocks per CU instead of 3).
constexpr int kStageRows4 = 72;
constexpr int kStagePitch4 = 32;

template <bool F4>
__device__ __forceinline__ int stage_at(int row, int piece) {
  if constexpr (F4) return row * kStagePitch4 + 4 * (piece ^ (row & 7));
  else return row * kStagePitch + 4 * piece;
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NKT, bool F4, int SR = kStageRows4>
__global__ __launch_bounds__(kThr) void block_dot_staged_kernel(long n, int nk, Rows<float> V,
                                                                const float* __restrict__ X,
                                                                long ldx, int nb,
                                                                double* __restrict__ part) {
  extern __shared__ double red[];  // staging tiles during the loop, then [4 waves][nk * nb]
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int r16 = lane & 15, g = lane >> 4;
  const long nw = (long)gridDim.x * (kThr / 64);
  const long wave = (long)blockIdx.x * (kThr / 64) + wid;
  const int nkt = (nk + 15) >> 4;
  const int slots = nk * nb;
  // F4: SR staged rows, the last 8 for the columns (SR = 24, 40 or 72 for up
  // to 16, 32 or 64 vectors); otherwise 64 vectors + 16 columns
  constexpr int SROWS = F4 ? SR : kStageRows;
  constexpr int VR = F4 ? SR - 8 : 64;  // staged vector rows
  static_assert(!F4 || 2 * NKT * 8 >= VR, "tiles cover the vector rows");
  float* stage =
      reinterpret_cast<float*>(red) + wid * SROWS * (F4 ? kStagePitch4 : kStagePitch);
  d4v acc[NKT];
#pragma unroll
  for (int t = 0; t < NKT; ++t) acc[t] = (d4v){0.0, 0.0, 0.0, 0.0};
  // F4: v_mfma_f64_4x4x4f64 (4 blocks of 4 x 4) on tiles of 8 vectors x 8
  // columns, block b = (vectors 4 (b >> 1) .., columns 4 (b & 1) ..); lane l
  // supplies A(b, i, k) with i = l % 4, b = (l / 4) % 4, k = l / 16, B(b, k, j)
  // with j = l % 4, and holds D(b, i = l / 16, j = l % 4) (measured layout)
  double acc8[2 * NKT];
#pragma unroll
  for (int t = 0; t < 2 * NKT; ++t) acc8[t] = 0.0;
  const int lb = (lane >> 2) & 3, li = lane & 3;
  constexpr int NI = SROWS / 8;  // load instructions per half: 8 rows each
  const int lrow = lane >> 3, lpiece = 4 * (lane & 7);
  const long nh = n >> 5;
  float4 ld[NI];
  auto issue = [&](long h) {
    const long e0 = h * 32 + lpiece;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int row = 8 * i + lrow;
      const bool isv = row < VR;
      const int c = row - VR;
      const bool ok = isv ? row < nk : c < nb;
      const float* src = isv ? V(row) : X + (long)(ok ? c : 0) * ldx;
      ld[i] = ok ? *reinterpret_cast<const float4*>(src + e0) : float4{0.f, 0.f, 0.f, 0.f};
    }
  };
  long h = wave;
  if (h < nh) issue(h);
  for (; h < nh; h += nw) {
#pragma unroll
    for (int i = 0; i < NI; ++i)
      *reinterpret_cast<float4*>(stage + stage_at<F4>(8 * i + lrow, lane & 7)) = ld[i];
    wave_lds_sync();
    if (h + nw < nh) issue(h + nw);
    float b[8];
    {
      const int brow = VR + (F4 ? 4 * (lb & 1) + li : r16);
      const float4 b0 = *reinterpret_cast<const float4*>(stage + stage_at<F4>(brow, 2 * g));
      const float4 b1 = *reinterpret_cast<const float4*>(stage + stage_at<F4>(brow, 2 * g + 1));
      b[0] = b0.x; b[1] = b0.y; b[2] = b0.z; b[3] = b0.w;
      b[4] = b1.x; b[5] = b1.y; b[6] = b1.z; b[7] = b1.w;
    }
    if constexpr (F4) {
#pragma unroll
      for (int t = 0; t < 2 * NKT; ++t) {
        if (8 * t < nk) {
          const int arow = 8 * t + 4 * (lb >> 1) + li;
          const float4 a0 = *reinterpret_cast<const float4*>(stage + stage_at<F4>(arow, 2 * g));
          const float4 a1 =
              *reinterpret_cast<const float4*>(stage + stage_at<F4>(arow, 2 * g + 1));
          const float a[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
#pragma unroll
          for (int e = 0; e < 8; ++e)
            acc8[t] = __builtin_amdgcn_mfma_f64_4x4x4f64((double)a[e], (double)b[e], acc8[t], 0,
                                                         0, 0);
        }
      }
    } else {
#pragma unroll
      for (int t = 0; t < NKT; ++t) {
        if (t < nkt) {
          const int arow = 16 * t + r16;
          const float4 a0 = *reinterpret_cast<const float4*>(stage + stage_at<F4>(arow, 2 * g));
          const float4 a1 =
              *reinterpret_cast<const float4*>(stage + stage_at<F4>(arow, 2 * g + 1));
          const float a[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
#pragma unroll
          for (int e = 0; e < 8; ++e)
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[e], (double)b[e], acc[t], 0,
                                                          0, 0);
        }
      }
    }
    wave_lds_sync();  // this half's reads are done before the next half is staged
  }
  __syncthreads();  // every wave is done with its staging tile
  double* mine = red + wid * slots;
  if constexpr (F4) {
    const int c = 4 * (lb & 1) + li;
#pragma unroll
    for (int t = 0; t < 2 * NKT; ++t) {
      const int k = 8 * t + 4 * (lb >> 1) + g;
      if (k < nk && c < nb) mine[k * nb + c] = acc8[t];
    }
  } else if (r16 < nb) {
#pragma unroll
    for (int t = 0; t < NKT; ++t) {
      if (t < nkt) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int k = t * 16 + g + 4 * q;
          if (k < nk) mine[k * nb + r16] = acc[t][q];
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < slots; i += kThr)
    part[(long)i * gridDim.x + blockIdx.x] =
        ((red[i] + red[slots + i]) + red[2 * slots + i]) + red[3 * slots + i];
}

// Gram slot of (c, d), c <= d < nb, in row order of the upper triangle.
__host__ __device__ inline int gram_slot(int c, int d, int nb) {
  return c * nb - c * (c - 1) / 2 + (d - c);
}

// Wave partials of the Gram matrix of the nb columns y (upper triangle, slot
// gram_slot) for one wave's chunk: part[slot * nw + wave].
template <int NB, int EPT, int VEC, typename Y>
__device__ __forceinline__ void gram_partials(const Y (&y)[NB][EPT][VEC], int nb, int lane,
                                              long nw, long wave, double* part) {
  constexpr int NG = NB * (NB + 1) / 2;
  constexpr int NGP = NG <= 4 ? 4 : NG <= 16 ? 16 : 64;
  double v[NGP];
#pragma unroll
  for (int g = 0; g < NGP; ++g) v[g] = 0.0;
  // element by element (few values live at once), each slot summed in (e, u)
  // order
#pragma unroll
  for (int e = 0; e < EPT; ++e)
#pragma unroll
    for (int u = 0; u < VEC; ++u) {
      double yd[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) yd[c] = (double)y[c][e][u];
#pragma unroll
      for (int c = 0; c < NB; ++c)
#pragma unroll
        for (int d = c; d < NB; ++d) v[c * NB - c * (c - 1) / 2 + (d - c)] += yd[c] * yd[d];
    }
  wave_reduce_scatter<NGP>(v, lane);
  const int idx = wave_scatter_index<NGP>(lane);
  const int ng = nb * (nb + 1) / 2;
  // map the NB-packed slot to the nb-packed one
  if ((lane & (64 / NGP - 1)) == 0 && idx < NG) {
    int c = 0, rem = idx;
    while (rem >= NB - c) { rem -= NB - c; ++c; }
    const int d = c + rem;
    if (d < nb) {
      const int slot = gram_slot(c, d, nb);
      if (slot < ng) part[(long)slot * nw + wave] = v[0];
    }
  }
}

// Elements per lane of a block update / scale wave for blocks of up to NB
// columns: 2, or 1 at NB = 8 (at most 8 columns' worth of double registers),
// so a 128^3 float launch has 4096 waves (8192 at NB = 8) for every NB.
template <int NB>
constexpr int block_ept() {
  return NB >= 8 ? 1 : 2;
}

// Y_c = X_c - sum_{k<nk} V_k C[k * nb + c] (c < nb; Y may alias X), and, when
// part != nullptr, wave partials of the Gram matrix of Y as stored.  Each
// wave owns one chunk of 64*VEC*EPT elements.
template <typename T, int VEC, int NB>
__global__ __launch_bounds__(kThr) void block_update_kernel(long n, int nk, Rows<T> V,
                                                            const double* __restrict__ C,
                                                            const T* X, long ldx, int nb, T* Y,
                                                            long ldy, double* part) {
  constexpr int EPT = block_ept<NB>();
  extern __shared__ double cs[];  // nk * nb
  for (int i = threadIdx.x; i < nk * nb; i += kThr) cs[i] = C[i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long nw = (long)gridDim.x * (kThr / 64);
  const long wave = (long)blockIdx.x * (kThr / 64) + (threadIdx.x >> 6);
  constexpr long CH = 64L * VEC * EPT;
  const long base = wave * CH;
  double y[NB][EPT][VEC];
#pragma unroll
  for (int c = 0; c < NB; ++c)
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const long i = base + ((long)e * 64 + lane) * VEC;
      if (c < nb && i < n) {
        vload<T, VEC>(X + (long)c * ldx, i, y[c][e]);
      } else {
#pragma unroll
        for (int u = 0; u < VEC; ++u) y[c][e][u] = 0.0;
      }
    }
  // PD = 8 / EPT basis vectors in flight per wave (8 loads of 16 bytes per lane,
  // 8 KB per wave), kept raw and widened element by element at the FMAs; each
  // y element takes its FMAs in basis order
  using RV = typename RawVec<T, VEC>::type;
  constexpr int PD = 8 / EPT > 0 ? 8 / EPT : 1;
  RV buf[PD][EPT];
  auto fetch = [&](int k, RV (&b)[EPT]) {
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const long i = base + ((long)e * 64 + lane) * VEC;
      b[e] = (k < nk && i < n) ? rload_nt<T, VEC>(V(k), i) : rzero<T, VEC>();
    }
  };
#pragma unroll
  for (int p = 0; p < PD; ++p) fetch(p, buf[p]);
  for (int k0 = 0; k0 < nk; k0 += PD) {
#pragma unroll
    for (int p = 0; p < PD; ++p) {
      const int k = k0 + p;
      if (k < nk) {  // uniform
        double cf[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) cf[c] = c < nb ? cs[k * nb + c] : 0.0;
#pragma unroll
        for (int e = 0; e < EPT; ++e)
#pragma unroll
          for (int u = 0; u < VEC; ++u) {
            const double a = rget<T, VEC>(buf[p][e], u);
#pragma unroll
            for (int c = 0; c < NB; ++c) y[c][e][u] -= cf[c] * a;
          }
        fetch(k + PD, buf[p]);
      }
    }
  }
  // round to the storage type first: the Gram is that of the stored columns
#pragma unroll
  for (int c = 0; c < NB; ++c)
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
#pragma unroll
      for (int u = 0; u < VEC; ++u) y[c][e][u] = (double)(T)y[c][e][u];
      const long i = base + ((long)e * 64 + lane) * VEC;
      if (c < nb && i < n) vstore<T, VEC>(Y + (long)c * ldy, i, y[c][e]);
    }
  if (part) gram_partials<NB, EPT, VEC>(y, nb, lane, nw, wave, part);
}

// Y_c = sum_{d <= c} X_d Rinv[d * nb + c] (Rinv upper triangular; Y may alias
// X) and optional Gram wave partials of Y.
template <typename T, int VEC, int NB>
__global__ __launch_bounds__(kThr) void block_scale_kernel(long n, const T* X, long ldx, int nb,
                                                           const double* __restrict__ Rinv, T* Y,
                                                           long ldy, double* part) {
  constexpr int EPT = block_ept<NB>();
  __shared__ double rs[kMaxBlock * kMaxBlock];
  for (int i = threadIdx.x; i < nb * nb; i += kThr) rs[i] = Rinv[i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long nw = (long)gridDim.x * (kThr / 64);
  const long wave = (long)blockIdx.x * (kThr / 64) + (threadIdx.x >> 6);
  constexpr long CH = 64L * VEC * EPT;
  const long base = wave * CH;
  // inputs kept raw and outputs in the storage type (the Gram is that of the
  // stored columns): half the registers of double copies
  using RV = typename RawVec<T, VEC>::type;
  RV x[NB][EPT];
  T y[NB][EPT][VEC];
#pragma unroll
  for (int c = 0; c < NB; ++c)
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const long i = base + ((long)e * 64 + lane) * VEC;
      x[c][e] = (c < nb && i < n) ? rload<T, VEC>(X + (long)c * ldx, i) : rzero<T, VEC>();
    }
#pragma unroll
  for (int e = 0; e < EPT; ++e)
#pragma unroll
    for (int u = 0; u < VEC; ++u) {
      double xd[NB];
#pragma unroll
      for (int d = 0; d < NB; ++d) xd[d] = rget<T, VEC>(x[d][e], u);
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        double t = 0.0;
#pragma unroll
        for (int d = 0; d <= c; ++d) t += xd[d] * (c < nb ? rs[d * nb + c] : 0.0);
        y[c][e][u] = (T)t;
      }
    }
#pragma unroll
  for (int c = 0; c < NB; ++c)
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const long i = base + ((long)e * 64 + lane) * VEC;
      if (c < nb && i < n) {
        double yd[VEC];
#pragma unroll
        for (int u = 0; u < VEC; ++u) yd[u] = (double)y[c][e][u];
        vstore<T, VEC>(Y + (long)c * ldy, i, yd);
      }
    }
  if (part) gram_partials<NB, EPT, VEC>(y, nb, lane, nw, wave, part);
}

// Upper Cholesky factor R of the Gram matrix G (packed upper triangle, already
// scaled) with deflation: a column whose remaining pivot is not above
// rel_tol * G_cc (or is not positive) is dependent on the ones before it; its
// row of R and its column of Rinv are zero and dropped[c] is set.  With
// Rprev (the first CholQR factor) the combined factor R * Rprev is stored in
// Bout (and its transpose in BTout).  One wave: lane 0 factors (nb <= 8) in
// LDS, lanes d < nb invert column d by back substitution, and every lane
// forms entries of the products.
__global__ __launch_bounds__(64) void block_chol_kernel(int nb, const double* __restrict__ G,
                                                        double rel_tol,
                                                        const double* __restrict__ Rprev,
                                                        double* __restrict__ Rout,
                                                        double* __restrict__ Rinv,
                                                        double* __restrict__ Bout,
                                                        double* __restrict__ BTout,
                                                        int* __restrict__ dropped) {
  __shared__ double R[kMaxBlock][kMaxBlock], Ri[kMaxBlock][kMaxBlock];
  const int t = threadIdx.x;
  if (t < kMaxBlock * kMaxBlock) {
    R[t / kMaxBlock][t % kMaxBlock] = 0.0;
    Ri[t / kMaxBlock][t % kMaxBlock] = 0.0;
  }
  __syncthreads();
  if (t == 0) {
    for (int c = 0; c < nb; ++c) {
      const double gcc = G[gram_slot(c, c, nb)];
      double piv = gcc;
      for (int k = 0; k < c; ++k) piv -= R[k][c] * R[k][c];
      const bool keep = piv > rel_tol * gcc && piv > 0.0 && gcc > 0.0;
      if (!keep) {
        if (dropped) dropped[c] = 1;
        continue;
      }
      const double r = sqrt(piv);
      R[c][c] = r;
      for (int d = c + 1; d < nb; ++d) {
        double v = G[gram_slot(c, d, nb)];
        for (int k = 0; k < c; ++k) v -= R[k][c] * R[k][d];
        R[c][d] = v / r;
      }
    }
  }
  __syncthreads();
  // column d of Rinv (zero rows of R stay zero)
  if (t < nb && R[t][t] != 0.0) {
    const int c = t;
    Ri[c][c] = 1.0 / R[c][c];
    for (int r = c - 1; r >= 0; --r) {
      if (R[r][r] == 0.0) continue;
      double v = 0.0;
      for (int k = r + 1; k <= c; ++k) v += R[r][k] * Ri[k][c];
      Ri[r][c] = -v / R[r][r];
    }
  }
  __syncthreads();
  if (t < nb * nb) {
    const int c = t / nb, d = t % nb;
    Rout[t] = R[c][d];
    Rinv[t] = Ri[c][d];
    if (Bout) {
      double v = 0.0;
      for (int k = 0; k < nb; ++k) v += R[c][k] * (Rprev ? Rprev[k * nb + d] : (k == d ? 1.0 : 0.0));
      Bout[t] = v;
      if (BTout) BTout[d * nb + c] = v;
    }
  }
}

// From the projection out[r * 2nb + j] (j < nb: <row_r, q_j>; j >= nb:
// <row_r, w_{j-nb}>) of step i over nx compressed-exchange fields xi_j (rows
// r < nx) and the nk = (i+1) nb basis vectors (rows nx + k), all inner
// products already scaled by dV:
//   Cx[j][c] = out[j][c] = dV <xi_j, q_c>: the compressed exchange of Q_i,
//     W = H_loc Q_i - xi Cx (the local part H_loc was applied to Q_i alone);
//   X[k][j] = dV <v_k, xi_j>, kept across the steps of a run (the rows of
//     block i come from this projection: out[j][d]);
//   D[k][c] = out[nx + k][nb + c] - sum_j X[k][j] Cx[j][c] = dV <v_k, W>;
//   A = D[last nb rows]                       (A_i = Q_i^T W)
//   C[k][c] = D[k][c] - sum_d G[k][d] A[d][c], plus A on the last nb rows
//     (G[k][d] = out[nx + k][d] = dV <v_k, q_d>)
// so that one sweep W = H_loc Q_i - [xi; V] [Cx; C] applies the exchange and
// removes Q_i A_i, Q_{i-1} B_{i-1}^T and the reorthogonalisation coefficients
// together; Asym = (A + A^T)/2 is the block of T.  The D values replace the
// W columns of out in place.  One block.
__global__ void block_coef_kernel(int nk, int nb, int nx, double* __restrict__ out,
                                  double* __restrict__ C, double* __restrict__ Asym,
                                  double* __restrict__ X) {
  const int k0 = nk - nb;
  const int w2 = 2 * nb;
  double* o = out + (size_t)nx * w2;  // the basis rows
  if (nx > 0) {
    for (int t = threadIdx.x; t < nb * nx; t += blockDim.x) {
      const int d = t / nx, j = t % nx;
      X[(size_t)(k0 + d) * nx + j] = out[j * w2 + d];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nk * nb; t += blockDim.x) {
      const int k = t / nb, c = t % nb;
      double v = o[k * w2 + nb + c];
      for (int j = 0; j < nx; ++j) v -= X[(size_t)k * nx + j] * out[j * w2 + c];
      o[k * w2 + nb + c] = v;
    }
    for (int t = threadIdx.x; t < nx * nb; t += blockDim.x) C[t] = out[(t / nb) * w2 + t % nb];
    __syncthreads();
  }
  double* Cb = C + (size_t)nx * nb;
  for (int t = threadIdx.x; t < nk * nb; t += blockDim.x) {
    const int k = t / nb, c = t % nb;
    double v = o[k * w2 + nb + c];
    for (int d = 0; d < nb; ++d) v -= o[k * w2 + d] * o[(k0 + d) * w2 + nb + c];
    if (k >= k0) v += o[k * w2 + nb + c];
    Cb[t] = v;
  }
  for (int t = threadIdx.x; t < nb * nb; t += blockDim.x) {
    const int r = t / nb, c = t % nb;
    Asym[t] = 0.5 * (o[(k0 + r) * w2 + nb + c] + o[(k0 + c) * w2 + nb + r]);
  }
}

__global__ void fill_int_kernel(int* p, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Zero-fill with a kernel rather than a memset node, so the sequence stays
// correct under graph capture.
int zero_ints(int* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, 0);
  GATO_LAUNCH_CHECK();
  return GATO_OK;
}

// The projection keeps 4 LDS slices of nk * nb doubles: allow up to 150 KB.
// Function-local static initialisation runs once and is thread-safe.
template <typename T, int VEC, int NB>
void allow_block_lds() {
  static const hipError_t done = hipFuncSetAttribute(
      (const void*)block_dot_kernel<T, VEC, NB>, hipFuncAttributeMaxDynamicSharedMemorySize,
      150 * 1024);
  (void)done;
}

template <typename T, int VEC, int NB>
int launch_block_dot(long n, int nk, Rows<T> V, const void* X, long ldx, int nb, double scale,
                     double* out, void* ws, hipStream_t s) {
  constexpr long CH = 64L * VEC * (16 / NB);
  const long groups = ((n + CH - 1) / CH + 3) / 4;
  const long slots = (long)nk * nb;
  long nblk = groups < kReduceBlocks ? groups : kReduceBlocks;
  if (nblk * slots > kPartialsCap) nblk = kPartialsCap / slots;
  GATO_REQUIRE(nblk >= 1 && 4 * slots * 8 <= 150 * 1024, GATO_ERR_UNSUPPORTED,
               "block projection of %d x %d vectors exceeds the workspace", nk, nb);
  const size_t sh = (size_t)4 * slots * sizeof(double);
  allow_block_lds<T, VEC, NB>();
  hipLaunchKernelGGL((block_dot_kernel<T, VEC, NB>), dim3((unsigned)nblk), dim3(kThr), sh, s, n,
                     nk, V, (const T*)X, ldx, nb, ws_partials(ws));
  GATO_LAUNCH_CHECK();
  return finalize_partials(ws_partials(ws), nblk, (int)slots, scale, false, out, s);
}

template <typename T, int NKT>
int launch_block_dot_mfma(long n, int nk, Rows<T> V, const void* X, long ldx, int nb,
                          double scale, double* out, void* ws, hipStream_t s) {
  const long slots = (long)nk * nb;
  const long ranges = n / 64;
  long nblk = 512;  // 2048 waves
  if (nblk * slots > kPartialsCap) nblk = kPartialsCap / slots;
  if (nblk * 4 > ranges) nblk = (ranges + 3) / 4;
  const size_t sh = (size_t)4 * slots * sizeof(double);
  GATO_REQUIRE(nblk >= 1 && sh <= 150 * 1024, GATO_ERR_UNSUPPORTED,
               "block projection of %d x %d vectors exceeds the workspace", nk, nb);
  static const hipError_t lds_set = hipFuncSetAttribute(
      (const void*)block_dot_mfma_kernel<T, NKT>, hipFuncAttributeMaxDynamicSharedMemorySize,
      150 * 1024);
  (void)lds_set;
  if constexpr (sizeof(T) == 4) {
    // float: the staged kernel, 768 blocks (3 per CU; 512, 1024, 1280 and 2048
    // measured slower), the 4 x 4 x 4 shape when the columns fit its 8
    if (nb <= 16 && nk <= 64) {
      const bool f4 = nb <= 8;
      // F4 stage rows by vector count: 24 (<= 16 vectors), 40 (<= 32), 72
      // (with <= 16 vectors the 24-row tile keeps every staged row busy: 4 / 8 /
      // 12 vectors and 8 columns at 128^3 take 19.2 / 23.8 / 29.6 us against
      // 30.0 / 29.7 / 36.6 us with the 72-row tile; from 16 vectors on the
      // three run within 1-4 %; 768 blocks stay the fastest for every size)
      const int sr = !f4 ? kStageRows : nk <= 16 ? 24 : nk <= 32 ? 40 : kStageRows4;
      long nb2 = 768;
      if (nb2 * slots > kPartialsCap) nb2 = kPartialsCap / slots;
      if (nb2 * 4 > n / 32) nb2 = (n / 32 + 3) / 4;
      const size_t stage_bytes = f4 ? (size_t)4 * sr * kStagePitch4 * sizeof(float)
                                    : (size_t)4 * kStageRows * kStagePitch * sizeof(float);
      const size_t sh2 = sh > stage_bytes ? sh : stage_bytes;
      const bool probed = probe_on(PROBE_BLOCK_DOT, (int)sizeof(T));
      if (probed) probe_begin(s);
      if (f4 && sr == 24)
        hipLaunchKernelGGL((block_dot_staged_kernel<1, true, 24>), dim3((unsigned)nb2),
                           dim3(kThr), sh2, s, n, nk, V, (const float*)X, ldx, nb,
                           ws_partials(ws));
      else if (f4 && sr == 40)
        hipLaunchKernelGGL((block_dot_staged_kernel<2, true, 40>), dim3((unsigned)nb2),
                           dim3(kThr), sh2, s, n, nk, V, (const float*)X, ldx, nb,
                           ws_partials(ws));
      else if (f4)
        hipLaunchKernelGGL((block_dot_staged_kernel<NKT, true>), dim3((unsigned)nb2), dim3(kThr),
                           sh2, s, n, nk, V, (const float*)X, ldx, nb, ws_partials(ws));
      else
        hipLaunchKernelGGL((block_dot_staged_kernel<NKT, false>), dim3((unsigned)nb2), dim3(kThr),
                           sh2, s, n, nk, V, (const float*)X, ldx, nb, ws_partials(ws));
      GATO_LAUNCH_CHECK();
      if (probed) probe_end(s, (double)(nk + nb) * n * sizeof(T));
      return finalize_partials(ws_partials(ws), nb2, (int)slots, scale, false, out, s);
    }
  }
  const bool probed = probe_on(PROBE_BLOCK_DOT, (int)sizeof(T));
  if (probed) probe_begin(s);
  hipLaunchKernelGGL((block_dot_mfma_kernel<T, NKT>), dim3((unsigned)nblk), dim3(kThr), sh, s, n,
                     nk, V, (const T*)X, ldx, nb, ws_partials(ws));
  GATO_LAUNCH_CHECK();
  if (probed) probe_end(s, (double)(nk + nb) * n * sizeof(T));
  return finalize_partials(ws_partials(ws), nblk, (int)slots, scale, false, out, s);
}

template <typename T, int VEC>
int block_dot_t(long n, int nk, Rows<T> V, const void* X, long ldx, int nb, double scale,
                double* out, void* ws, hipStream_t s) {
  // The matrix-core kernel when the rows are 64-element aligned (VEC > 1
  // guarantees 16-byte aligned rows), in launches of 64 vectors (4 tiles),
  // each re-reading X.  Measured at 128^3 float, 8 columns: 128 vectors in
  // one 8-tile launch (one wave per SIMD) took 332 us, in two 4-tile
  // launches 306 us.
  if (VEC > 1 && n % 64 == 0 && nb <= 16) {
    for (int k0 = 0; k0 < nk; k0 += 64) {
      const int kc = nk - k0 < 64 ? nk - k0 : 64;
      const int st = launch_block_dot_mfma<T, 4>(n, kc, V.shift(k0), X, ldx, nb, scale,
                                                 out + (long)k0 * nb, ws, s);
      if (st) return st;
    }
    return GATO_OK;
  }
  if (nb <= 2) return launch_block_dot<T, VEC, 2>(n, nk, V, X, ldx, nb, scale, out, ws, s);
  if (nb <= 4) return launch_block_dot<T, VEC, 4>(n, nk, V, X, ldx, nb, scale, out, ws, s);
  if (nb <= 8) return launch_block_dot<T, VEC, 8>(n, nk, V, X, ldx, nb, scale, out, ws, s);
  return launch_block_dot<T, VEC, 16>(n, nk, V, X, ldx, nb, scale, out, ws, s);
}

// Gram partial count per wave grid of the update/scale kernels.
template <int VEC, int NB>
long block_waves(long n) {
  constexpr long CH = 64L * VEC * block_ept<NB>();
  return (((n + CH - 1) / CH + 3) / 4) * 4;
}

template <typename T, int VEC, int NB>
int launch_block_update(long n, int nk, Rows<T> V, const double* C, const void* X, long ldx,
                        int nb, void* Y, long ldy, double scale, double* gram, void* ws,
                        hipStream_t s) {
  const long nw = block_waves<VEC, NB>(n);
  const int ng = nb * (nb + 1) / 2;
  GATO_REQUIRE(nw * ng <= kPartialsCap && nw / 4 <= 0x7fffffffL, GATO_ERR_UNSUPPORTED,
               "block of %d vectors of %ld elements exceeds the reduction workspace", nb, n);
  GATO_REQUIRE((size_t)nk * nb * 8 <= 64 * 1024, GATO_ERR_UNSUPPORTED,
               "%d x %d update coefficients exceed LDS", nk, nb);
  const size_t sh = (size_t)(nk * nb > 0 ? nk * nb : 1) * sizeof(double);
  const bool probed = probe_on(PROBE_BLOCK_UPDATE, (int)sizeof(T));
  if (probed) probe_begin(s);
  hipLaunchKernelGGL((block_update_kernel<T, VEC, NB>), dim3((unsigned)(nw / 4)), dim3(kThr), sh, s,
                     n, nk, V, C, (const T*)X, ldx, nb, (T*)Y, ldy,
                     gram ? ws_partials(ws) : nullptr);
  GATO_LAUNCH_CHECK();
  if (probed) probe_end(s, (double)(nk + 2 * nb) * n * sizeof(T));
  if (!gram) return GATO_OK;
  return finalize_partials(ws_partials(ws), nw, ng, scale, false, gram, s);
}

template <typename T, int VEC, int NB>
int launch_block_scale(long n, const void* X, long ldx, int nb, const double* Rinv, void* Y,
                       long ldy, double scale, double* gram, void* ws, hipStream_t s) {
  const long nw = block_waves<VEC, NB>(n);
  const int ng = nb * (nb + 1) / 2;
  GATO_REQUIRE(nw * ng <= kPartialsCap && nw / 4 <= 0x7fffffffL, GATO_ERR_UNSUPPORTED,
               "block of %d vectors of %ld elements exceeds the reduction workspace", nb, n);
  hipLaunchKernelGGL((block_scale_kernel<T, VEC, NB>), dim3((unsigned)(nw / 4)), dim3(kThr), 0, s,
                     n, (const T*)X, ldx, nb, Rinv, (T*)Y, ldy, gram ? ws_partials(ws) : nullptr);
  GATO_LAUNCH_CHECK();
  if (!gram) return GATO_OK;
  return finalize_partials(ws_partials(ws), nw, ng, scale, false, gram, s);
}

#define GATO_BLOCK_DISPATCH(FN, ...)                                                 \
  do {                                                                               \
    if (nb <= 1) return FN<T, VEC, 1>(__VA_ARGS__);                                  \
    if (nb <= 2) return FN<T, VEC, 2>(__VA_ARGS__);                                  \
    if (nb <= 4) return FN<T, VEC, 4>(__VA_ARGS__);                                  \
    return FN<T, VEC, 8>(__VA_ARGS__);                                               \
  } while (0)

template <typename T, int VEC>
int block_update_t(long n, int nk, Rows<T> V, const double* C, const void* X, long ldx, int nb,
                   void* Y, long ldy, double scale, double* gram, void* ws, hipStream_t s) {
  GATO_BLOCK_DISPATCH(launch_block_update, n, nk, V, C, X, ldx, nb, Y, ldy, scale, gram, ws, s);
}

template <typename T, int VEC>
int block_scale_t(long n, const void* X, long ldx, int nb, const double* Rinv, void* Y, long ldy,
                  double scale, double* gram, void* ws, hipStream_t s) {
  GATO_BLOCK_DISPATCH(launch_block_scale, n, X, ldx, nb, Rinv, Y, ldy, scale, gram, ws, s);
}

bool vec_ok(int dtype, long n, std::initializer_list<long> lds, std::initializer_list<const void*> ps) {
  const int vec = dtype == GATO_F32 ? 4 : 2;
  if (n % vec) return false;
  for (long l : lds)
    if (l % vec) return false;
  for (const void* p : ps)
    if (!aligned16(p)) return false;
  return true;
}

}  // namespace

// out[r * nb + c] = scale * <row_r, X_c> over the rows [P (np fields, stride
// ldp); V (nk fields, stride ldv)]
int block_dot2(int dtype, long n, int np, const void* P, long ldp, int nk, const void* V, long ldv,
               const void* X, long ldx, int nb, double scale, double* out, void* ws,
               hipStream_t s) {
  GATO_REQUIRE(n >= 1 && np >= 0 && nk >= 0 && np + nk >= 1 && nb >= 1 && nb <= 2 * kMaxBlock,
               GATO_ERR_VALUE, "bad sizes n=%ld rows=%d+%d nb=%d", n, np, nk, nb);
  GATO_REQUIRE((np == 0 || P) && (nk == 0 || V) && X && out && ws, GATO_ERR_VALUE, "null pointer");
  const int nt = np + nk;
  const bool v = vec_ok(dtype, n, {np > 0 ? ldp : 0, nk > 1 ? ldv : 0, nb > 1 ? ldx : 0},
                        {np > 0 ? P : nullptr, nk > 0 ? V : nullptr, X});
  if (dtype == GATO_F32) {
    const Rows<float> R{(const float*)P, ldp, np, (const float*)V, ldv};
    return v ? block_dot_t<float, 4>(n, nt, R, X, ldx, nb, scale, out, ws, s)
             : block_dot_t<float, 1>(n, nt, R, X, ldx, nb, scale, out, ws, s);
  }
  const Rows<double> R{(const double*)P, ldp, np, (const double*)V, ldv};
  return v ? block_dot_t<double, 2>(n, nt, R, X, ldx, nb, scale, out, ws, s)
           : block_dot_t<double, 1>(n, nt, R, X, ldx, nb, scale, out, ws, s);
}

// out[k * nb + c] = scale * <V_k, X_c>
int block_dot(int dtype, long n, int nk, const void* V, long ldv, const void* X, long ldx, int nb,
              double scale, double* out, void* ws, hipStream_t s) {
  GATO_REQUIRE(nk >= 1, GATO_ERR_VALUE, "bad sizes n=%ld nk=%d nb=%d", n, nk, nb);
  return block_dot2(dtype, n, 0, nullptr, 0, nk, V, ldv, X, ldx, nb, scale, out, ws, s);
}

// Y = X - [P; V] C (C: (np + nk) x nb, row r = coefficients of row r of
// [P; V]); gram (optional, nb(nb+1)/2 doubles, packed upper triangle) =
// scale * Y^T Y
int block_update2(int dtype, long n, int np, const void* P, long ldp, int nk, const void* V,
                  long ldv, const double* C, const void* X, long ldx, int nb, void* Y, long ldy,
                  double scale, double* gram, void* ws, hipStream_t s) {
  GATO_REQUIRE(n >= 1 && np >= 0 && nk >= 0 && nb >= 1 && nb <= kMaxBlock, GATO_ERR_VALUE,
               "bad sizes n=%ld rows=%d+%d nb=%d (nb <= %d)", n, np, nk, nb, kMaxBlock);
  GATO_REQUIRE(X && Y && ws && (np == 0 || P) && (nk == 0 || V) && (np + nk == 0 || C),
               GATO_ERR_VALUE, "null pointer");
  const int nt = np + nk;
  const bool v = vec_ok(dtype, n,
                        {np > 0 ? ldp : 0, nk > 1 ? ldv : 0, nb > 1 ? ldx : 0, nb > 1 ? ldy : 0},
                        {np > 0 ? P : nullptr, nk > 0 ? V : nullptr, X, Y});
  if (dtype == GATO_F32) {
    const Rows<float> R{(const float*)P, ldp, np, (const float*)V, ldv};
    return v ? block_update_t<float, 4>(n, nt, R, C, X, ldx, nb, Y, ldy, scale, gram, ws, s)
             : block_update_t<float, 1>(n, nt, R, C, X, ldx, nb, Y, ldy, scale, gram, ws, s);
  }
  const Rows<double> R{(const double*)P, ldp, np, (const double*)V, ldv};
  return v ? block_update_t<double, 2>(n, nt, R, C, X, ldx, nb, Y, ldy, scale, gram, ws, s)
           : block_update_t<double, 1>(n, nt, R, C, X, ldx, nb, Y, ldy, scale, gram, ws, s);
}

// Y = X - V C (C: nk x nb, row k = coefficients of V_k); gram as block_update2
int block_update(int dtype, long n, int nk, const void* V, long ldv, const double* C,
                 const void* X, long ldx, int nb, void* Y, long ldy, double scale, double* gram,
                 void* ws, hipStream_t s) {
  return block_update2(dtype, n, 0, nullptr, 0, nk, V, ldv, C, X, ldx, nb, Y, ldy, scale, gram, ws,
                       s);
}

int block_scale(int dtype, long n, const void* X, long ldx, int nb, const double* Rinv, void* Y,
                long ldy, double scale, double* gram, void* ws, hipStream_t s) {
  GATO_REQUIRE(n >= 1 && nb >= 1 && nb <= kMaxBlock, GATO_ERR_VALUE, "bad sizes n=%ld nb=%d", n,
               nb);
  GATO_REQUIRE(X && Y && Rinv && ws, GATO_ERR_VALUE, "null pointer");
  const bool v = vec_ok(dtype, n, {nb > 1 ? ldx : 0, nb > 1 ? ldy : 0}, {X, Y});
  if (dtype == GATO_F32)
    return v ? block_scale_t<float, 4>(n, X, ldx, nb, Rinv, Y, ldy, scale, gram, ws, s)
             : block_scale_t<float, 1>(n, X, ldx, nb, Rinv, Y, ldy, scale, gram, ws, s);
  return v ? block_scale_t<double, 2>(n, X, ldx, nb, Rinv, Y, ldy, scale, gram, ws, s)
           : block_scale_t<double, 1>(n, X, ldx, nb, Rinv, Y, ldy, scale, gram, ws, s);
}

namespace {
// packed upper triangle of the symmetric part of M (n x n row-major)
__global__ __launch_bounds__(64) void sym_pack_kernel(int n, const double* __restrict__ M,
                                                      double* __restrict__ G) {
  for (int t = threadIdx.x; t < n * n; t += 64) {
    const int c = t / n, d = t % n;
    if (c <= d) G[gram_slot(c, d, n)] = 0.5 * (M[c * n + d] + M[d * n + c]);
  }
}
}  // namespace

// xi <- xi R^{-1}, R^T R the symmetric part of M = coef[0 .. n^2): the factor
// and inverse of the compressed exchange, on the device (no host round
// trip).  coef holds >= 4 n^2 doubles (M, its packed triangle, R, R^{-1}).
int ace_factor(int dtype, long len, int n, void* xi, double* coef, void* ws, hipStream_t s) {
  GATO_REQUIRE(n >= 1 && n <= kMaxBlock, GATO_ERR_VALUE, "compressed exchange of %d orbitals", n);
  double* G = coef + (size_t)n * n;
  double* R = G + (size_t)n * n;
  double* Ri = R + (size_t)n * n;
  hipLaunchKernelGGL(sym_pack_kernel, dim3(1), dim3(64), 0, s, n, (const double*)coef, G);
  GATO_LAUNCH_CHECK();
  hipLaunchKernelGGL(block_chol_kernel, dim3(1), dim3(64), 0, s, n, (const double*)G, 0.0,
                     (const double*)nullptr, R, Ri, (double*)nullptr, (double*)nullptr,
                     (int*)nullptr);
  GATO_LAUNCH_CHECK();
  return block_scale(dtype, len, xi, len, n, Ri, xi, len, 1.0, nullptr, ws, s);
}

// Device scratch of one block orthonormalisation / Lanczos step (doubles).
struct BlockScratch {
  double* gram;    // nb(nb+1)/2
  double* R1;      // nb*nb
  double* R1inv;   // nb*nb
  double* R2;      // nb*nb
  double* R2inv;   // nb*nb
  double* BT;      // nb*nb: transpose of the last B
  double* proj;    // nk_max * 2nb
  double* C;       // nk_max * nb
  static size_t doubles(int nb, int nk_max) {
    return (size_t)nb * (nb + 1) / 2 + 5 * (size_t)nb * nb + 3 * (size_t)nk_max * nb;
  }
  BlockScratch(double* base, int nb, int nk_max) {
    gram = base;
    R1 = gram + nb * (nb + 1) / 2;
    R1inv = R1 + nb * nb;
    R2 = R1inv + nb * nb;
    R2inv = R2 + nb * nb;
    BT = R2inv + nb * nb;
    proj = BT + nb * nb;
    C = proj + (size_t)nk_max * 2 * nb;
  }
};

// Q R = X by CholQR2 (Y may alias X): B (nb x nb, row-major upper) = R, BT its
// transpose, dropped[c] set for dependent columns (their Q column is zero).
int block_orthonormalize(int dtype, long n, int nb, const void* X, long ldx, void* Y, long ldy,
                         double scale, double rel_tol, BlockScratch& sc, double* B, double* BT,
                         int* dropped, void* ws, hipStream_t s) {
  // copy + Gram, factor, scale + Gram, factor, scale
  int st = block_update(dtype, n, 0, nullptr, 0, nullptr, X, ldx, nb, Y, ldy, scale, sc.gram, ws, s);
  if (st) return st;
  hipLaunchKernelGGL(block_chol_kernel, dim3(1), dim3(64), 0, s, nb, sc.gram, rel_tol, nullptr,
                     sc.R1, sc.R1inv, nullptr, nullptr, dropped);
  GATO_LAUNCH_CHECK();
  st = block_scale(dtype, n, Y, ldy, nb, sc.R1inv, Y, ldy, scale, sc.gram, ws, s);
  if (st) return st;
  // the second pass only refines: its factor is close to the identity
  hipLaunchKernelGGL(block_chol_kernel, dim3(1), dim3(64), 0, s, nb, sc.gram, 0.0, sc.R1, sc.R2,
                     sc.R2inv, B, BT, nullptr);
  GATO_LAUNCH_CHECK();
  return block_scale(dtype, n, Y, ldy, nb, sc.R2inv, Y, ldy, scale, nullptr, ws, s);
}

// One block Lanczos step i on V (blocks 0..i orthonormal, block i+1 holds
// W = H Q_i on entry and Q_{i+1} on return; with nx > 0 compressed-exchange
// fields xi, W = H_loc Q_i without the exchange, which this step subtracts,
// and X (rows of nx doubles per basis vector, kept by the caller across the
// steps of a run) collects dV <v_k, xi_j>).  Writes Asym (T's diagonal
// block), B (B_i), BT and dropped flags.  The three-term recurrence's
// W - Q_{i-1} B_{i-1}^T (lanczos.py:146-148) is not a pass of its own: the
// projection of W on the whole basis carries the Q_{i-1} component
// (<Q_{i-1}, W> = B_{i-1}^T up to rounding) and the update removes it with
// the others, one read of the basis and of W instead of two, and one rounding
// of W to its storage type instead of two.
int block_lanczos_orth(int dtype, long n, int i, int nb, void* V, double scale, double rel_tol,
                       BlockScratch& sc, double* Asym, double* B, int* dropped, void* ws,
                       hipStream_t s, int nx, const void* xi, double* X) {
  const size_t es = dtype == GATO_F32 ? 4 : 8;
  char* Vb = (char*)V;
  void* Qi = Vb + (size_t)i * nb * n * es;
  void* W = Vb + (size_t)(i + 1) * nb * n * es;
  int st;
  const int nk = (i + 1) * nb;
  // one pass: <xi_j, q_c>, <xi_j, w_c>, <v_k, q_c> and <v_k, w_c> for the nx
  // compressed-exchange fields and the nk basis vectors
  st = block_dot2(dtype, n, nx, xi, n, nk, V, n, Qi, n, 2 * nb, scale, sc.proj, ws, s);
  if (st) return st;
  hipLaunchKernelGGL(block_coef_kernel, dim3(1), dim3(256), 0, s, nk, nb, nx, sc.proj, sc.C, Asym,
                     X);
  GATO_LAUNCH_CHECK();
  // W -= [xi; V] [Cx; C] (the compressed exchange of Q_i, Q_i A_i, Q_{i-1}
  // B_{i-1}^T and the reorthogonalisation components), Gram of W
  st = block_update2(dtype, n, nx, xi, n, nk, V, n, sc.C, W, n, nb, W, n, scale, sc.gram, ws, s);
  if (st) return st;
  hipLaunchKernelGGL(block_chol_kernel, dim3(1), dim3(64), 0, s, nb, sc.gram, rel_tol, nullptr,
                     sc.R1, sc.R1inv, nullptr, nullptr, dropped);
  GATO_LAUNCH_CHECK();
  st = block_scale(dtype, n, W, n, nb, sc.R1inv, W, n, scale, sc.gram, ws, s);
  if (st) return st;
  hipLaunchKernelGGL(block_chol_kernel, dim3(1), dim3(64), 0, s, nb, sc.gram, 0.0, sc.R1, sc.R2,
                     sc.R2inv, B, sc.BT, nullptr);
  GATO_LAUNCH_CHECK();
  return block_scale(dtype, n, W, n, nb, sc.R2inv, W, n, scale, nullptr, ws, s);
}

// ---------------------------------------------------------------------------
// Host dense symmetric eigensolver for the K lowest pairs: Householder
// reduction to tridiagonal form (Golub & Van Loan, Algorithm 8.3.1) keeping
// the reflectors, the K lowest eigenvalues of the tridiagonal by bisection,
// its K eigenvectors by inverse iteration, and the back-transform of
// those K vectors by the reflectors: O(m^3) for the reduction only, instead of
// also accumulating all m eigenvectors.  A is m x m row-major and is
// overwritten.  Ascending eigenvalues; each eigenvector's largest-magnitude
// component positive (the gauge of the scalar path).
// ---------------------------------------------------------------------------
int sym_eigh_lowest(int m, std::vector<double> A, int K, double* evals, double* evecs) {
  std::vector<double> d(m), e(m, 0.0), p(m), w(m);
  std::vector<double> refl((size_t)m * m, 0.0), rbeta(m, 0.0);  // reflector k: refl[k*m + 0..len)
  auto a = [&](int r, int c) -> double& { return A[(size_t)r * m + c]; };
  for (int k = 0; k + 2 < m; ++k) {
    const int len = m - k - 1;
    double* v = &refl[(size_t)k * m];
    double xn2 = 0.0;
    for (int j = 0; j < len; ++j) xn2 += a(k + 1 + j, k) * a(k + 1 + j, k);
    const double x0 = a(k + 1, k);
    if (xn2 == 0.0 || xn2 == x0 * x0) {  // column already reduced
      e[k] = x0;
      continue;
    }
    const double alpha = x0 >= 0 ? -std::sqrt(xn2) : std::sqrt(xn2);
    for (int j = 0; j < len; ++j) v[j] = a(k + 1 + j, k);
    v[0] -= alpha;
    double vn2 = 0.0;
    for (int j = 0; j < len; ++j) vn2 += v[j] * v[j];
    const double beta = 2.0 / vn2;
    rbeta[k] = beta;
    // p = beta A22 v; w = p - (beta p.v / 2) v; A22 -= v w^T + w v^T
    double pv = 0.0;
    for (int r = 0; r < len; ++r) {
      const double* row = &A[(size_t)(k + 1 + r) * m + k + 1];
      double t = 0.0;
      for (int c = 0; c < len; ++c) t += row[c] * v[c];
      p[r] = beta * t;
      pv += p[r] * v[r];
    }
    const double Kc = 0.5 * beta * pv;
    for (int r = 0; r < len; ++r) w[r] = p[r] - Kc * v[r];
    for (int r = 0; r < len; ++r) {
      double* row = &A[(size_t)(k + 1 + r) * m + k + 1];
      const double vr = v[r], wr = w[r];
      for (int c = 0; c < len; ++c) row[c] -= vr * w[c] + wr * v[c];
    }
    e[k] = alpha;
  }
  if (m >= 2) e[m - 2] = a(m - 1, m - 2);
  for (int i = 0; i < m; ++i) d[i] = a(i, i);
  tridiag_lowest_eigvals(m, d.data(), e.data(), K, evals);
  tridiag_inverse_iteration(m, d.data(), e.data(), K, evals, evecs);
  // back-transform: x = P_0 P_1 ... P_{m-3} y
  for (int k = m - 3; k >= 0; --k) {
    if (rbeta[k] == 0.0) continue;
    const double* v = &refl[(size_t)k * m];
    const int len = m - k - 1;
    for (int q = 0; q < K; ++q) {
      double t = 0.0;
      for (int j = 0; j < len; ++j) t += v[j] * evecs[(size_t)(k + 1 + j) * K + q];
      t *= rbeta[k];
      for (int j = 0; j < len; ++j) evecs[(size_t)(k + 1 + j) * K + q] -= t * v[j];
    }
  }
  for (int q = 0; q < K; ++q) {
    double big = 0.0;
    int at = 0;
    for (int j = 0; j < m; ++j)
      if (std::fabs(evecs[(size_t)j * K + q]) > big) { big = std::fabs(evecs[(size_t)j * K + q]); at = j; }
    if (evecs[(size_t)at * K + q] < 0)
      for (int j = 0; j < m; ++j) evecs[(size_t)j * K + q] = -evecs[(size_t)j * K + q];
  }
  return GATO_OK;
}

// The K lowest eigenvalues (ascending) of the symmetric tridiagonal (d, e) by
// bisection on Sturm counts, as tridiag_lowest_eigvals, with the K brackets
// advanced in lockstep: each pass over the matrix evaluates the counts of up
// to four midpoints, so the divisions of four independent recurrences overlap
// instead of waiting on each other.  Each bracket starts from the Gershgorin
// interval and stops at the rounding level of ||T||.
void tridiag_lowest_eigvals_lockstep(int m, const double* a, const double* b, int K,
                                     double* out) {
  const double eps = 2.220446049250313e-16;
  double lo = a[0], hi = a[0], bmax2 = 0.0;
  for (int i = 0; i < m; ++i) {
    double r = 0.0;
    if (i > 0) r += std::fabs(b[i - 1]);
    if (i + 1 < m) r += std::fabs(b[i]);
    lo = std::fmin(lo, a[i] - r);
    hi = std::fmax(hi, a[i] + r);
    if (i + 1 < m) bmax2 = std::fmax(bmax2, b[i] * b[i]);
  }
  const double tn = std::fmax(std::fabs(lo), std::fabs(hi));
  const double pivmin = std::fmax(1e-300, 1e-290 * bmax2);
  lo -= 2.0 * eps * tn + pivmin;
  hi += 2.0 * eps * tn + pivmin;
  thread_local std::vector<double> b2;
  b2.assign(m, 0.0);
  for (int i = 0; i + 1 < m; ++i) b2[i + 1] = b[i] * b[i];
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int nc = std::min(4, K - k0);
    double l[4], h[4], mid[4];
    bool act[4];
    for (int c = 0; c < 4; ++c) {
      l[c] = lo;
      h[c] = hi;
      act[c] = c < nc;
    }
    for (int it = 0; it < 200; ++it) {
      bool any = false;
      for (int c = 0; c < 4; ++c) {
        if (act[c]) {
          mid[c] = 0.5 * (l[c] + h[c]);
          if (h[c] - l[c] <= 2.0 * eps * std::fmax(std::fabs(l[c]), std::fabs(h[c])) + pivmin ||
              mid[c] == l[c] || mid[c] == h[c])
            act[c] = false;
        }
        if (!act[c]) mid[c] = hi;  // evaluated and ignored
        any |= act[c];
      }
      if (!any) break;
      double q[4];
      int cnt[4];
      for (int c = 0; c < 4; ++c) {
        q[c] = a[0] - mid[c];
        if (std::fabs(q[c]) < pivmin) q[c] = -pivmin;
        cnt[c] = q[c] < 0;
      }
      for (int i = 1; i < m; ++i) {
        const double ai = a[i], bi = b2[i];
        for (int c = 0; c < 4; ++c) {
          double t = (ai - mid[c]) - bi / q[c];
          if (std::fabs(t) < pivmin) t = -pivmin;
          q[c] = t;
          cnt[c] += t < 0;
        }
      }
      for (int c = 0; c < 4; ++c)
        if (act[c]) {
          if (cnt[c] > k0 + c) h[c] = mid[c];
          else l[c] = mid[c];
        }
    }
    for (int c = 0; c < nc; ++c) out[k0 + c] = 0.5 * (l[c] + h[c]);
  }
}

// The K lowest eigenpairs of a symmetric band matrix of half-bandwidth w held
// as its lower band Lb[r * (w + 2) + d] = a(r, r - d), d = 0 .. w + 1 (the
// last diagonal zero on entry: room for the bulge).  Reduction to tridiagonal
// form by Givens rotations of adjacent indices, the bulge each one makes
// chased down the band (Schwarz, Numer. Math. 12, 231 (1968)); each rotation
// updates the lower band only, O(m^2 w) in all.  Then bisection and inverse
// iteration on the tridiagonal, the rotations applied back to those K vectors,
// and the gauge of sym_eigh_lowest (largest component positive).  Lb is
// overwritten.
int band_eigh_lowest(int m, int w, double* Lb, int K, double* evals, double* evecs) {
  const int LD = w + 2;
  auto L = [&](int r, int c) -> double& { return Lb[(size_t)r * LD + (r - c)]; };
  struct Rot {
    int p;
    double c, s;
  };
  thread_local std::vector<Rot> rots;
  rots.clear();
  // similarity by the rotation (c, s) of the indices (p, p + 1):
  // row p <- c row p + s row q, row q <- -s row p + c row q, and the columns
  auto rotate = [&](int p, double c, double s) {
    const int q = p + 1;
    for (int j = std::max(0, p - w); j < p; ++j) {
      double& x = L(p, j);
      double& y = L(q, j);
      const double xv = x, yv = y;
      x = c * xv + s * yv;
      y = c * yv - s * xv;
    }
    const double app = L(p, p), aqq = L(q, q), aqp = L(q, p);
    const double cc = c * c, ss = s * s, cs = c * s;
    L(p, p) = cc * app + 2.0 * cs * aqp + ss * aqq;
    L(q, q) = ss * app - 2.0 * cs * aqp + cc * aqq;
    L(q, p) = cs * (aqq - app) + (cc - ss) * aqp;
    const int hi = std::min(m - 1, q + w);
    for (int j = q + 1; j <= hi; ++j) {
      double& x = L(j, p);
      double& y = L(j, q);
      const double xv = x, yv = y;
      x = c * xv + s * yv;
      y = c * yv - s * xv;
    }
    rots.push_back({p, c, s});
  };
  // zero a(p + 1, k) against a(p, k), k < p
  auto annihilate = [&](int p, int k) {
    const double x = L(p, k), y = L(p + 1, k);
    const double big = std::fmax(std::fabs(x), std::fabs(y));
    const double r = (big > 1e-150 && big < 1e150) ? std::sqrt(x * x + y * y) : std::hypot(x, y);
    rotate(p, x / r, y / r);
    L(p + 1, k) = 0.0;
  };
  for (int k = 0; w > 1 && k + 2 < m; ++k) {
    for (int i = std::min(k + w, m - 1); i >= k + 2; --i) {
      if (L(i, k) == 0.0) continue;
      annihilate(i - 1, k);
      // the rotation of (i - 1, i) fills (i + w, i - 1); chase it down the band
      int p = i - 1, q = i + w;
      while (q < m && L(q, p) != 0.0) {
        annihilate(q - 1, p);
        p = q - 1;
        q += w;
      }
    }
  }
  thread_local std::vector<double> d, e;
  d.assign(m, 0.0);
  e.assign(m, 0.0);
  for (int i = 0; i < m; ++i) d[i] = L(i, i);
  for (int i = 0; i + 1 < m; ++i) e[i] = L(i + 1, i);
  tridiag_lowest_eigvals_lockstep(m, d.data(), e.data(), K, evals);
  tridiag_inverse_iteration(m, d.data(), e.data(), K, evals, evecs);
  // x = G_1 G_2 ... G_N y: the last rotation first
  for (size_t r = rots.size(); r-- > 0;) {
    const Rot& g = rots[r];
    double* x0 = evecs + (size_t)g.p * K;
    double* x1 = evecs + (size_t)(g.p + 1) * K;
    for (int q = 0; q < K; ++q) {
      const double u = x0[q], v = x1[q];
      x0[q] = g.c * u - g.s * v;
      x1[q] = g.s * u + g.c * v;
    }
  }
  for (int q = 0; q < K; ++q) {
    double big = 0.0;
    int at = 0;
    for (int j = 0; j < m; ++j)
      if (std::fabs(evecs[(size_t)j * K + q]) > big) { big = std::fabs(evecs[(size_t)j * K + q]); at = j; }
    if (evecs[(size_t)at * K + q] < 0)
      for (int j = 0; j < m; ++j) evecs[(size_t)j * K + q] = -evecs[(size_t)j * K + q];
  }
  return GATO_OK;
}

// Rayleigh-Ritz on the block-tridiagonal T of s steps of block size nb
// (A, B: s blocks of nb x nb, row-major; each A_i symmetric, its lower
// triangle read) over the columns that were not dropped (drop: s+1 rows of nb
// flags; a dropped column of block i+1 is a zero basis vector and is left out
// of T).  The kept matrix goes straight into band storage.  S (m x K,
// m = s*nb) holds zero rows for dropped columns.
int block_ritz(int nb, int s, const double* A, const double* B, const int* drop, int K,
               double* evals, double* S) {
  const int m = s * nb;
  thread_local std::vector<int> kidx;
  kidx.assign(m, -1);
  int mk = 0;
  for (int j = 0; j < m; ++j)
    if (!drop[j]) kidx[j] = mk++;
  GATO_REQUIRE(mk >= K, GATO_ERR_VALUE,
               "only %d independent Krylov vectors for %d eigenpairs", mk, K);
  // entries of the lower triangle: (row, col, value) of A_i (c <= r) and B_i
  // (block i+1 row r, block i column c); the half-bandwidth of the kept
  // matrix (nb when every B_i is upper triangular, as Cholesky QR makes it)
  int w = 0;
  auto visit = [&](auto&& f) {
    for (int i = 0; i < s; ++i)
      for (int r = 0; r < nb; ++r) {
        const int gr = i * nb + r;
        for (int c = 0; c <= r; ++c) f(gr, i * nb + c, A[(size_t)i * nb * nb + r * nb + c]);
        if (i + 1 < s)
          for (int c = 0; c < nb; ++c)
            f((i + 1) * nb + r, i * nb + c, B[(size_t)i * nb * nb + r * nb + c]);
      }
  };
  visit([&](int r, int c, double v) {
    if (v != 0.0 && kidx[r] >= 0 && kidx[c] >= 0) w = std::max(w, kidx[r] - kidx[c]);
  });
  const int LD = w + 2;
  thread_local std::vector<double> Lb;
  Lb.assign((size_t)mk * LD, 0.0);
  visit([&](int r, int c, double v) {
    const int kr = kidx[r], kc = kidx[c];
    if (v != 0.0 && kr >= 0 && kc >= 0) Lb[(size_t)kr * LD + (kr - kc)] = v;
  });
  thread_local std::vector<double> Sk;
  Sk.assign((size_t)mk * K, 0.0);
  const int st = band_eigh_lowest(mk, w, Lb.data(), K, evals, Sk.data());
  if (st) return st;
  std::fill(S, S + (size_t)m * K, 0.0);
  for (int j = 0; j < m; ++j)
    if (kidx[j] >= 0)
      for (int k = 0; k < K; ++k) S[(size_t)j * K + k] = Sk[(size_t)kidx[j] * K + k];
  return GATO_OK;
}

}  // namespace gato

using namespace gato;

extern "C" int gato_block_dot(int dtype, long n, int nk, const void* V, long ldv, const void* X,
                              long ldx, int nb, double scale, double* out, void* ws,
                              void* stream) {
  clear_ok();
  return block_dot(dtype, n, nk, V, ldv, X, ldx, nb, scale, out, ws, as_stream(stream));
}

extern "C" int gato_block_update(int dtype, long n, int nk, const void* V, long ldv,
                                 const double* C, const void* X, long ldx, int nb, void* Y,
                                 long ldy, double scale, double* gram, void* ws, void* stream) {
  clear_ok();
  return block_update(dtype, n, nk, V, ldv, C, X, ldx, nb, Y, ldy, scale, gram, ws,
                      as_stream(stream));
}

extern "C" size_t gato_block_scratch_bytes(int nb, int s) {
  return (BlockScratch::doubles(nb, (s + 1) * nb + kMaxBlock) + 2 * (size_t)s * nb * nb +
          (size_t)(s + 1) * nb + (size_t)(s + 1) * nb * 64 + (size_t)(s + 1) * nb * kMaxBlock) *
             sizeof(double) +
         256;
}

namespace {
struct BlockLayout {
  // the step scratch, with room for kMaxBlock compressed-exchange rows ahead
  // of the basis in the projection and the update coefficients
  BlockScratch sc;
  // step records of RD = 2 nb^2 + nb doubles: A_i | B_i | the flags of block
  // i + 1 (nb ints), so one copy stages a step's share of T
  double* rec;
  int* drop0;    // flags of block 0 (nb ints; nb doubles of room)
  double* S;     // ((s + 1) * nb) * 64 Ritz coefficients
  double* X;     // ((s + 1) * nb) * kMaxBlock: dV <v_k, xi_j>
  size_t nbb, RD;
  BlockLayout(void* base, int nb, int s)
      : sc((double*)base, nb, (s + 1) * nb + kMaxBlock), nbb((size_t)nb * nb),
        RD(2 * (size_t)nb * nb + nb) {
    rec = (double*)base + BlockScratch::doubles(nb, (s + 1) * nb + kMaxBlock);
    drop0 = (int*)(rec + (size_t)s * RD);
    S = (double*)drop0 + nb;
    X = S + (size_t)(s + 1) * nb * 64;
  }
  double* A(int i) const { return rec + (size_t)i * RD; }
  double* B(int i) const { return rec + (size_t)i * RD + nbb; }
  int* drop(int block) const { return block == 0 ? drop0 : (int*)(rec + (size_t)(block - 1) * RD + 2 * nbb); }
};

// contiguous A (s nb x nb), B (s nb x nb) and flags ((s + 1) x nb) from step
// records (host copies)
void unpack_records(const double* rec, const int* drop0, int nb, int s, double* A, double* B,
                    int* drop) {
  const size_t nbb = (size_t)nb * nb, RD = 2 * nbb + nb;
  for (int i = 0; i < s; ++i) {
    std::memcpy(A + i * nbb, rec + i * RD, sizeof(double) * nbb);
    std::memcpy(B + i * nbb, rec + i * RD + nbb, sizeof(double) * nbb);
    std::memcpy(drop + (size_t)(i + 1) * nb, rec + i * RD + 2 * nbb, sizeof(int) * nb);
  }
  std::memcpy(drop, drop0, sizeof(int) * nb);
}
}  // namespace

extern "C" int gato_block_orthonormalize(int dtype, long n, int nb, const void* X, long ldx,
                                         void* Y, long ldy, double scale, double rel_tol,
                                         double* R, int* dropped, void* scratch, void* ws,
                                         void* stream) {
  clear_ok();
  GATO_REQUIRE(X && Y && R && dropped && scratch && ws, GATO_ERR_VALUE, "null pointer");
  GATO_REQUIRE(nb >= 1 && nb <= kMaxBlock, GATO_ERR_VALUE, "block size %d (1..%d)", nb, kMaxBlock);
  hipStream_t s = as_stream(stream);
  BlockScratch sc((double*)scratch, nb, nb);
  const int st = zero_ints(dropped, nb, s);
  if (st) return st;
  return block_orthonormalize(dtype, n, nb, X, ldx, Y, ldy, scale, rel_tol, sc, R, sc.BT, dropped,
                              ws, s);
}

extern "C" int gato_block_lanczos_step(int dtype, long n, int i, int nb, void* V, double scale,
                                       double rel_tol, const double* Bprev, double* A, double* B,
                                       int* dropped, void* scratch, void* ws, void* stream) {
  clear_ok();
  GATO_REQUIRE(V && A && B && dropped && scratch && ws, GATO_ERR_VALUE, "null pointer");
  GATO_REQUIRE(i >= 0 && nb >= 1 && nb <= kMaxBlock, GATO_ERR_VALUE, "bad step %d / block %d", i,
               nb);
  hipStream_t s = as_stream(stream);
  BlockScratch sc((double*)scratch, nb, (i + 1) * nb);
  (void)Bprev;  // the projection pass removes the Q_{i-1} component (block_lanczos_orth)
  int st = zero_ints(dropped, nb, s);
  if (st) return st;
  return block_lanczos_orth(dtype, n, i, nb, V, scale, rel_tol, sc, A, B, dropped, ws, s, 0, nullptr,
                            nullptr);
}

extern "C" size_t gato_block_step_scratch_bytes(int nb, int nk_max) {
  return BlockScratch::doubles(nb, nk_max) * sizeof(double);
}

extern "C" int gato_block_ritz(int nb, int s, const double* A, const double* B, const int* dropped,
                               int K, double* evals, double* S) {
  clear_ok();
  GATO_REQUIRE(nb >= 1 && s >= 1 && K >= 1 && A && B && dropped && evals && S, GATO_ERR_VALUE,
               "bad arguments");
  return block_ritz(nb, s, A, B, dropped, K, evals, S);
}

extern "C" int gato_sym_eigh(int m, const double* A, int K, double* evals, double* evecs) {
  clear_ok();
  GATO_REQUIRE(m >= 1 && K >= 1 && K <= m && A && evals && evecs, GATO_ERR_VALUE,
               "bad arguments m=%d K=%d", m, K);
  std::vector<double> a(A, A + (size_t)m * m);
  const int st = sym_eigh_lowest(m, std::move(a), K, evals, evecs);
  if (st) set_error(st, "symmetric eigensolver did not converge");
  return st;
}

namespace {
// Page-locked host staging for the per-step T blocks of the residual checks
// (an asynchronous copy into pageable memory would wait for the copy, and so
// for the step that is meant to run behind the check).  One buffer per host
// thread, grown on demand and kept for the thread's lifetime.
struct PinnedStage {
  void* p = nullptr;
  size_t bytes = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // recorded after the last copy of a run that returned without waiting for
  // it (the step behind the last check, the Ritz coefficients): the next run
  // of this thread waits for it, and for the last step events, before it
  // reuses the buffer, whatever stream it runs on
  hipEvent_t done = nullptr;
  bool pending = false;
  ~PinnedStage() {
    if (p) (void)hipHostFree(p);
    for (hipEvent_t e : {ev[0], ev[1], done})
      if (e) (void)hipEventDestroy(e);
  }
  int reserve(size_t want) {
    if (!ev[0]) {
      GATO_HIP_CHECK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
      GATO_HIP_CHECK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
      GATO_HIP_CHECK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    }
    if (pending) {
      GATO_HIP_CHECK(hipEventSynchronize(done));
      pending = false;
    }
    // the step copies of a run that ended early (events never recorded return
    // at once)
    GATO_HIP_CHECK(hipEventSynchronize(ev[0]));
    GATO_HIP_CHECK(hipEventSynchronize(ev[1]));
    if (want <= bytes) return GATO_OK;
    if (p) GATO_HIP_CHECK(hipHostFree(p));
    p = nullptr;
    bytes = 0;
    GATO_HIP_CHECK(hipHostMalloc(&p, want, hipHostMallocDefault));
    bytes = want;
    return GATO_OK;
  }
};
thread_local PinnedStage g_stage;

// Largest residual norm ||F y_k - theta_k y_k|| of the K Ritz pairs of T_s:
// the basis is orthonormal, so it is ||B_{s-1} S_last[:, k]|| with S_last the
// coefficients of the last block (the coupling T_s leaves out).
double ritz_residual(int nb, int s, const double* B, const double* S, int K) {
  const double* Bl = B + (size_t)(s - 1) * nb * nb;
  double worst = 0.0;
  for (int k = 0; k < K; ++k) {
    double r2 = 0.0;
    for (int r = 0; r < nb; ++r) {
      double t = 0.0;
      for (int c = 0; c < nb; ++c) t += Bl[r * nb + c] * S[(size_t)((s - 1) * nb + c) * K + k];
      r2 += t * t;
    }
    worst = std::max(worst, std::sqrt(r2));
  }
  return worst;
}

// The block Lanczos loop.  With res_tol > 0 the Krylov dimension adapts:
// from s_min steps on, the host forms the Ritz pairs of T after step i while
// the GPU runs step i+1 (each step's share of T is staged through page-locked
// memory by one copy behind an event), and the loop stops once every wanted
// pair has a residual below res_tol.  The result is the Ritz pairs of that
// check; the step that ran behind it is not waited for or used.  A run that
// never meets res_tol ends on the check after step s_max - 1 (the last step
// then only overlaps it).  Checks beyond 96 vectors go to every other step.
// The run ends without a host synchronisation: the Ritz coefficients go to the
// device from page-locked memory.  *s_used gets the steps in the result and
// *res_out the residual of the last check (-1 when none ran).  With res_tol =
// 0 every step runs, the final T is solved after a synchronisation, and the
// call returns with the Ritz vectors written.
int block_lanczos_run(gato_fock* f, const void* X0, long ldx, int nb, int s_max, int s_min,
                      double res_tol, int K, double dV, double rel_tol, void* V_buf,
                      void* scratch, double* A_out, double* B_out, int* dropped_out,
                      double* evals, void* eigenstates, int* s_used, double* res_out, void* ws,
                      hipStream_t s) {
  const int dt = fock_dtype(f);
  const long n = fock_n(f);
  const size_t es = dt == GATO_F32 ? 4 : 8;
  char* Vb = (char*)V_buf;
  BlockLayout lay(scratch, nb, s_max);
  const size_t nbb = lay.nbb, RD = lay.RD;
  int st = zero_ints((int*)lay.rec, (int)(s_max * RD * 2 + nb), s);  // records and drop0
  if (st) return st;
  // Q_0 R = X0
  st = block_orthonormalize(dt, n, nb, X0, ldx, V_buf, n, dV, rel_tol, lay.sc, lay.sc.R2,
                            lay.sc.BT, lay.drop(0), ws, s);
  if (st) return st;
  const bool adapt = res_tol > 0.0 && s_max > 1;
  // a compressed exchange of up to kMaxBlock fields is applied inside the
  // projection and update passes (a larger one by the operator itself)
  const void* xi = nullptr;
  const int nx_op = fock_ace(f, &xi);
  const int nx = nx_op <= kMaxBlock ? nx_op : 0;
  double *hRec = nullptr, *hS = nullptr;
  int* hD0 = nullptr;
  if (adapt) {
    st = g_stage.reserve(sizeof(double) * (s_max * RD + nb + (size_t)s_max * nb * K));
    if (st) return st;
    hRec = (double*)g_stage.p;
    hD0 = (int*)(hRec + s_max * RD);
    hS = (double*)hD0 + nb;
  }
  auto step = [&](int i) -> int {
    // W = H Q_i into block i+1 (all columns together)
    int r = fock_apply_block(f, nb, Vb + (size_t)i * nb * n * es,
                             Vb + (size_t)(i + 1) * nb * n * es, ws, s, nx == 0);
    if (r) return r;
    r = block_lanczos_orth(dt, n, i, nb, V_buf, dV, rel_tol, lay.sc, lay.A(i), lay.B(i),
                           lay.drop(i + 1), ws, s, nx, xi, lay.X);
    if (r || !adapt) return r;
    // stage record i (A_i, B_i, the flags of block i+1); block 0's flags with step 0
    if (i == 0)
      GATO_HIP_CHECK(hipMemcpyAsync(hD0, lay.drop0, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
    GATO_HIP_CHECK(hipMemcpyAsync(hRec + i * RD, lay.A(i), sizeof(double) * RD,
                                  hipMemcpyDeviceToHost, s));
    GATO_HIP_CHECK(hipEventRecord(g_stage.ev[i & 1], s));
    return GATO_OK;
  };
  *res_out = -1.0;
  st = step(0);
  if (st) return st;
  std::vector<double> hA((size_t)s_max * nbb), hB((size_t)s_max * nbb), Sc, ev_c(K);
  std::vector<int> hD((size_t)(s_max + 1) * nb);
  int s_chk = 0;  // steps of the last completed check
  for (int i = 0; i + 1 < s_max; ++i) {
    st = step(i + 1);  // runs on the GPU while the host checks step i
    if (st) return st;
    const int sc_ = i + 1, mc = sc_ * nb;
    const bool last = sc_ + 1 == s_max;
    if (!adapt || sc_ < s_min || mc < K || (mc > 96 && sc_ % 2 && !last)) continue;
    GATO_HIP_CHECK(hipEventSynchronize(g_stage.ev[i & 1]));
    unpack_records(hRec, hD0, nb, sc_, hA.data(), hB.data(), hD.data());
    int kept = 0;
    for (int j = 0; j < mc; ++j) kept += hD[j] ? 0 : 1;
    if (kept < K) continue;
    Sc.assign((size_t)mc * K, 0.0);
    st = block_ritz(nb, sc_, hA.data(), hB.data(), hD.data(), K, ev_c.data(), Sc.data());
    if (st) return st;
    *res_out = ritz_residual(nb, sc_, hB.data(), Sc.data(), K);
    s_chk = sc_;
    if (*res_out < res_tol) break;
  }
  if (adapt && s_chk > 0) {
    // finish on the last check: its T, Ritz values and coefficients
    *s_used = s_chk;
    const int m = s_chk * nb;
    std::memcpy(A_out, hA.data(), sizeof(double) * s_chk * nbb);
    std::memcpy(B_out, hB.data(), sizeof(double) * s_chk * nbb);
    std::memcpy(dropped_out, hD.data(), sizeof(int) * (size_t)(s_chk + 1) * nb);
    std::memcpy(evals, ev_c.data(), sizeof(double) * K);
    std::memcpy(hS, Sc.data(), sizeof(double) * (size_t)m * K);
    GATO_HIP_CHECK(hipMemcpyAsync(lay.S, hS, sizeof(double) * (size_t)m * K,
                                  hipMemcpyHostToDevice, s));
    GATO_HIP_CHECK(hipEventRecord(g_stage.done, s));
    g_stage.pending = true;
    return ritz_combine(dt, n, m, K, V_buf, n, lay.S, eigenstates, n, s);
  }
  // every step, then the whole T
  *s_used = s_max;
  const int m = s_max * nb;
  std::vector<double> recs((size_t)s_max * RD);
  std::vector<int> d0(nb);
  GATO_HIP_CHECK(hipMemcpyAsync(recs.data(), lay.rec, sizeof(double) * recs.size(),
                                hipMemcpyDeviceToHost, s));
  GATO_HIP_CHECK(hipMemcpyAsync(d0.data(), lay.drop0, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
  GATO_HIP_CHECK(hipStreamSynchronize(s));
  unpack_records(recs.data(), d0.data(), nb, s_max, A_out, B_out, dropped_out);
  std::vector<double> S((size_t)m * K);
  st = block_ritz(nb, s_max, A_out, B_out, dropped_out, K, evals, S.data());
  if (st) return st;
  GATO_HIP_CHECK(hipMemcpyAsync(lay.S, S.data(), sizeof(double) * S.size(), hipMemcpyHostToDevice,
                                s));
  st = ritz_combine(dt, n, m, K, V_buf, n, lay.S, eigenstates, n, s);
  if (st) return st;
  GATO_HIP_CHECK(hipStreamSynchronize(s));
  return GATO_OK;
}
}  // namespace

extern "C" int gato_block_lanczos(gato_fock* f, const void* X0, long ldx, int nb, int s_steps,
                                  int K, double dV, double rel_tol, void* V_buf, void* scratch,
                                  double* A_out, double* B_out, int* dropped_out, double* evals,
                                  void* eigenstates, void* ws, void* stream) {
  clear_ok();
  GATO_REQUIRE(f && X0 && V_buf && scratch && A_out && B_out && dropped_out && evals &&
                   eigenstates && ws,
               GATO_ERR_VALUE, "bad arguments");
  GATO_REQUIRE(nb >= 1 && nb <= kMaxBlock && s_steps >= 1 && K >= 1 && K <= 16 &&
                   K <= s_steps * nb,
               GATO_ERR_VALUE, "bad sizes b=%d s=%d K=%d", nb, s_steps, K);
  int s_used = 0;
  double res = 0.0;
  return block_lanczos_run(f, X0, ldx, nb, s_steps, s_steps, 0.0, K, dV, rel_tol, V_buf, scratch,
                           A_out, B_out, dropped_out, evals, eigenstates, &s_used, &res, ws,
                           as_stream(stream));
}

extern "C" int gato_block_lanczos_adaptive(gato_fock* f, const void* X0, long ldx, int nb,
                                           int s_max, int s_min, double res_tol, int K, double dV,
                                           double rel_tol, void* V_buf, void* scratch,
                                           double* A_out, double* B_out, int* dropped_out,
                                           double* evals, void* eigenstates, int* s_used,
                                           double* res_out, void* ws, void* stream) {
  clear_ok();
  GATO_REQUIRE(f && X0 && V_buf && scratch && A_out && B_out && dropped_out && evals &&
                   eigenstates && s_used && res_out && ws,
               GATO_ERR_VALUE, "bad arguments");
  GATO_REQUIRE(nb >= 1 && nb <= kMaxBlock && s_max >= 1 && s_min >= 1 && K >= 1 && K <= 16 &&
                   K <= s_max * nb && res_tol >= 0.0,
               GATO_ERR_VALUE, "bad sizes b=%d s=%d..%d K=%d tol=%g", nb, s_min, s_max, K,
               res_tol);
  return block_lanczos_run(f, X0, ldx, nb, s_max, s_min, res_tol, K, dV, rel_tol, V_buf, scratch,
                           A_out, B_out, dropped_out, evals, eigenstates, s_used, res_out, ws,
                           as_stream(stream));
}
