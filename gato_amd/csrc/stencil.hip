// Fused finite-difference Laplacian / Hamiltonian / Fock-local apply.
//
// out = (acc ? out : 0) + c_lap * lap(psi) + V * psi - sub_coef * sub
//
// Reference: operators.py:40-130 (7-point and 13-point stencils, Dirichlet
// zero padding and periodic wrap), hamiltonian.py:25-27 and scf.py:106-109.
//
// Design (gfx950): a 64 (z) x 8 (y) tile of threads marches along x.  Each
// thread keeps the 2R+1 values psi[x-R..x+R] of its (y, z) column in a
// register queue, so the x-neighbours cost no extra loads; the y/z neighbours
// of the current plane come from an LDS tile with an R-wide halo.  A wave
// covers 64 consecutive z (256 B of float, 512 B of double per load), so
// every global access is fully coalesced.  Each block handles a chunk of
// XC planes; blockIdx.z enumerates (batch, chunk).
#include <array>
#include <atomic>
#include <map>
#include <mutex>
#include <type_traits>

#include "common.hpp"

namespace gato {
int scale_by_partials(int dtype, long n, const void* w, const void* ws, long np, double scale,
                      double* beta, void* v, hipStream_t s);
int scale_by_partials_at(int dtype, long n, const void* w, const double* part, long np,
                         double scale, double* beta, void* v, hipStream_t s);
int normalize(int dtype, long n, const void* w, double scale, double* norm_out, void* v, void* ws,
              hipStream_t s);
int stencil_apply(int dtype, int N, int batch, double h, int order, int boundary, double c_lap,
                  const void* psi, const void* V, int accumulate, const void* sub,
                  const double* sub_coef, void* out, hipStream_t s);
namespace {

constexpr int TZ = 64;
constexpr int TY = 8;

template <typename T, int R, bool PER>
__device__ __forceinline__ T ldp(const T* __restrict__ p, int N, int x, int y, int z) {
  if (PER) {
    x = (x + N) % N; y = (y + N) % N; z = (z + N) % N;
  } else {
    if (x < 0 || x >= N || y < 0 || y >= N || z < 0 || z >= N) return T(0);
  }
  return p[((size_t)x * N + y) * N + z];
}

template <typename T, int R, bool PER>
__global__ __launch_bounds__(TZ * TY) void stencil_kernel(
    const T* __restrict__ psi, const T* __restrict__ V, T* __restrict__ out,
    const T* __restrict__ sub, const double* __restrict__ sub_coef, int N, int XC, int nxc,
    T c0, T c1, T c2, int accumulate) {
  __shared__ T tile[TY + 2 * R][TZ + 2 * R];
  const int tz = threadIdx.x, ty = threadIdx.y;
  const int zb = blockIdx.x * TZ, yb = blockIdx.y * TY;
  const int z = zb + tz, y = yb + ty;
  const int b = blockIdx.z / nxc;
  const int x0 = (blockIdx.z % nxc) * XC;
  const int x1 = min(x0 + XC, N);
  const size_t NN = (size_t)N * N;
  const size_t boff = (size_t)b * NN * N;
  const T* p = psi + boff;
  const bool active = (z < N) && (y < N);
  // Lanes outside the grid still load (zero or periodic image) so that their
  // LDS cells hold the correct neighbour values for the lanes inside.

  T q[2 * R + 1];
#pragma unroll
  for (int i = 0; i < 2 * R; ++i) q[i + 1] = ldp<T, R, PER>(p, N, x0 - R + i, y, z);

  const T sc = (sub != nullptr) ? (T)sub_coef[0] : T(0);

  for (int x = x0; x < x1; ++x) {
#pragma unroll
    for (int i = 0; i < 2 * R; ++i) q[i] = q[i + 1];
    q[2 * R] = ldp<T, R, PER>(p, N, x + R, y, z);
    const T center = q[R];

    __syncthreads();  // previous plane's readers are done with the tile
    tile[ty + R][tz + R] = center;
    if (ty < R) {
      tile[ty][tz + R] = ldp<T, R, PER>(p, N, x, yb - R + ty, z);
    } else if (ty < 2 * R) {
      const int r = ty - R;
      tile[TY + R + r][tz + R] = ldp<T, R, PER>(p, N, x, yb + TY + r, z);
    }
    if (tz < R) {
      tile[ty + R][tz] = ldp<T, R, PER>(p, N, x, y, zb - R + tz);
    } else if (tz < 2 * R) {
      const int r = tz - R;
      tile[ty + R][TZ + R + r] = ldp<T, R, PER>(p, N, x, y, zb + TZ + r);
    }
    __syncthreads();

    if (active) {
      const int cy = ty + R, cz = tz + R;
      T s1 = q[R - 1] + q[R + 1] + tile[cy - 1][cz] + tile[cy + 1][cz] + tile[cy][cz - 1] +
             tile[cy][cz + 1];
      T lap = c0 * center + c1 * s1;
      if (R == 2) {
        T s2 = q[0] + q[4] + tile[cy - 2][cz] + tile[cy + 2][cz] + tile[cy][cz - 2] +
               tile[cy][cz + 2];
        lap += c2 * s2;
      }
      const size_t idx = boff + ((size_t)x * N + y) * N + z;
      T r = lap;
      if (V != nullptr) r += V[((size_t)x * N + y) * N + z] * center;
      if (accumulate) r += out[idx];
      if (sub != nullptr) r -= sc * sub[idx];
      out[idx] = r;
    }
  }
}

// ---------------------------------------------------------------------------
// Vector kernel (N a multiple of the 16-byte vector width VEC = 16/sizeof(T)).
// Thread (tx, ty) of a 32 x 8 block owns VEC consecutive z of RYT consecutive
// y rows of one x plane, so every field access is one 16-byte load or store.
// The in-plane y neighbours of its rows come from RYT + 2R row loads shared by
// the rows, the z halo from two 2-element loads, and the x neighbours from
// the planes x +- 1, 2, which the neighbouring blocks of the same XCD have
// just read: blocks are ordered (z-tile, y-tile, x, batch) fastest first and
// remapped so each XCD (blocks are dealt to the 8 XCDs round-robin) works
// through one contiguous range of planes, keeping them in its L2.
// ---------------------------------------------------------------------------
constexpr int VTX = 32;  // z lanes per block
constexpr int VTY = 8;   // y threads per block
constexpr int RYT = 2;   // y rows per thread

template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; using half = float2; static constexpr int n = 4; };
template <> struct Vec16<double> { using type = double2; using half = double2; static constexpr int n = 2; };

template <typename T, int VEC>
__device__ __forceinline__ void vld(const T* __restrict__ p, long i, bool ok, T* o) {
  using V = typename Vec16<T>::type;
  if (ok) {
    const V v = *reinterpret_cast<const V*>(p + i);
    const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
    for (int k = 0; k < VEC; ++k) o[k] = e[k];
  } else {
#pragma unroll
    for (int k = 0; k < VEC; ++k) o[k] = T(0);
  }
}

// two consecutive elements at an even index
template <typename T>
__device__ __forceinline__ void hld(const T* __restrict__ p, long i, bool ok, T* o) {
  if (ok) {
    if constexpr (sizeof(T) == 4) {
      const float2 v = *reinterpret_cast<const float2*>(p + i);
      o[0] = v.x; o[1] = v.y;
    } else {
      const double2 v = *reinterpret_cast<const double2*>(p + i);
      o[0] = v.x; o[1] = v.y;
    }
  } else {
    o[0] = T(0); o[1] = T(0);
  }
}

__device__ __forceinline__ int wrap(int i, int N) { return ((i % N) + N) % N; }

// NORM: also block partials of sum out^2 (double) to npart[blockIdx.x], the
// first half of a fused renormalisation.  With pin (NORM only), out is also
// scaled by 1 / sqrt(pscale * sum pin[0..npin)): the deferred renormalisation
// of the input, whose norm partials the previous launch left in pin.  Every
// block sums pin in the same order as scale_inv_partials_kernel, so the
// factor is the one the separate scaling pass would apply.
template <typename T, int R, bool PER, bool NORM = false>
__global__ __launch_bounds__(VTX * VTY) void stencil_vec_kernel(
    const T* __restrict__ psi, const T* __restrict__ V, T* __restrict__ out,
    const T* __restrict__ sub, const double* __restrict__ sub_coef, int N, int nzt, int nyt,
    int remap, T c0, T c1, T c2, int accumulate, double* __restrict__ npart = nullptr,
    const double* __restrict__ pin = nullptr, long npin = 0, double pscale = 0.0) {
  constexpr int VEC = Vec16<T>::n;
  const long nb = (long)gridDim.x;
  const long lb = remap ? (long)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : (long)blockIdx.x;
  const int zt = (int)(lb % nzt);
  long r1 = lb / nzt;
  const int yt = (int)(r1 % nyt);
  r1 /= nyt;
  const int x = (int)(r1 % N);
  const long b = r1 / N;
  const int z0 = (zt * VTX + threadIdx.x) * VEC;
  const int yb = (yt * VTY + threadIdx.y) * RYT;
  const bool live = z0 < N && yb < N;
  if (!NORM && !live) return;
  double nacc = 0.0;
  const long NN = (long)N * N, P = NN * N;
  const T* p = psi + b * P;
  // x neighbour planes and z halo positions (Dirichlet: validity; periodic: wrapped)
  int xs[2 * R];
  bool xok[2 * R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    xs[k] = PER ? wrap(x - R + k, N) : x - R + k;
    xs[R + k] = PER ? wrap(x + 1 + k, N) : x + 1 + k;
    xok[k] = PER || xs[k] >= 0;
    xok[R + k] = PER || xs[R + k] < N;
  }
  const int zlo = PER ? wrap(z0 - 2, N) : z0 - 2;
  const int zhi = PER ? wrap(z0 + VEC, N) : z0 + VEC;
  const bool zlo_ok = PER || z0 > 0, zhi_ok = PER || z0 + VEC < N;
  // rows yb-R .. yb+RYT+R-1 of plane x
  T row[RYT + 2 * R][VEC];
#pragma unroll
  for (int k = 0; k < RYT + 2 * R; ++k) {
    const int y = yb - R + k;
    const int yy = PER ? wrap(y, N) : y;
    vld<T, VEC>(p, ((long)x * N + yy) * N + z0, live && (PER || (y >= 0 && y < N)), row[k]);
  }
  // NORM with pin: the inverse input norm is summed while the rows load (a
  // variant that also issued the halo, neighbour and V loads first was not
  // faster)
  T isc = T(1);
  if (NORM && pin != nullptr) {
    __shared__ double pred[16];
    __shared__ double s_inv;
    const int tid = threadIdx.x + VTX * threadIdx.y;
    KSum acc;
    acc.init();
    ksum_strided(acc, pin, npin, tid, VTX * VTY);
    const double t = block_sum(acc.value(), pred);
    if (tid == 0) {
      const double nrm = sqrt(fmax(t * pscale, 0.0));
      s_inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    }
    __syncthreads();
    isc = (T)s_inv;
  }
  if (live) {
  const T sc = (sub != nullptr) ? (T)sub_coef[0] : T(0);
#pragma unroll
  for (int ry = 0; ry < RYT; ++ry) {
    const int y = yb + ry;
    if (y >= N) break;
    const long rowoff = (long)y * N;
    const long c = (long)x * NN + rowoff + z0;
    T lo[2], hi[2];
    hld<T>(p, (long)x * NN + rowoff + zlo, zlo_ok, lo);
    hld<T>(p, (long)x * NN + rowoff + zhi, zhi_ok, hi);
    T xn[2 * R][VEC];
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) vld<T, VEC>(p, (long)xs[k] * NN + rowoff + z0, xok[k], xn[k]);
    T vv[VEC], oo[VEC], ss[VEC];
    vld<T, VEC>(V, (long)x * NN + rowoff + z0, V != nullptr, vv);
    vld<T, VEC>(out + b * P, c, accumulate != 0, oo);
    vld<T, VEC>(sub + b * P, c, sub != nullptr, ss);
    // z line: lo[0], lo[1], centre[0..VEC-1], hi[0], hi[1]
    T zz[VEC + 4];
    zz[0] = lo[0];
    zz[1] = lo[1];
#pragma unroll
    for (int k = 0; k < VEC; ++k) zz[2 + k] = row[ry + R][k];
    zz[VEC + 2] = hi[0];
    zz[VEC + 3] = hi[1];
    T res[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
      const T ctr = row[ry + R][k];
      T s1 = zz[k + 1] + zz[k + 3] + row[ry + R - 1][k] + row[ry + R + 1][k] + xn[R - 1][k] +
             xn[R][k];
      T r = c0 * ctr + c1 * s1;
      if (R == 2) {
        const T s2 = zz[k] + zz[k + 4] + row[ry][k] + row[ry + 4][k] + xn[0][k] + xn[3][k];
        r += c2 * s2;
      }
      r += vv[k] * ctr;
      r += oo[k];
      r -= sc * ss[k];
      res[k] = NORM ? r * isc : r;
    }
    using Vt = typename Vec16<T>::type;
    Vt o;
    T* oe = reinterpret_cast<T*>(&o);
#pragma unroll
    for (int k = 0; k < VEC; ++k) oe[k] = res[k];
    *reinterpret_cast<Vt*>(out + b * P + c) = o;
    if (NORM) {
#pragma unroll
      for (int k = 0; k < VEC; ++k) nacc += (double)res[k] * (double)res[k];
    }
  }
  }  // live
  if (NORM) {
    __shared__ double red[16];
    const double t = block_sum(nacc, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) npart[blockIdx.x] = t;
  }
}

// ---------------------------------------------------------------------------
// x-marching vector kernel.  A 32 x 8 block owns the same (y, z) tile as
// stencil_vec_kernel (VEC consecutive z of RYT rows per thread) and walks XC
// consecutive x planes.  Each plane is read from memory once: a thread keeps
// the 2R+1 planes around the current one for its own rows in registers (the
// x neighbours), and the block stages the current plane with its y and z
// halos in LDS (the in-plane neighbours).  The rows, halos and V of the
// planes PF ahead are loaded PF iterations early (a ring of PF prefetch
// stages in registers), so each wave keeps PF planes of loads in flight while
// it computes; the loads a chunk does not need are predicated off.  Two block
// orders (blocks are dealt to the 8 XCDs round-robin): remap 1 orders them
// (x-chunk, z-tile, y-tile, batch) fastest first so that each XCD works
// through consecutive chunks of the same tiles (the R planes each chunk
// re-reads at its start were just read by its neighbour chunk on the same
// XCD); remap 2 gives each XCD an eighth of the spatial units for every field
// of the batch, so the shared V of a unit comes from memory once and from
// the XCD's L2 for the other fields.  Measured on 16 float fields of 128^3
// (order 4): 8-plane chunks in the field-inner order 62.2 us against 64.8 us
// for the best chunking in the other order; 4 fields 17.4 us against 24.7.
// RY = 1 gives each thread one y row in 512-thread blocks (120 instead of 156
// VGPRs, 4 waves per SIMD instead of 3) over the same 16-row tile: 4 fields
// 16.4 us against 17.3, 16 fields no faster (62.3 against 62.8 us).
// ---------------------------------------------------------------------------
template <typename T, int R, bool PER, int PF, int RY>
__global__ __launch_bounds__(VTX * VTY * RYT / RY) void stencil_march_kernel(
    const T* __restrict__ psi, const T* __restrict__ V, T* __restrict__ out,
    const T* __restrict__ sub, const double* __restrict__ sub_coef, int N, int nzt, int nyt,
    int nxc, int XC, int remap, T c0, T c1, T c2, int accumulate) {
  constexpr int VEC = Vec16<T>::n;
  constexpr int TZW = VTX * VEC;     // z extent of the tile
  constexpr int MY = VTY * RYT / RY;  // y threads: RY rows each, the same 16-row tile
  constexpr int TYR = MY * RY;        // y extent of the tile
  constexpr int LW = TZW + 2 * VEC;  // LDS row: VEC pad | tile | VEC pad (R <= VEC)
  constexpr int LR = TYR + 2 * R;    // LDS rows: the tile's rows and R halo rows each side
  static_assert(R <= VEC && 2 * R <= MY, "halo wider than the padding");
  __shared__ __align__(16) T tile[LR * LW];
  using Vt = typename Vec16<T>::type;

  const long nb = (long)gridDim.x;
  int xc, zt, yt;
  long b;
  if (remap == 2) {
    // field-inner order: XCD k owns the k-th eighth of the (y-tile, z-tile,
    // x-chunk) units, y-tile fastest, for every field, fields fastest; the
    // units' V (and the y halos of neighbouring tiles) then come from memory
    // once per XCD and from its L2 for the other fields
    const long S = (long)nxc * nzt * nyt, q = blockIdx.x / 8;
    const long nf = nb / S;
    b = q % nf;
    long u = (long)(blockIdx.x % 8) * (S / 8) + q / nf;
    yt = (int)(u % nyt);
    u /= nyt;
    zt = (int)(u % nzt);
    xc = (int)(u / nzt);
  } else {
    long lb = remap ? (long)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : (long)blockIdx.x;
    xc = (int)(lb % nxc);
    lb /= nxc;
    zt = (int)(lb % nzt);
    lb /= nzt;
    yt = (int)(lb % nyt);
    b = lb / nyt;
  }
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int tid = ty * VTX + tx;
  const int zb = zt * TZW, z0 = zb + tx * VEC;
  const int yb0 = yt * TYR, yb = yb0 + ty * RY;
  const int x0 = xc * XC;
  const int xe = x0 + XC < N ? x0 + XC : N;
  const long NN = (long)N * N, P = NN * N;
  const T* p = psi + b * P;
  T* o = out + b * P;
  const T* sb = sub != nullptr ? sub + b * P : nullptr;
  const T sc = (sub != nullptr) ? (T)sub_coef[0] : T(0);
  // own vector column: z0 .. z0+VEC-1 (wrapped for the periodic case; N % VEC == 0)
  const bool zin = z0 < N;
  const int zc = PER ? (z0 % N) : z0;
  const bool zok = PER || zin;
  // the halo row this thread stages (threads ty < 2R) and the z-halo scalar
  // (threads tid < LR * 2R)
  const int hy = ty < R ? yb0 - R + ty : yb0 + TYR + (ty - R);
  const int hrow = ty < R ? ty : R + TYR + (ty - R);
  const bool has_h = ty < 2 * R;
  const int zr = tid / (2 * R), zk = tid % (2 * R);
  const bool has_z = tid < LR * 2 * R;
  const int zy = yb0 - R + zr;
  const int zz = zk < R ? zb - R + zk : zb + TZW + (zk - R);
  const int zcol = zk < R ? VEC - R + zk : VEC + TZW + (zk - R);

  auto plane_ok = [&](int xp, int& xw) {
    xw = PER ? wrap(xp, N) : xp;
    return PER || (xp >= 0 && xp < N);
  };
  auto row_ok = [&](int y, int& yw) {
    yw = PER ? wrap(y, N) : y;
    return PER || (y >= 0 && y < N);
  };
  // need: false for a plane the chunk does not use (its registers get zeros)
  auto load_own = [&](int xp, T (&w)[RY][VEC], bool need) {
    int xw;
    const bool xo = plane_ok(xp, xw) && need;
#pragma unroll
    for (int ry = 0; ry < RY; ++ry) {
      int yw;
      const bool yo = row_ok(yb + ry, yw);
      vld<T, VEC>(p, ((long)xw * N + yw) * N + zc, xo && yo && zok, w[ry]);
    }
  };
  auto load_halo = [&](int xp, T (&hv)[VEC], T& zs, bool need) {
    int xw;
    const bool xo = plane_ok(xp, xw) && need;
    int yw;
    const bool yo = row_ok(hy, yw);
    vld<T, VEC>(p, ((long)xw * N + yw) * N + zc, has_h && xo && yo && zok, hv);
    int yzw;
    const bool yzo = row_ok(zy, yzw);
    const int zw = PER ? wrap(zz, N) : zz;
    const bool zo = PER || (zz >= 0 && zz < N);
    zs = (has_z && xo && yzo && zo) ? p[((long)xw * N + yzw) * N + zw] : T(0);
  };
  auto load_v = [&](int xp, T (&vv)[RY][VEC], bool need) {
#pragma unroll
    for (int ry = 0; ry < RY; ++ry) {
      const int y = yb + ry;
      vld<T, VEC>(V, ((long)xp * N + y) * N + z0,
                  need && V != nullptr && xp < N && y < N && zin, vv[ry]);
    }
  };

  // Register rings, indexed at compile time: the x loop is unrolled by the
  // ring length, so every load lands in the slot where it is used and no
  // register is copied while its load is in flight (a copy of such a
  // register makes the wave wait for the load, which would serialise the
  // prefetch with the computation).  Own rows: plane p in slot
  // (p - (x0 - R)) % L, L >= 2R + 1 + PF (the 2R + 1 planes around the
  // current one and PF planes ahead).  Halos and V: plane p in slot
  // (p - x0) % H, H = PF + 1 (H divides L).
  constexpr int H = PF + 1;
  constexpr int L = ((2 * R + 1 + PF + H - 1) / H) * H;
  T own[L][RY][VEC];
  T hal[H][VEC], zsr[H], vr[H][RY][VEC];
#pragma unroll
  for (int k = 0; k < 2 * R + PF; ++k) load_own(x0 - R + k, own[k], x0 - R + k <= xe - 1 + R);
#pragma unroll
  for (int d = 0; d < PF; ++d) {
    const bool need = x0 + d < xe;
    load_halo(x0 + d, hal[d], zsr[d], need);
    load_v(x0 + d, vr[d], need);
  }

  for (int xb = x0; xb < xe; xb += L) {
#pragma unroll
    for (int u = 0; u < L; ++u) {
      const int x = xb + u;
      if (x >= xe) break;  // uniform over the block
      {
        // the loads of the plane PF ahead
        const bool need = x + PF < xe;
        load_own(x + R + PF, own[(u + 2 * R + PF) % L], need);
        load_halo(x + PF, hal[(u + PF) % H], zsr[(u + PF) % H], need);
        load_v(x + PF, vr[(u + PF) % H], need);
      }
      const int cs = (u + R) % L;  // slot of plane x
      // stage plane x with its halos
#pragma unroll
      for (int ry = 0; ry < RY; ++ry) {
        Vt t;
        T* te = reinterpret_cast<T*>(&t);
#pragma unroll
        for (int k = 0; k < VEC; ++k) te[k] = own[cs][ry][k];
        *reinterpret_cast<Vt*>(&tile[(R + ty * RY + ry) * LW + VEC + tx * VEC]) = t;
      }
      if (has_h) {
        Vt t;
        T* te = reinterpret_cast<T*>(&t);
#pragma unroll
        for (int k = 0; k < VEC; ++k) te[k] = hal[u % H][k];
        *reinterpret_cast<Vt*>(&tile[hrow * LW + VEC + tx * VEC]) = t;
      }
      if (has_z) tile[zr * LW + zcol] = zsr[u % H];
      __syncthreads();
      const int m1 = (u + R - 1) % L, p1 = (u + R + 1) % L;
      const int m2 = (u + L + R - 2) % L, p2 = (u + R + 2) % L;
      // outputs of the thread's rows
#pragma unroll
      for (int ry = 0; ry < RY; ++ry) {
        const int y = yb + ry;
        const int lr = R + ty * RY + ry;
        // z line: 3 VEC-wide LDS vectors around the own column
        T zl[3 * VEC];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const Vt t = *reinterpret_cast<const Vt*>(&tile[lr * LW + (tx + q) * VEC]);
          const T* te = reinterpret_cast<const T*>(&t);
#pragma unroll
          for (int k = 0; k < VEC; ++k) zl[q * VEC + k] = te[k];
        }
        T yn[2 * R][VEC];
#pragma unroll
        for (int d = 0; d < R; ++d) {
          const Vt a = *reinterpret_cast<const Vt*>(&tile[(lr - 1 - d) * LW + VEC + tx * VEC]);
          const Vt c = *reinterpret_cast<const Vt*>(&tile[(lr + 1 + d) * LW + VEC + tx * VEC]);
          const T* ae = reinterpret_cast<const T*>(&a);
          const T* ce = reinterpret_cast<const T*>(&c);
#pragma unroll
          for (int k = 0; k < VEC; ++k) {
            yn[d][k] = ae[k];
            yn[R + d][k] = ce[k];
          }
        }
        const bool act = zin && y < N;
        const long c = ((long)x * N + y) * N + z0;
        T oo[VEC], ss[VEC];
        vld<T, VEC>(o, c, act && accumulate != 0, oo);
        vld<T, VEC>(sb, c, act && sb != nullptr, ss);
        T res[VEC];
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
          const T ctr = own[cs][ry][k];
          T s1 = zl[VEC + k - 1] + zl[VEC + k + 1] + yn[0][k] + yn[R][k] + own[m1][ry][k] +
                 own[p1][ry][k];
          T r = c0 * ctr + c1 * s1;
          if (R == 2) {
            const T s2 = zl[VEC + k - 2] + zl[VEC + k + 2] + yn[1][k] + yn[R + 1][k] +
                         own[m2][ry][k] + own[p2][ry][k];
            r += c2 * s2;
          }
          r += vr[u % H][ry][k] * ctr;
          r += oo[k];
          r -= sc * ss[k];
          res[k] = r;
        }
        if (act) {
          Vt t;
          T* te = reinterpret_cast<T*>(&t);
#pragma unroll
          for (int k = 0; k < VEC; ++k) te[k] = res[k];
          *reinterpret_cast<Vt*>(o + c) = t;
        }
      }
      __syncthreads();  // the tile is rewritten for the next plane
    }
  }
}

// ---------------------------------------------------------------------------
// The same x-marching kernel with every global access through buffer
// resources (`buffer_load_dwordx4` / `buffer_store_dwordx4`).  A resource is
// a base address and a byte range; the hardware returns zeros for a load
// whose offset lies outside the range and drops such a store.  Each plane
// gets its own resources (base + x N^2, scalar arithmetic; a range of 0 for a
// plane outside the domain), and each thread keeps 32-bit byte offsets of its
// rows within a plane, set beyond every range for rows outside the domain.
// So the loop computes no 64-bit addresses, no per-lane bounds and no
// zero-selects: in the kernel above more than half of the vector
// instructions are that bookkeeping (875 integer, select and move against 789
// floating-point instructions in the unrolled loop).  The accumulate and
// subtract inputs load under uniform branches.  The plane tile is double
// buffered in LDS, so each plane costs one barrier instead of two.  The
// arithmetic is the kernel above's, term by term, so the results are the
// same bits.
// ---------------------------------------------------------------------------
constexpr unsigned kOOB = 0x80000000u;  // beyond every plane range (N^2 * 8 < 2^31)
typedef unsigned int gato_v4u32 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           0x00020000);
}

template <typename T>
__device__ __forceinline__ void bld(__amdgpu_buffer_rsrc_t r, unsigned off, T* o) {
  if constexpr (sizeof(T) == 4) {
    const float4 v = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  } else {
    const double2 v =
        __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    o[0] = v.x; o[1] = v.y;
  }
}

template <typename T>
__device__ __forceinline__ T bld1(__amdgpu_buffer_rsrc_t r, unsigned off) {
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  else
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

template <typename T>
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, unsigned off, const T* a) {
  if constexpr (sizeof(T) == 4) {
    const float4 v = {a[0], a[1], a[2], a[3]};
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(gato_v4u32, v), r, off, 0, 0);
  } else {
    const double2 v = {a[0], a[1]};
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(gato_v4u32, v), r, off, 0, 0);
  }
}

template <typename T, int R, bool PER, int PF, int RY>
__global__ __launch_bounds__(VTX * VTY * RYT / RY) void stencil_marchb_kernel(
    const T* __restrict__ psi, const T* __restrict__ V, T* __restrict__ out,
    const T* __restrict__ sub, const double* __restrict__ sub_coef, int N, int nzt, int nyt,
    int nxc, int XC, int remap, T c0, T c1, T c2, int accumulate) {
  constexpr int VEC = Vec16<T>::n;
  constexpr int TZW = VTX * VEC;
  constexpr int MY = VTY * RYT / RY;
  constexpr int TYR = MY * RY;
  constexpr int LW = TZW + 2 * VEC;
  constexpr int LR = TYR + 2 * R;
  static_assert(R <= VEC && 2 * R <= MY, "halo wider than the padding");
  // two tiles: plane x is staged into tile[(x - x0) & 1] while slower waves
  // may still read plane x - 1 from the other, so one barrier per plane
  __shared__ __align__(16) T tiles[2][LR * LW];
  using Vt = typename Vec16<T>::type;

  const long nb = (long)gridDim.x;
  int xc, zt, yt;
  long b;
  if (remap == 2) {
    const long S = (long)nxc * nzt * nyt, q = blockIdx.x / 8;
    const long nf = nb / S;
    b = q % nf;
    long u = (long)(blockIdx.x % 8) * (S / 8) + q / nf;
    yt = (int)(u % nyt);
    u /= nyt;
    zt = (int)(u % nzt);
    xc = (int)(u / nzt);
  } else {
    long lb = remap ? (long)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : (long)blockIdx.x;
    xc = (int)(lb % nxc);
    lb /= nxc;
    zt = (int)(lb % nzt);
    lb /= nzt;
    yt = (int)(lb % nyt);
    b = lb / nyt;
  }
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int tid = ty * VTX + tx;
  const int zb = zt * TZW, z0 = zb + tx * VEC;
  const int yb0 = yt * TYR, yb = yb0 + ty * RY;
  const int x0 = xc * XC;
  const int xe = x0 + XC < N ? x0 + XC : N;
  const long NN = (long)N * N, P = NN * N;
  const unsigned PB = (unsigned)(NN * sizeof(T));  // bytes of one plane
  const T* p = psi + b * P;
  T* o = out + b * P;
  const T* sb = sub != nullptr ? sub + b * P : nullptr;
  const T sc = (sub != nullptr) ? (T)sub_coef[0] : T(0);
  const bool zin = z0 < N;
  const int zc = PER ? (z0 % N) : z0;
  const bool zok = PER || zin;
  auto row_ok = [&](int y, int& yw) {
    yw = PER ? wrap(y, N) : y;
    return PER || (y >= 0 && y < N);
  };
  auto boff = [&](bool ok, int y, int z) {
    return ok ? (unsigned)(((long)y * N + z) * (long)sizeof(T)) : kOOB;
  };
  // byte offsets within a plane, kOOB for what lies outside the domain
  unsigned own_off[RY], v_off[RY], o_off[RY];
#pragma unroll
  for (int ry = 0; ry < RY; ++ry) {
    int yw;
    const bool ok = row_ok(yb + ry, yw);
    own_off[ry] = boff(ok && zok, yw, zc);
    const int y = yb + ry;
    v_off[ry] = boff(y < N && zin, y, z0);
    o_off[ry] = boff(zin && y < N, y, z0);
  }
  const int hy = ty < R ? yb0 - R + ty : yb0 + TYR + (ty - R);
  const int hrow = ty < R ? ty : R + TYR + (ty - R);
  const bool has_h = ty < 2 * R;
  unsigned h_off;
  {
    int yw;
    const bool ok = row_ok(hy, yw);
    h_off = boff(has_h && ok && zok, yw, zc);
  }
  const int zr = tid / (2 * R), zk = tid % (2 * R);
  const bool has_z = tid < LR * 2 * R;
  const int zy = yb0 - R + zr;
  const int zz = zk < R ? zb - R + zk : zb + TZW + (zk - R);
  const int zcol = zk < R ? VEC - R + zk : VEC + TZW + (zk - R);
  unsigned zs_off;
  {
    int yzw;
    const bool yzo = row_ok(zy, yzw);
    const int zw = PER ? wrap(zz, N) : zz;
    const bool zo = PER || (zz >= 0 && zz < N);
    zs_off = boff(has_z && yzo && zo, yzw, zw);
  }
  // the resources of plane xp of a field (range 0 outside the domain or the chunk)
  auto prs = [&](const T* f, int xp, bool need) {
    const int xw = PER ? wrap(xp, N) : xp;
    const bool in = need && (PER || (xp >= 0 && xp < N));
    return plane_rsrc(f + (in ? (long)xw * NN : 0), in ? PB : 0u);
  };

  auto load_own = [&](int xp, T (&w)[RY][VEC], bool need) {
    const __amdgpu_buffer_rsrc_t r = prs(p, xp, need);
#pragma unroll
    for (int ry = 0; ry < RY; ++ry) bld<T>(r, own_off[ry], w[ry]);
  };
  auto load_halo = [&](int xp, T (&hv)[VEC], T& zs, bool need) {
    const __amdgpu_buffer_rsrc_t r = prs(p, xp, need);
    bld<T>(r, h_off, hv);
    zs = bld1<T>(r, zs_off);
  };
  auto load_v = [&](int xp, T (&vv)[RY][VEC], bool need) {
    const __amdgpu_buffer_rsrc_t r = prs(V, xp, need && V != nullptr && xp < N);
#pragma unroll
    for (int ry = 0; ry < RY; ++ry) bld<T>(r, v_off[ry], vv[ry]);
  };

  constexpr int H = PF + 1;
  constexpr int L = ((2 * R + 1 + PF + H - 1) / H) * H;
  T own[L][RY][VEC];
  T hal[H][VEC], zsr[H], vr[H][RY][VEC];
#pragma unroll
  for (int k = 0; k < 2 * R + PF; ++k) load_own(x0 - R + k, own[k], x0 - R + k <= xe - 1 + R);
#pragma unroll
  for (int d = 0; d < PF; ++d) {
    const bool need = x0 + d < xe;
    load_halo(x0 + d, hal[d], zsr[d], need);
    load_v(x0 + d, vr[d], need);
  }

  for (int xb = x0; xb < xe; xb += L) {
#pragma unroll
    for (int u = 0; u < L; ++u) {
      const int x = xb + u;
      if (x >= xe) break;  // uniform over the block
      {
        const bool need = x + PF < xe;
        load_own(x + R + PF, own[(u + 2 * R + PF) % L], need);
        load_halo(x + PF, hal[(u + PF) % H], zsr[(u + PF) % H], need);
        load_v(x + PF, vr[(u + PF) % H], need);
      }
      const int cs = (u + R) % L;
      T* tile = tiles[(x - x0) & 1];
#pragma unroll
      for (int ry = 0; ry < RY; ++ry) {
        Vt t;
        T* te = reinterpret_cast<T*>(&t);
#pragma unroll
        for (int k = 0; k < VEC; ++k) te[k] = own[cs][ry][k];
        *reinterpret_cast<Vt*>(&tile[(R + ty * RY + ry) * LW + VEC + tx * VEC]) = t;
      }
      if (has_h) {
        Vt t;
        T* te = reinterpret_cast<T*>(&t);
#pragma unroll
        for (int k = 0; k < VEC; ++k) te[k] = hal[u % H][k];
        *reinterpret_cast<Vt*>(&tile[hrow * LW + VEC + tx * VEC]) = t;
      }
      if (has_z) tile[zr * LW + zcol] = zsr[u % H];
      __syncthreads();
      const int m1 = (u + R - 1) % L, p1 = (u + R + 1) % L;
      const int m2 = (u + L + R - 2) % L, p2 = (u + R + 2) % L;
      const __amdgpu_buffer_rsrc_t ro = prs(o, x, true);
      T oo[RY][VEC], ss[RY][VEC];
      if (accumulate != 0) {
#pragma unroll
        for (int ry = 0; ry < RY; ++ry) bld<T>(ro, o_off[ry], oo[ry]);
      } else {
#pragma unroll
        for (int ry = 0; ry < RY; ++ry)
#pragma unroll
          for (int k = 0; k < VEC; ++k) oo[ry][k] = T(0);
      }
      if (sb != nullptr) {
        const __amdgpu_buffer_rsrc_t rs = prs(sb, x, true);
#pragma unroll
        for (int ry = 0; ry < RY; ++ry) bld<T>(rs, o_off[ry], ss[ry]);
      } else {
#pragma unroll
        for (int ry = 0; ry < RY; ++ry)
#pragma unroll
          for (int k = 0; k < VEC; ++k) ss[ry][k] = T(0);
      }
#pragma unroll
      for (int ry = 0; ry < RY; ++ry) {
        const int lr = R + ty * RY + ry;
        T zl[3 * VEC];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const Vt t = *reinterpret_cast<const Vt*>(&tile[lr * LW + (tx + q) * VEC]);
          const T* te = reinterpret_cast<const T*>(&t);
#pragma unroll
          for (int k = 0; k < VEC; ++k) zl[q * VEC + k] = te[k];
        }
        T yn[2 * R][VEC];
#pragma unroll
        for (int d = 0; d < R; ++d) {
          const Vt a = *reinterpret_cast<const Vt*>(&tile[(lr - 1 - d) * LW + VEC + tx * VEC]);
          const Vt c = *reinterpret_cast<const Vt*>(&tile[(lr + 1 + d) * LW + VEC + tx * VEC]);
          const T* ae = reinterpret_cast<const T*>(&a);
          const T* ce = reinterpret_cast<const T*>(&c);
#pragma unroll
          for (int k = 0; k < VEC; ++k) {
            yn[d][k] = ae[k];
            yn[R + d][k] = ce[k];
          }
        }
        T res[VEC];
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
          const T ctr = own[cs][ry][k];
          T s1 = zl[VEC + k - 1] + zl[VEC + k + 1] + yn[0][k] + yn[R][k] + own[m1][ry][k] +
                 own[p1][ry][k];
          T r = c0 * ctr + c1 * s1;
          if (R == 2) {
            const T s2 = zl[VEC + k - 2] + zl[VEC + k + 2] + yn[1][k] + yn[R + 1][k] +
                         own[m2][ry][k] + own[p2][ry][k];
            r += c2 * s2;
          }
          r += vr[u % H][ry][k] * ctr;
          r += oo[ry][k];
          r -= sc * ss[ry][k];
          res[k] = r;
        }
        bst<T>(ro, o_off[ry], res);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Barrier-free x-marching stencil.  Every wave streams its own strip of 64
// consecutive 16-byte vectors of a plane (the plane is contiguous, so a strip
// is 64 / (N / VEC) whole or partial rows) through XC planes of one field, and
// no wave ever waits for another: no LDS and no barrier.
// - x neighbours: a register ring of the thread's own vector on planes
//   x - R .. x + R + PF, each plane loaded once per chunk, PF planes ahead;
// - z neighbours: the neighbouring lanes' vectors through wave-wide DPP shifts
//   (wave_shr:1 / wave_shl:1); a lane at a row end takes 0 (Dirichlet) or the
//   row's other end (periodic, loaded), and a row cut by the strip's first or
//   last lane loads its R halo elements;
// - y neighbours: the vectors R rows above and below, loaded PF planes ahead;
//   they are the rows the neighbouring waves (or this one) stream from memory,
//   so they come from L1 / L2.
// Blocks hold blockDim.x / 64 strips.  Order 2 is the field-inner XCD split of
// the marching kernels: XCD k owns an eighth of the (chunk, strip) units, strip
// fastest, for every field, so the shared V and the y halos of adjacent strips
// come from that XCD's L2.  Same expression in the same order as the other
// variants, so the same bits.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T lane_from_left(T v) {  // lane l gets lane l - 1's value
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(
        T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
  } else {
    const int2 w = __builtin_bit_cast(int2, v);
    int2 r;
    r.x = __builtin_amdgcn_update_dpp(0, w.x, 0x138, 0xf, 0xf, true);
    r.y = __builtin_amdgcn_update_dpp(0, w.y, 0x138, 0xf, 0xf, true);
    return __builtin_bit_cast(T, r);
  }
}

template <typename T>
__device__ __forceinline__ T lane_from_right(T v) {  // lane l gets lane l + 1's value
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(
        T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, true));
  } else {
    const int2 w = __builtin_bit_cast(int2, v);
    int2 r;
    r.x = __builtin_amdgcn_update_dpp(0, w.x, 0x130, 0xf, 0xf, true);
    r.y = __builtin_amdgcn_update_dpp(0, w.y, 0x130, 0xf, 0xf, true);
    return __builtin_bit_cast(T, r);
  }
}

template <typename T, int R, bool PER, int PF>
__global__ __launch_bounds__(512) void stencil_marchr_kernel(
    const T* __restrict__ psi, const T* __restrict__ V, T* __restrict__ out,
    const T* __restrict__ sub, const double* __restrict__ sub_coef, int N, int nst, int nxc,
    int XC, int remap, T c0, T c1, T c2, int accumulate) {
  constexpr int VEC = Vec16<T>::n;
  static_assert(R <= VEC, "halo wider than one vector");
  const long nb = (long)gridDim.x;
  int xc, st;
  long b;
  if (remap == 2) {
    const long S = (long)nxc * nst, qb = blockIdx.x / 8;
    const long nf = nb / S;
    b = qb % nf;
    const long u = (long)(blockIdx.x % 8) * (S / 8) + qb / nf;
    st = (int)(u % nst);
    xc = (int)(u / nst);
  } else {
    long lb = remap ? (long)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : (long)blockIdx.x;
    st = (int)(lb % nst);
    lb /= nst;
    xc = (int)(lb % nxc);
    b = lb / nxc;
  }
  const int lane = threadIdx.x & 63;
  const long NN = (long)N * N, P = NN * N;
  const long Qv = NN / VEC;  // vectors per plane
  const long q = (long)st * blockDim.x + threadIdx.x;
  const bool act = q < Qv;
  const int e0 = act ? (int)(q * VEC) : 0;  // element of the plane
  const int y = e0 / N, z0 = e0 % N;
  const unsigned PB = (unsigned)(NN * sizeof(T));
  const T* p = psi + b * P;
  T* o = out + b * P;
  const T* sb = sub != nullptr ? sub + b * P : nullptr;
  const T sc = (sub != nullptr) ? (T)sub_coef[0] : T(0);
  const int x0 = xc * XC;
  const int xe = x0 + XC < N ? x0 + XC : N;
  auto eoff = [&](bool ok, long e) { return ok ? (unsigned)(e * (long)sizeof(T)) : kOOB; };
  const unsigned own_off = eoff(act, e0);
  // y neighbours: [d] = row y - 1 - d, [R + d] = row y + 1 + d
  unsigned y_off[2 * R];
#pragma unroll
  for (int d = 0; d < R; ++d) {
    const int ym = y - 1 - d, yp = y + 1 + d;
    const bool okm = act && (PER || ym >= 0), okp = act && (PER || yp < N);
    y_off[d] = eoff(okm, (long)(PER ? wrap(ym, N) : ym) * N + z0);
    y_off[R + d] = eoff(okp, (long)(PER ? wrap(yp, N) : yp) * N + z0);
  }
  // z halo elements this lane loads: [k] = z0 - 1 - k, [R + k] = z0 + VEC + k
  const bool row_start = z0 == 0, row_end = z0 + VEC == N;
  const bool load_l = act && ((row_start && PER) || (!row_start && lane == 0));
  const bool load_r = act && ((row_end && PER) || (!row_end && (lane == 63 || q + 1 == Qv)));
  unsigned z_off[2 * R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    z_off[k] = eoff(load_l, (long)y * N + (PER ? wrap(z0 - 1 - k, N) : z0 - 1 - k));
    z_off[R + k] = eoff(load_r, (long)y * N + (PER ? wrap(z0 + VEC + k, N) : z0 + VEC + k));
  }
  const bool zload = __any(load_l || load_r);  // uniform over the wave
  auto prs = [&](const T* f, int xp, bool need) {
    const int xw = PER ? wrap(xp, N) : xp;
    const bool in = need && (PER || (xp >= 0 && xp < N));
    return plane_rsrc(f + (in ? (long)xw * NN : 0), in ? PB : 0u);
  };

  constexpr int H = PF + 1;
  constexpr int L = ((2 * R + 1 + PF + H - 1) / H) * H;
  T own[L][VEC];
  T yv[H][2 * R][VEC], vr[H][VEC], zh[H][2 * R];
  auto load_plane = [&](int xp, int h, bool need) {
    const __amdgpu_buffer_rsrc_t r = prs(p, xp, need);
#pragma unroll
    for (int d = 0; d < 2 * R; ++d) bld<T>(r, y_off[d], yv[h][d]);
    if (zload) {
#pragma unroll
      for (int k = 0; k < 2 * R; ++k) zh[h][k] = bld1<T>(r, z_off[k]);
    } else {
#pragma unroll
      for (int k = 0; k < 2 * R; ++k) zh[h][k] = T(0);
    }
    const __amdgpu_buffer_rsrc_t rv = prs(V, xp, need && V != nullptr && xp < N);
    bld<T>(rv, own_off, vr[h]);
  };
#pragma unroll
  for (int k = 0; k < 2 * R + PF; ++k) {
    const __amdgpu_buffer_rsrc_t r = prs(p, x0 - R + k, x0 - R + k <= xe - 1 + R);
    bld<T>(r, own_off, own[k]);
  }
#pragma unroll
  for (int d = 0; d < PF; ++d) load_plane(x0 + d, d, x0 + d < xe);

  for (int xb = x0; xb < xe; xb += L) {
#pragma unroll
    for (int u = 0; u < L; ++u) {
      const int x = xb + u;
      if (x >= xe) break;  // uniform over the block
      {
        const bool need = x + PF < xe;
        const __amdgpu_buffer_rsrc_t r = prs(p, x + R + PF, need);
        bld<T>(r, own_off, own[(u + 2 * R + PF) % L]);
        load_plane(x + PF, (u + PF) % H, need);
      }
      const int cs = (u + R) % L, h = u % H;
      const int m1 = (u + R - 1) % L, p1 = (u + R + 1) % L;
      const int m2 = (u + L + R - 2) % L, p2 = (u + R + 2) % L;
      // z row of this lane: zz[R + k] = element z0 + k, halos on both sides
      T zz[VEC + 2 * R];
#pragma unroll
      for (int k = 0; k < VEC; ++k) zz[R + k] = own[cs][k];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const T fl = lane_from_left(own[cs][VEC - 1 - k]);
        const T fr = lane_from_right(own[cs][k]);
        zz[R - 1 - k] = (row_start || lane == 0) ? zh[h][k] : fl;
        zz[R + VEC + k] = (row_end || lane == 63) ? zh[h][R + k] : fr;
      }
      const __amdgpu_buffer_rsrc_t ro = prs(o, x, true);
      T oo[VEC], ss[VEC];
      if (accumulate != 0) {
        bld<T>(ro, own_off, oo);
      } else {
#pragma unroll
        for (int k = 0; k < VEC; ++k) oo[k] = T(0);
      }
      if (sb != nullptr) {
        const __amdgpu_buffer_rsrc_t rs = prs(sb, x, true);
        bld<T>(rs, own_off, ss);
      } else {
#pragma unroll
        for (int k = 0; k < VEC; ++k) ss[k] = T(0);
      }
      T res[VEC];
#pragma unroll
      for (int k = 0; k < VEC; ++k) {
        const T ctr = own[cs][k];
        T s1 = zz[R + k - 1] + zz[R + k + 1] + yv[h][0][k] + yv[h][R][k] + own[m1][k] +
               own[p1][k];
        T r = c0 * ctr + c1 * s1;
        if (R == 2) {
          const T s2 = zz[R + k - 2] + zz[R + k + 2] + yv[h][1][k] + yv[h][R + 1][k] +
                       own[m2][k] + own[p2][k];
          r += c2 * s2;
        }
        r += vr[h][k] * ctr;
        r += oo[k];
        r -= sc * ss[k];
        res[k] = r;
      }
      bst<T>(ro, own_off, res);
    }
  }
}

// ---------------------------------------------------------------------------
// Variant choice for the vector path.  A variant is 0 (stencil_vec_kernel) or
// XC + 100 (PF - 1) + 1000 order + 10000 y + 100000 kind with XC > 0:
// XC-plane chunks with PF planes of prefetch in either block order, kind 0
// (stencil_march_kernel) or 1 (stencil_marchb_kernel) with y = 1 for one row
// per thread, or kind 2 (stencil_marchr_kernel) with y = 0 / 1 / 2 for strips
// of 256 / 512 / 128 vectors.  All of them evaluate the
// same expression in the same order.  The first call of a shape (dtype, N,
// batch, order, boundary, and which of V / accumulate / sub are present)
// times every candidate on scratch fields on the caller's stream and keeps
// the fastest for the process; calls under stream capture, or with tuning
// switched off, take the measured rule in static_variant().
// ---------------------------------------------------------------------------
std::atomic<int> g_autotune{1};
std::atomic<int> g_forced_variant{-1};
std::mutex g_tune_mu;
std::map<std::array<int, 8>, int> g_tuned;

// Large batches: the x-marching kernel with 32-plane chunks, when they still
// give >= 512 blocks (2 per CU).  Measured on 128^3 float (order 4, GB/s of
// algorithmic traffic): 16 fields 4171 against 3390 for the one-plane kernel;
// shorter chunks (2-16 planes) 3370-3670.  Single fields and small batches
// keep the one-plane kernel: there the chunks get short and their R-plane
// re-reads and serial plane chain cost more than the marching saves (1 field:
// 9.8 us at 2-plane chunks against 10.6 us; 4 fields: 20.0 us against 17.7 us).
int static_variant(long tiles, int N) {
  return tiles * ((N + 31) / 32) >= 512 ? 32 : 0;
}

template <typename T, int R, bool PER>
int run_variant(int variant, int N, int batch, int nzt, int nyt, T c0, T c1, T c2,
                const void* psi, const void* V, int accumulate, const void* sub,
                const double* sub_coef, void* out, hipStream_t s) {
  if (variant > 0) {
    const int XC = variant % 100, PF = variant % 1000 / 100 + 1, order = variant % 10000 / 1000;
    const int ry1 = variant % 100000 / 10000;  // 1: one row per thread (512-thread blocks)
    const int kb = variant / 100000;  // 1: buffer-resource loads and stores; 2: barrier-free
    GATO_REQUIRE(XC > 0 && PF >= 1 && PF <= (kb == 2 ? 2 : 3) && order <= 1 &&
                     ry1 <= (kb == 2 ? 2 : 1) && kb <= 2,
                 GATO_ERR_VALUE, "unknown stencil variant %d", variant);
    GATO_REQUIRE(kb == 0 || (long)N * N * (long)sizeof(T) < 0x80000000L, GATO_ERR_UNSUPPORTED,
                 "stencil variant %d needs planes under 2 GiB", variant);
    if (kb == 2) {
      // strips of 256 (ry1 = 0), 512 (1) or 128 (2) vectors per block
      const int nthr = ry1 == 0 ? 256 : ry1 == 1 ? 512 : 128;
      constexpr int VEC = Vec16<T>::n;
      const long Qv = (long)N * N / VEC;
      const int nst = (int)((Qv + nthr - 1) / nthr);
      const int nxc = (N + XC - 1) / XC;
      const long nbr = (long)nst * nxc * batch;
      GATO_REQUIRE(nbr <= 0x7fffffffL, GATO_ERR_UNSUPPORTED, "stencil grid too large");
      const int remap = (order == 1 && ((long)nst * nxc) % 8 == 0) ? 2 : (nbr % 8 == 0) ? 1 : 0;
      if (PF == 1)
        hipLaunchKernelGGL((stencil_marchr_kernel<T, R, PER, 1>), dim3((unsigned)nbr), dim3(nthr),
                           0, s, (const T*)psi, (const T*)V, (T*)out, (const T*)sub, sub_coef, N,
                           nst, nxc, XC, remap, c0, c1, c2, accumulate);
      else
        hipLaunchKernelGGL((stencil_marchr_kernel<T, R, PER, 2>), dim3((unsigned)nbr), dim3(nthr),
                           0, s, (const T*)psi, (const T*)V, (T*)out, (const T*)sub, sub_coef, N,
                           nst, nxc, XC, remap, c0, c1, c2, accumulate);
      GATO_LAUNCH_CHECK();
      return GATO_OK;
    }
    const int nxc = (N + XC - 1) / XC;
    const long nbm = (long)nzt * nyt * batch * nxc;
    GATO_REQUIRE(nbm <= 0x7fffffffL, GATO_ERR_UNSUPPORTED, "stencil grid too large");
    // order 1: the field-inner XCD split when the spatial units divide by 8
    const int remap = (order == 1 && ((long)nzt * nyt * nxc) % 8 == 0) ? 2
                      : (nbm % 8 == 0) ? 1 : 0;
    auto launch = [&](auto pf) {
      constexpr int P = decltype(pf)::value;
      if (kb && ry1)
        hipLaunchKernelGGL((stencil_marchb_kernel<T, R, PER, P, 1>), dim3((unsigned)nbm),
                           dim3(VTX, VTY * RYT), 0, s, (const T*)psi, (const T*)V, (T*)out,
                           (const T*)sub, sub_coef, N, nzt, nyt, nxc, XC, remap, c0, c1, c2,
                           accumulate);
      else if (kb)
        hipLaunchKernelGGL((stencil_marchb_kernel<T, R, PER, P, RYT>), dim3((unsigned)nbm),
                           dim3(VTX, VTY), 0, s, (const T*)psi, (const T*)V, (T*)out,
                           (const T*)sub, sub_coef, N, nzt, nyt, nxc, XC, remap, c0, c1, c2,
                           accumulate);
      else if (ry1)
        hipLaunchKernelGGL((stencil_march_kernel<T, R, PER, P, 1>), dim3((unsigned)nbm),
                           dim3(VTX, VTY * RYT), 0, s, (const T*)psi, (const T*)V, (T*)out,
                           (const T*)sub, sub_coef, N, nzt, nyt, nxc, XC, remap, c0, c1, c2,
                           accumulate);
      else
        hipLaunchKernelGGL((stencil_march_kernel<T, R, PER, P, RYT>), dim3((unsigned)nbm),
                           dim3(VTX, VTY), 0, s, (const T*)psi, (const T*)V, (T*)out,
                           (const T*)sub, sub_coef, N, nzt, nyt, nxc, XC, remap, c0, c1, c2,
                           accumulate);
    };
    if (PF == 1) launch(std::integral_constant<int, 1>{});
    else if (PF == 2) launch(std::integral_constant<int, 2>{});
    else launch(std::integral_constant<int, 3>{});
    GATO_LAUNCH_CHECK();
    return GATO_OK;
  }
  const long nb = (long)nzt * nyt * N * batch;
  GATO_REQUIRE(nb <= 0x7fffffffL, GATO_ERR_UNSUPPORTED, "stencil grid too large");
  const int remap = (nb % 8 == 0) ? 1 : 0;
  hipLaunchKernelGGL((stencil_vec_kernel<T, R, PER>), dim3((unsigned)nb), dim3(VTX, VTY), 0, s,
                     (const T*)psi, (const T*)V, (T*)out, (const T*)sub, sub_coef, N, nzt, nyt,
                     remap, c0, c1, c2, accumulate);
  GATO_LAUNCH_CHECK();
  return GATO_OK;
}

// Times the candidates (the one-plane kernel; LDS-tiled marching with chunks
// of 4-32 planes, 1-3 planes of prefetch, either block order and one or two y
// rows per thread; barrier-free wave strips of 128-512 vectors with 16- or
// 32-plane chunks; those that give >= 256 blocks) on
// zero-filled scratch fields of the caller's shape:
// one warm-up and ten timed launches each.  Returns -1 if anything fails
// (the caller then takes the static rule).
template <typename T, int R, bool PER>
int tune_variant(int N, int batch, int nzt, int nyt, bool hasV, int accumulate, bool hasSub,
                 hipStream_t s) {
  const size_t field = (size_t)N * N * N * sizeof(T);
  void *in = nullptr, *o = nullptr, *v = nullptr, *sb = nullptr, *coef = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int best = -1;
  bool ok = hipMalloc(&in, field * batch) == hipSuccess &&
            hipMalloc(&o, field * batch) == hipSuccess &&
            (!hasV || hipMalloc(&v, field) == hipSuccess) &&
            (!hasSub || (hipMalloc(&sb, field * batch) == hipSuccess &&
                         hipMalloc(&coef, sizeof(double)) == hipSuccess)) &&
            hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
  if (ok) {
    ok = hipMemsetAsync(in, 0, field * batch, s) == hipSuccess &&
         hipMemsetAsync(o, 0, field * batch, s) == hipSuccess &&
         (!v || hipMemsetAsync(v, 0, field, s) == hipSuccess) &&
         (!sb || (hipMemsetAsync(sb, 0, field * batch, s) == hipSuccess &&
                  hipMemsetAsync(coef, 0, sizeof(double), s) == hipSuccess));
  }
  const long tiles = (long)nzt * nyt * batch;
  float best_ms = 0.f;
  const int cand[] = {0,      4,      8,      16,     32,     104,    108,    116,    132,
                      208,    216,    232,    264,    1004,   1008,   1016,   1032,   1108,
                      1116,   1132,   1216,   1232,   10016,  10032,  11004,  11008,  11016,
                      11032,  100008, 100032, 101008, 101016, 110008, 110016, 110032, 111004,
                      111008, 111016, 111032, 110108, 110132, 111108, 111116, 111208,
                      200032, 201016, 201032, 211016, 211032, 220032, 221016, 221032};
  constexpr int ncand = (int)(sizeof(cand) / sizeof(cand[0]));
  for (int c = 0; ok && c < ncand; ++c) {
    const int var = cand[c];
    const int xc = var % 100;
    if (var > 0 && tiles * ((N + xc - 1) / xc) < 256) continue;
    auto launch = [&]() {
      return run_variant<T, R, PER>(var, N, batch, nzt, nyt, T(1), T(1), T(1), in, v, accumulate,
                                    sb, (const double*)coef, o, s);
    };
    ok = launch() == GATO_OK && hipEventRecord(e0, s) == hipSuccess;
    for (int r = 0; ok && r < 10; ++r) ok = launch() == GATO_OK;
    float ms = 0.f;
    ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
         hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    if (ok && (best < 0 || ms < best_ms)) {
      best = var;
      best_ms = ms;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  for (void* q : {in, o, v, sb, coef})
    if (q) (void)hipFree(q);
  if (!ok) {
    (void)hipGetLastError();
    clear_ok();
    return -1;
  }
  return best;
}

template <typename T, int R, bool PER>
int choose_variant(int N, int batch, int nzt, int nyt, const void* V, int accumulate,
                   const void* sub, hipStream_t s) {
  const int forced = g_forced_variant.load();
  if (forced >= 0) return forced;
  const long tiles = (long)nzt * nyt * batch;
  const int fallback = static_variant(tiles, N);
  if (!g_autotune.load()) return fallback;
  const std::array<int, 8> key = {(int)sizeof(T), N, batch, R, PER ? 1 : 0, V != nullptr,
                                  accumulate != 0, sub != nullptr};
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) return it->second;
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) {
    (void)hipGetLastError();
    return fallback;
  }
  int v = tune_variant<T, R, PER>(N, batch, nzt, nyt, V != nullptr, accumulate, sub != nullptr, s);
  if (v < 0) v = fallback;
  std::lock_guard<std::mutex> g(g_tune_mu);
  return g_tuned.emplace(key, v).first->second;  // the first tuning of a shape wins
}

// One field: out = stencil(psi) with the block partials of sum out^2 written
// to the workspace, then one launch that sums them (fixed order) and scales
// out by 1 / sqrt(dV sum).  Returns GATO_ERR_UNSUPPORTED when the field does
// not take the vector path (the caller then runs the two steps separately).
template <typename T, int R, bool PER>
int launch_stencil_normalize(int N, double h, double c_lap, const void* psi, const void* V,
                             const void* sub, const double* sub_coef, void* out, double dV,
                             double* norm_out, void* ws, hipStream_t s) {
  double nb1, nb2, ctr;
  if (R == 1) {
    nb1 = 1.0; nb2 = 0.0; ctr = 3.0 * -2.0;
  } else {
    nb1 = 16.0 / 12.0; nb2 = -1.0 / 12.0; ctr = 3.0 * (-30.0 / 12.0);
  }
  const double sc = c_lap / (h * h);
  constexpr int VEC = Vec16<T>::n;
  const bool aligned = ((uintptr_t)psi % 16 == 0) && ((uintptr_t)out % 16 == 0) &&
                       ((uintptr_t)V % 16 == 0) && ((uintptr_t)sub % 16 == 0);
  const int nzt = (N / VEC + VTX - 1) / VTX;
  const int nyt = (N + VTY * RYT - 1) / (VTY * RYT);
  const long nb = (long)nzt * nyt * N;
  if (N % VEC != 0 || !aligned || nb > kPartialsCap) return GATO_ERR_UNSUPPORTED;
  const int remap = (nb % 8 == 0) ? 1 : 0;
  hipLaunchKernelGGL((stencil_vec_kernel<T, R, PER, true>), dim3((unsigned)nb), dim3(VTX, VTY), 0,
                     s, (const T*)psi, (const T*)V, (T*)out, (const T*)sub, sub_coef, N, nzt, nyt,
                     remap, (T)(sc * ctr), (T)(sc * nb1), (T)(sc * nb2), 0, ws_partials(ws));
  GATO_LAUNCH_CHECK();
  const long n = (long)N * N * N;
  return scale_by_partials(sizeof(T) == 4 ? GATO_F32 : GATO_F64, n, out, ws, nb, dV,
                           norm_out ? norm_out : ws_stage(ws), out, s);
}

// n_steps Euler steps of imaginary time, psi <- normalise(psi + c_lap lap(psi)
// + Vdt psi), between the buffers a and b (step k reads a when k is even):
// one launch per step.  Step k scales its result by the inverse norm of its
// input, summed from the partials step k-1 wrote (the two halves of the
// partials area alternate), and writes the partials of its own result; one
// final pass normalises the last result.  Because the step is linear, step k
// then writes exactly the unnormalised update of the normalised state.  The
// +psi term enters as +1 on the centre weight.  The input is taken as
// normalised (step 0 applies no scale).
template <typename T, int R, bool PER>
int launch_imag_time(int N, double h, double c_lap, const void* Vdt, void* a, void* b,
                     int n_steps, double dV, double* norm_out, void* ws, hipStream_t s) {
  double nb1, nb2, ctr;
  if (R == 1) {
    nb1 = 1.0; nb2 = 0.0; ctr = 3.0 * -2.0;
  } else {
    nb1 = 16.0 / 12.0; nb2 = -1.0 / 12.0; ctr = 3.0 * (-30.0 / 12.0);
  }
  const double sc = c_lap / (h * h);
  constexpr int VEC = Vec16<T>::n;
  const bool aligned = ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0) &&
                       ((uintptr_t)Vdt % 16 == 0);
  const int nzt = (N / VEC + VTX - 1) / VTX;
  const int nyt = (N + VTY * RYT - 1) / (VTY * RYT);
  const long nb = (long)nzt * nyt * N;
  if (N % VEC != 0 || !aligned || 2 * nb > kPartialsCap) return GATO_ERR_UNSUPPORTED;
  const int remap = (nb % 8 == 0) ? 1 : 0;
  double* part[2] = {ws_partials(ws), ws_partials(ws) + kPartialsCap / 2};
  T* src = static_cast<T*>(a);
  T* dst = static_cast<T*>(b);
  for (int k = 0; k < n_steps; ++k) {
    hipLaunchKernelGGL((stencil_vec_kernel<T, R, PER, true>), dim3((unsigned)nb), dim3(VTX, VTY),
                       0, s, src, (const T*)Vdt, dst, (const T*)nullptr, (const double*)nullptr, N,
                       nzt, nyt, remap, (T)(sc * ctr + 1.0), (T)(sc * nb1), (T)(sc * nb2), 0,
                       part[k & 1], k > 0 ? part[(k - 1) & 1] : (const double*)nullptr, nb, dV);
    GATO_LAUNCH_CHECK();
    T* t = src;
    src = dst;
    dst = t;
  }
  return scale_by_partials_at(sizeof(T) == 4 ? GATO_F32 : GATO_F64, (long)N * N * N, src,
                              part[(n_steps - 1) & 1], nb, dV,
                              norm_out ? norm_out : ws_stage(ws), src, s);
}

__global__ void set_double_kernel(double* p, double v) { p[0] = v; }

template <typename T, int R, bool PER>
int launch_stencil(int N, int batch, double h, double c_lap, const void* psi, const void* V,
                   int accumulate, const void* sub, const double* sub_coef, void* out,
                   hipStream_t s) {
  double nb1, nb2, ctr;
  if (R == 1) {
    nb1 = 1.0; nb2 = 0.0; ctr = 3.0 * -2.0;           // operators.py:30-32
  } else {
    nb1 = 16.0 / 12.0; nb2 = -1.0 / 12.0; ctr = 3.0 * (-30.0 / 12.0);  // operators.py:34-37
  }
  const double sc = c_lap / (h * h);
  constexpr int VEC = Vec16<T>::n;
  const bool aligned = ((uintptr_t)psi % 16 == 0) && ((uintptr_t)out % 16 == 0) &&
                       ((uintptr_t)V % 16 == 0) && ((uintptr_t)sub % 16 == 0);
  if (N % VEC == 0 && aligned) {
    const int nzt = (N / VEC + VTX - 1) / VTX;
    const int nyt = (N + VTY * RYT - 1) / (VTY * RYT);
    const int variant = choose_variant<T, R, PER>(N, batch, nzt, nyt, V, accumulate, sub, s);
    return run_variant<T, R, PER>(variant, N, batch, nzt, nyt, (T)(sc * ctr), (T)(sc * nb1),
                                  (T)(sc * nb2), psi, V, accumulate, sub, sub_coef, out, s);
  }
  // Chunk of x-planes per block: enough blocks to fill 256 CUs several times.
  const long tiles = (long)((N + TZ - 1) / TZ) * ((N + TY - 1) / TY) * batch;
  int XC = 32;
  while (XC > 8 && tiles * ((N + XC - 1) / XC) < 2048) XC /= 2;
  const int nxc = (N + XC - 1) / XC;
  dim3 grid((N + TZ - 1) / TZ, (N + TY - 1) / TY, nxc * batch);
  dim3 block(TZ, TY, 1);
  hipLaunchKernelGGL((stencil_kernel<T, R, PER>), grid, block, 0, s, (const T*)psi, (const T*)V,
                     (T*)out, (const T*)sub, sub_coef, N, XC, nxc, (T)(sc * ctr), (T)(sc * nb1),
                     (T)(sc * nb2), accumulate);
  GATO_LAUNCH_CHECK();
  return GATO_OK;
}

template <typename T>
int stencil_dispatch(int N, int batch, double h, int order, int boundary, double c_lap,
                     const void* psi, const void* V, int accumulate, const void* sub,
                     const double* sub_coef, void* out, hipStream_t s) {
  if (order == 2) {
    if (boundary == GATO_PERIODIC)
      return launch_stencil<T, 1, true>(N, batch, h, c_lap, psi, V, accumulate, sub, sub_coef, out, s);
    return launch_stencil<T, 1, false>(N, batch, h, c_lap, psi, V, accumulate, sub, sub_coef, out, s);
  }
  if (boundary == GATO_PERIODIC)
    return launch_stencil<T, 2, true>(N, batch, h, c_lap, psi, V, accumulate, sub, sub_coef, out, s);
  return launch_stencil<T, 2, false>(N, batch, h, c_lap, psi, V, accumulate, sub, sub_coef, out, s);
}

}  // namespace

int stencil_normalize(int dtype, int N, double h, int order, int boundary, double c_lap,
                      const void* psi, const void* V, const void* sub, const double* sub_coef,
                      void* out, double dV, double* norm_out, void* ws, hipStream_t s) {
  GATO_REQUIRE(order == 2 || order == 4, GATO_ERR_VALUE,
               "Unknown order %d; expected 2 or 4.", order);
  GATO_REQUIRE(boundary == GATO_DIRICHLET || boundary == GATO_PERIODIC, GATO_ERR_VALUE,
               "Unknown boundary %d; expected 'dirichlet' or 'periodic'.", boundary);
  GATO_REQUIRE(N >= 1 && psi != nullptr && out != nullptr && ws != nullptr && dV > 0,
               GATO_ERR_VALUE, "bad arguments");
  GATO_REQUIRE(!(boundary == GATO_PERIODIC && order == 4 && N < 2), GATO_ERR_VALUE,
               "periodic order-4 stencil needs N >= 2");
  GATO_REQUIRE(sub == nullptr || sub_coef != nullptr, GATO_ERR_VALUE, "sub without sub_coef");
  GATO_REQUIRE(dtype == GATO_F32 || dtype == GATO_F64, GATO_ERR_VALUE, "unknown dtype %d", dtype);
  const bool per = boundary == GATO_PERIODIC, f32 = dtype == GATO_F32;
  int st;
  if (order == 2)
    st = per ? (f32 ? launch_stencil_normalize<float, 1, true>(N, h, c_lap, psi, V, sub, sub_coef, out, dV, norm_out, ws, s)
                    : launch_stencil_normalize<double, 1, true>(N, h, c_lap, psi, V, sub, sub_coef, out, dV, norm_out, ws, s))
             : (f32 ? launch_stencil_normalize<float, 1, false>(N, h, c_lap, psi, V, sub, sub_coef, out, dV, norm_out, ws, s)
                    : launch_stencil_normalize<double, 1, false>(N, h, c_lap, psi, V, sub, sub_coef, out, dV, norm_out, ws, s));
  else
    st = per ? (f32 ? launch_stencil_normalize<float, 2, true>(N, h, c_lap, psi, V, sub, sub_coef, out, dV, norm_out, ws, s)
                    : launch_stencil_normalize<double, 2, true>(N, h, c_lap, psi, V, sub, sub_coef, out, dV, norm_out, ws, s))
             : (f32 ? launch_stencil_normalize<float, 2, false>(N, h, c_lap, psi, V, sub, sub_coef, out, dV, norm_out, ws, s)
                    : launch_stencil_normalize<double, 2, false>(N, h, c_lap, psi, V, sub, sub_coef, out, dV, norm_out, ws, s));
  if (st != GATO_ERR_UNSUPPORTED) return st;
  clear_ok();
  st = stencil_apply(dtype, N, 1, h, order, boundary, c_lap, psi, V, 0, sub, sub_coef, out, s);
  if (st) return st;
  return normalize(dtype, (long)N * N * N, out, dV, norm_out, out, ws, s);
}

int imag_time_steps(int dtype, int N, double h, int order, int boundary, double dt,
                    const void* Vdt, void* a, void* b, int n_steps, double dV, double* norm_out,
                    void* ws, hipStream_t s) {
  GATO_REQUIRE(order == 2 || order == 4, GATO_ERR_VALUE,
               "Unknown order %d; expected 2 or 4.", order);
  GATO_REQUIRE(boundary == GATO_DIRICHLET || boundary == GATO_PERIODIC, GATO_ERR_VALUE,
               "Unknown boundary %d; expected 'dirichlet' or 'periodic'.", boundary);
  GATO_REQUIRE(N >= 1 && n_steps >= 0 && a != nullptr && b != nullptr && a != b &&
                   ws != nullptr && dV > 0,
               GATO_ERR_VALUE, "bad arguments");
  GATO_REQUIRE(!(boundary == GATO_PERIODIC && order == 4 && N < 2), GATO_ERR_VALUE,
               "periodic order-4 stencil needs N >= 2");
  GATO_REQUIRE(dtype == GATO_F32 || dtype == GATO_F64, GATO_ERR_VALUE, "unknown dtype %d", dtype);
  if (n_steps == 0) return GATO_OK;
  const double c_lap = 0.5 * dt;
  const bool per = boundary == GATO_PERIODIC, f32 = dtype == GATO_F32;
#define GATO_IT(T, R, P) launch_imag_time<T, R, P>(N, h, c_lap, Vdt, a, b, n_steps, dV, norm_out, ws, s)
  int st;
  if (order == 2)
    st = per ? (f32 ? GATO_IT(float, 1, true) : GATO_IT(double, 1, true))
             : (f32 ? GATO_IT(float, 1, false) : GATO_IT(double, 1, false));
  else
    st = per ? (f32 ? GATO_IT(float, 2, true) : GATO_IT(double, 2, true))
             : (f32 ? GATO_IT(float, 2, false) : GATO_IT(double, 2, false));
#undef GATO_IT
  if (st != GATO_ERR_UNSUPPORTED) return st;
  // other fields: the stencil (+psi through the subtract operand with
  // coefficient -1, held in the staging area) and gato_normalize per step
  clear_ok();
  double* minus_one = ws_stage(ws) + 1;
  hipLaunchKernelGGL(set_double_kernel, dim3(1), dim3(1), 0, s, minus_one, -1.0);
  GATO_LAUNCH_CHECK();
  const long n = (long)N * N * N;
  void* src = a;
  void* dst = b;
  for (int k = 0; k < n_steps; ++k) {
    st = stencil_apply(dtype, N, 1, h, order, boundary, c_lap, src, Vdt, 0, src, minus_one, dst, s);
    if (!st) st = normalize(dtype, n, dst, dV, norm_out, dst, ws, s);
    if (st) return st;
    void* t = src;
    src = dst;
    dst = t;
  }
  return GATO_OK;
}

int stencil_apply(int dtype, int N, int batch, double h, int order, int boundary, double c_lap,
                  const void* psi, const void* V, int accumulate, const void* sub,
                  const double* sub_coef, void* out, hipStream_t s) {
  GATO_REQUIRE(order == 2 || order == 4, GATO_ERR_VALUE,
               "Unknown order %d; expected 2 or 4.", order);
  GATO_REQUIRE(boundary == GATO_DIRICHLET || boundary == GATO_PERIODIC, GATO_ERR_VALUE,
               "Unknown boundary %d; expected 'dirichlet' or 'periodic'.", boundary);
  GATO_REQUIRE(N >= 1 && batch >= 1, GATO_ERR_VALUE, "bad grid N=%d batch=%d", N, batch);
  GATO_REQUIRE(!(boundary == GATO_PERIODIC && order == 4 && N < 2), GATO_ERR_VALUE,
               "periodic order-4 stencil needs N >= 2");
  GATO_REQUIRE(psi != nullptr && out != nullptr, GATO_ERR_VALUE, "null field pointer");
  GATO_REQUIRE(sub == nullptr || sub_coef != nullptr, GATO_ERR_VALUE, "sub without sub_coef");
  if (dtype == GATO_F32)
    return stencil_dispatch<float>(N, batch, h, order, boundary, c_lap, psi, V, accumulate, sub,
                                   sub_coef, out, s);
  if (dtype == GATO_F64)
    return stencil_dispatch<double>(N, batch, h, order, boundary, c_lap, psi, V, accumulate, sub,
                                    sub_coef, out, s);
  set_error(GATO_ERR_VALUE, "unknown dtype %d", dtype);
  return GATO_ERR_VALUE;
}

bool autotune_on() { return g_autotune.load() != 0; }

}  // namespace gato

extern "C" void gato_set_autotune(int on) { gato::g_autotune.store(on != 0); }

extern "C" void gato_stencil_force_variant(int variant) {
  gato::g_forced_variant.store(variant);
}

extern "C" int gato_stencil_normalize(int dtype, int N, double h, int order, int boundary,
                                      double c_lap, const void* psi, const void* V,
                                      const void* sub, const double* sub_coef, void* out,
                                      double dV, double* norm_out, void* ws, void* stream) {
  gato::clear_ok();
  return gato::stencil_normalize(dtype, N, h, order, boundary, c_lap, psi, V, sub, sub_coef, out,
                                 dV, norm_out, ws, gato::as_stream(stream));
}

extern "C" int gato_imag_time_steps(int dtype, int N, double h, int order, int boundary,
                                    double dt, const void* Vdt, void* a, void* b, int n_steps,
                                    double dV, double* norm_out, void* ws, void* stream) {
  gato::clear_ok();
  return gato::imag_time_steps(dtype, N, h, order, boundary, dt, Vdt, a, b, n_steps, dV, norm_out,
                               ws, gato::as_stream(stream));
}

extern "C" int gato_stencil_apply(int dtype, int N, int batch, double h, int order, int boundary,
                                  double c_lap, const void* psi, const void* V, int accumulate,
                                  const void* sub, const double* sub_coef, void* out,
                                  void* stream) {
  gato::clear_ok();
  return gato::stencil_apply(dtype, N, batch, h, order, boundary, c_lap, psi, V, accumulate, sub,
                             sub_coef, out, gato::as_stream(stream));
}
