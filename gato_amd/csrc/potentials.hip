// External potentials on the grid and the HGH separable nonlocal part, with
// their Hellmann-Feynman position derivatives.
//
// Reference:
//   potentials.py:16-71          softened Coulomb, multi-centre sum, harmonic
//   pseudopotentials.py:201-253  HGH local part (erf-screened Coulomb + Gaussian
//                                polynomial, r = sqrt(r^2 + 1e-30))
//   pseudopotentials.py:265-383  projectors p_li(r) Y_lm and their separable apply
//   water.py:181-241             the force: jax.grad of the BO energy at fixed
//                                orbitals (here written out analytically)
//
// Every value is evaluated in double and stored as the field type.  The
// nonlocal projectors are sampled once per geometry on a per-atom box that
// holds every point where a projector exceeds cutoff_tol times its peak; the
// apply touches only that box: one kernel per atom for boxes of at most 8192
// points, else an overlap pass and a rank-n update that sums the overlap
// partials itself.  Gradients recompute the projector derivatives on the
// fly instead of storing them.
#include <cmath>
#include <vector>

#include "common.hpp"

namespace gato {
namespace {

constexpr int kThr = 256;
constexpr int kMaxCenters = 32;
constexpr int kMaxCh = 8;

struct Centers {
  int n;
  double p[kMaxCenters][GATO_CENTER_PARAMS];
};

__device__ __forceinline__ double center_value(const double* c, double dx, double dy, double dz) {
  const int kind = (int)c[0];
  const double r2 = dx * dx + dy * dy + dz * dz;
  if (kind == GATO_CENTER_SOFT_COULOMB) {
    return -c[4] / sqrt(r2 + c[5] * c[5]);
  } else if (kind == GATO_CENTER_HGH_LOCAL) {
    const double Z = c[4], rl = c[5];
    const double r = sqrt(r2 + 1e-30);
    const double erf_term = -Z * erf(r / (1.4142135623730951 * rl)) / r;
    const double rho2 = (r / rl) * (r / rl);
    const double poly = c[6] + c[7] * rho2 + c[8] * rho2 * rho2 + c[9] * rho2 * rho2 * rho2;
    return erf_term + exp(-0.5 * rho2) * poly;
  } else if (kind == GATO_CENTER_SLATER_1S) {
    const double Z = c[4], sf = c[5];
    return Z * sqrt(Z) * exp(-Z * sqrt(r2 + sf * sf));
  } else {
    const double w = c[4];
    return 0.5 * w * w * r2;
  }
}

// (1/d) dV/dd so that dV/dR_a = -(1/d dV/dd) * (r - R)_a
__device__ __forceinline__ double center_dvdr_over_r(const double* c, double dx, double dy, double dz) {
  const int kind = (int)c[0];
  const double r2 = dx * dx + dy * dy + dz * dz;
  if (kind == GATO_CENTER_SOFT_COULOMB) {
    const double q = r2 + c[5] * c[5];
    return c[4] / (q * sqrt(q));
  } else if (kind == GATO_CENTER_HGH_LOCAL) {
    const double Z = c[4], rl = c[5];
    const double r = sqrt(r2 + 1e-30);
    const double a = 1.4142135623730951 * rl;
    const double e = erf(r / a);
    const double de = 1.1283791670955126 * exp(-(r / a) * (r / a)) / a;
    const double d_erf = -Z * (de / r - e / (r * r));
    const double rho2 = (r / rl) * (r / rl);
    const double poly = c[6] + c[7] * rho2 + c[8] * rho2 * rho2 + c[9] * rho2 * rho2 * rho2;
    const double dpoly = c[7] + 2.0 * c[8] * rho2 + 3.0 * c[9] * rho2 * rho2;
    const double drho2 = 2.0 * r / (rl * rl);
    const double d_poly = exp(-0.5 * rho2) * (-0.5 * drho2 * poly + dpoly * drho2);
    return (d_erf + d_poly) / r;
  } else if (kind == GATO_CENTER_SLATER_1S) {
    const double Z = c[4], q = sqrt(r2 + c[5] * c[5]);
    return -Z * Z * sqrt(Z) * exp(-Z * q) / q;
  } else {
    return c[4] * c[4];
  }
}

template <typename T>
__global__ void local_potential_kernel(int N, double L, Centers cs, int accumulate, T* __restrict__ V) {
  const double h = L / N;
  const long n = (long)N * N * N;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int z = (int)(i % N), y = (int)((i / N) % N), x = (int)(i / ((long)N * N));
    const double px = -L / 2 + (x + 0.5) * h, py = -L / 2 + (y + 0.5) * h,
                 pz = -L / 2 + (z + 0.5) * h;
    double v = accumulate ? (double)V[i] : 0.0;
    for (int c = 0; c < cs.n; ++c)
      v += center_value(cs.p[c], px - cs.p[c][1], py - cs.p[c][2], pz - cs.p[c][3]);
    V[i] = (T)v;
  }
}

// partials[(c*3+a)*nblk + b] = block sums of rho_i dV_c/dR_a ; block (chunk, centre).
// finalize_partials scales by dV into grad[c*3+a].
template <typename T>
__global__ __launch_bounds__(kThr) void local_grad_kernel(int N, double L, Centers cs,
                                                          const T* __restrict__ rho,
                                                          double* partials) {
  __shared__ double red[16];
  const double h = L / N, dV = h * h * h;
  const int c = blockIdx.y;
  const double* cp = cs.p[c];
  const long n = (long)N * N * N;
  KSum g[3];
  for (int a = 0; a < 3; ++a) g[a].init();
  const long stride = (long)gridDim.x * kThr;
  for (long i = (long)blockIdx.x * kThr + threadIdx.x; i < n; i += stride) {
    const int z = (int)(i % N), y = (int)((i / N) % N), x = (int)(i / ((long)N * N));
    const double dx = -L / 2 + (x + 0.5) * h - cp[1];
    const double dy = -L / 2 + (y + 0.5) * h - cp[2];
    const double dz = -L / 2 + (z + 0.5) * h - cp[3];
    const double f = -(double)rho[i] * center_dvdr_over_r(cp, dx, dy, dz);
    g[0].add(f * dx);
    g[1].add(f * dy);
    g[2].add(f * dz);
  }
  const int nblk = gridDim.x;
  for (int a = 0; a < 3; ++a) {
    const double v = block_sum(g[a].value(), red);
    if (threadIdx.x == 0) partials[(long)(c * 3 + a) * nblk + blockIdx.x] = v;
  }
}

int load_centers(Centers& cs, int n, const double* centers) {
  GATO_REQUIRE(n >= 0 && n <= kMaxCenters, GATO_ERR_UNSUPPORTED,
               "%d centres; at most %d per call", n, kMaxCenters);
  cs.n = n;
  for (int c = 0; c < n; ++c) {
    for (int k = 0; k < GATO_CENTER_PARAMS; ++k) cs.p[c][k] = centers[c * GATO_CENTER_PARAMS + k];
    const int kind = (int)cs.p[c][0];
    GATO_REQUIRE(kind == GATO_CENTER_SOFT_COULOMB || kind == GATO_CENTER_HGH_LOCAL ||
                     kind == GATO_CENTER_HARMONIC || kind == GATO_CENTER_SLATER_1S,
                 GATO_ERR_VALUE, "unknown centre kind %d", kind);
  }
  return GATO_OK;
}

// ---------------------------------------------------------------------------
// HGH nonlocal projectors
// ---------------------------------------------------------------------------
struct NLChan {
  int ell, m, power;
  double r, norm;
};

struct NLAtom {
  int lo[3], dims[3];
  long npts, off;  // values of channel c at off + c*npts + local index
  int nch, ch0;    // channels ch0 .. ch0+nch-1 (global numbering)
  double pos[3];
  NLChan ch[kMaxCh];
  double h[kMaxCh * kMaxCh];  // coupling block of this atom
};

// Box-local index -> grid coordinates.  A box holds at most N^3 <= 512^3 < 2^31
// points, so 32-bit division is exact.
__device__ __forceinline__ void box_coords(const NLAtom& A, long li, int& x, int& y, int& z) {
  const int bz = A.dims[2], by = A.dims[1];
  const int l = (int)li, q = l / bz;
  z = A.lo[2] + (l - q * bz);
  y = A.lo[1] + q % by;
  x = A.lo[0] + q / by;
}

// value of p_{l i}(r) Y_lm and (optionally) its derivative wrt the centre R
__device__ __forceinline__ double proj_value(const NLChan& c, double dx, double dy, double dz) {
  const double r = sqrt(dx * dx + dy * dy + dz * dz + 1e-30);
  double rad = c.norm * exp(-(r * r) / (2.0 * c.r * c.r));
  for (int k = 0; k < c.power; ++k) rad *= r;
  if (c.ell == 0) return rad * 0.28209479177387814;  // 1/sqrt(4 pi)
  const double cc = 0.4886025119029199;             // sqrt(3/(4 pi))
  const double comp = (c.m == -1) ? dy : (c.m == 0 ? dz : dx);
  return rad * cc * comp / r;
}

__device__ __forceinline__ void proj_grad_R(const NLChan& c, double dx, double dy, double dz,
                                            double* g) {
  const double r = sqrt(dx * dx + dy * dy + dz * dz + 1e-30);
  const double ex = exp(-(r * r) / (2.0 * c.r * c.r));
  const double il2 = 1.0 / (c.r * c.r);
  if (c.ell == 0) {
    // f = norm r^p e ; f' = norm e (p r^{p-1} - r^{p+1}/r_l^2)
    double rp = 1.0;
    for (int k = 0; k < c.power; ++k) rp *= r;  // r^p
    double fprime = -c.norm * ex * rp * r * il2;
    if (c.power >= 1) {
      double rpm1 = 1.0;
      for (int k = 0; k < c.power - 1; ++k) rpm1 *= r;
      fprime += c.norm * ex * c.power * rpm1;
    }
    const double s = fprime * 0.28209479177387814 / r;
    g[0] = -s * dx; g[1] = -s * dy; g[2] = -s * dz;
  } else {
    const double cc = 0.4886025119029199;
    const int q = c.power - 1;  // g(r) = norm r^q e
    double rq = 1.0;
    for (int k = 0; k < q; ++k) rq *= r;
    const double gval = c.norm * rq * ex;
    double gp = -c.norm * ex * rq * r * il2;  // -norm e r^{q+1}/r_l^2
    if (q >= 1) {
      double rqm1 = 1.0;
      for (int k = 0; k < q - 1; ++k) rqm1 *= r;
      gp += c.norm * ex * q * rqm1;
    }
    const int a = (c.m == -1) ? 1 : (c.m == 0 ? 2 : 0);
    const double comp = (a == 0) ? dx : (a == 1 ? dy : dz);
    const double s = cc * gp / r * comp;
    double gr[3] = {s * dx, s * dy, s * dz};
    gr[a] += cc * gval;
    g[0] = -gr[0]; g[1] = -gr[1]; g[2] = -gr[2];
  }
}

template <typename T>
__global__ void nl_values_kernel(int N, double L, NLAtom A, T* __restrict__ vals) {
  const double h = L / N;
  const long total = A.npts * A.nch;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const int c = (int)(t / A.npts);
    const long li = t % A.npts;
    int x, y, z;
    box_coords(A, li, x, y, z);
    const double dx = -L / 2 + (x + 0.5) * h - A.pos[0];
    const double dy = -L / 2 + (y + 0.5) * h - A.pos[1];
    const double dz = -L / 2 + (z + 0.5) * h - A.pos[2];
    vals[A.off + (long)c * A.npts + li] = (T)proj_value(A.ch[c], dx, dy, dz);
  }
}

// Block partials of sum_box p_c psi_v (grid: chunks x nv); MODE 1 also of the
// derivative overlaps.  partials[((v*nslot + s)*nblk + b]; nl_finalize_kernel
// scatters the sums into o and dov.
template <typename T, int MODE>
__global__ __launch_bounds__(kThr) void nl_overlap_kernel(int N, double L, NLAtom A,
                                                          const T* __restrict__ vals,
                                                          const T* __restrict__ psi, long ld,
                                                          double* partials) {
  __shared__ double red[16];
  const int v = blockIdx.y;
  const T* p = psi + (long)v * ld;
  const double h = L / N;
  constexpr int NS = (MODE == 0) ? kMaxCh : 4 * kMaxCh;
  KSum acc[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) acc[k].init();
  const long stride = (long)gridDim.x * kThr;
  for (long li = (long)blockIdx.x * kThr + threadIdx.x; li < A.npts; li += stride) {
    int x, y, z;
    box_coords(A, li, x, y, z);
    const double ps = (double)p[((long)x * N + y) * N + z];
#pragma unroll
    for (int c = 0; c < kMaxCh; ++c) {
      if (c < A.nch) {
        acc[c].add((double)vals[A.off + (long)c * A.npts + li] * ps);
        if (MODE == 1) {
          const double dx = -L / 2 + (x + 0.5) * h - A.pos[0];
          const double dy = -L / 2 + (y + 0.5) * h - A.pos[1];
          const double dz = -L / 2 + (z + 0.5) * h - A.pos[2];
          double g[3];
          proj_grad_R(A.ch[c], dx, dy, dz, g);
          acc[kMaxCh + 3 * c + 0].add(g[0] * ps);
          acc[kMaxCh + 3 * c + 1].add(g[1] * ps);
          acc[kMaxCh + 3 * c + 2].add(g[2] * ps);
        }
      }
    }
  }
  const int nblk = gridDim.x;
  const int nslot = (MODE == 0) ? A.nch : 4 * A.nch;
  for (int s = 0; s < nslot; ++s) {
    const int k = (s < A.nch) ? s : kMaxCh + (s - A.nch);
    double val = 0.0;
#pragma unroll
    for (int kk = 0; kk < NS; ++kk)
      if (kk == k) val = acc[kk].value();
    const double b = block_sum(val, red);
    if (threadIdx.x == 0) partials[((long)v * nslot + s) * nblk + blockIdx.x] = b;
  }
}

// o[v*n_proj + ch0 + c] = dV * sum_box p_c psi_v, and (nslot > nch) the
// derivative overlaps dov[(v*n_proj + ch0 + c)*3 + a]; one 64-lane block per
// (vector, slot), each lane summing a fixed stride of the nblk block partials,
// then a fixed shuffle tree.
__global__ __launch_bounds__(64) void nl_finalize_kernel(const double* __restrict__ partials,
                                                         int nblk, int nv, int nslot, int nch,
                                                         int ch0, int n_proj, double dV,
                                                         double* __restrict__ o,
                                                         double* __restrict__ dov) {
  const int t = blockIdx.x;
  const int vv = t / nslot, s = t % nslot;
  KSum sum;
  sum.init();
  const double* pp = partials + (long)t * nblk;
  for (int b = threadIdx.x; b < nblk; b += 64) sum.add(pp[b]);
  const double r = wave_sum(sum.value()) * dV;
  if (threadIdx.x != 0) return;
  if (s < nch) {
    o[(long)vv * n_proj + ch0 + s] = r;
  } else {
    const int c = (s - nch) / 3, a = (s - nch) % 3;
    dov[((long)vv * n_proj + ch0 + c) * 3 + a] = r;
  }
}

// out_v[box] += sum_c (sum_q h_cq o_vq) p_c      (grid: chunks x nv)
template <typename T>
__global__ __launch_bounds__(kThr) void nl_add_kernel(int N, NLAtom A, const T* __restrict__ vals,
                                                      const double* __restrict__ o, int n_proj,
                                                      T* __restrict__ out, long ldo) {
  __shared__ double cf[kMaxCh];
  const int v = blockIdx.y;
  if (threadIdx.x < A.nch) {
    double s = 0.0;
    for (int q = 0; q < A.nch; ++q) s += A.h[threadIdx.x * kMaxCh + q] * o[(long)v * n_proj + A.ch0 + q];
    cf[threadIdx.x] = s;
  }
  __syncthreads();
  T* dst = out + (long)v * ldo;
  const long stride = (long)gridDim.x * kThr;
  for (long li = (long)blockIdx.x * kThr + threadIdx.x; li < A.npts; li += stride) {
    int x, y, z;
    box_coords(A, li, x, y, z);
    double s = 0.0;
    for (int c = 0; c < A.nch; ++c) s += cf[c] * (double)vals[A.off + (long)c * A.npts + li];
    const long idx = ((long)x * N + y) * N + z;
    dst[idx] = (T)((double)dst[idx] + s);
  }
}

// out_v[box] += sum_c cf_c p_c after nl_overlap_kernel<T, 0>: every block first
// forms o_c = dV sum_b partials[(v*nch + c)*nblk + b] itself (wave c % 4 sums
// the nblk block partials in one fixed order, so all blocks agree bit for bit
// and no finalize launch is needed), then cf = H o; block 0 also stores o.
template <typename T>
__global__ __launch_bounds__(kThr) void nl_add_reduce_kernel(int N, NLAtom A,
                                                             const T* __restrict__ vals,
                                                             const double* __restrict__ partials,
                                                             int nblk, double dV, int n_proj,
                                                             double* __restrict__ o,
                                                             T* __restrict__ out, long ldo) {
  __shared__ double ov[kMaxCh], cf[kMaxCh];
  const int v = blockIdx.y;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int c = wave; c < A.nch; c += kThr / 64) {
    const double* pp = partials + ((long)v * A.nch + c) * nblk;
    KSum acc;
    acc.init();
    for (int b = lane; b < nblk; b += 64) acc.add(pp[b]);
    const double t = wave_sum(acc.value());
    if (lane == 0) ov[c] = t * dV;
  }
  __syncthreads();
  if (threadIdx.x < A.nch) {
    double t = 0.0;
    for (int q = 0; q < A.nch; ++q) t += A.h[threadIdx.x * kMaxCh + q] * ov[q];
    cf[threadIdx.x] = t;
    if (blockIdx.x == 0) o[(long)v * n_proj + A.ch0 + threadIdx.x] = ov[threadIdx.x];
  }
  __syncthreads();
  T* dst = out + (long)v * ldo;
  const long stride = (long)gridDim.x * kThr;
  for (long li = (long)blockIdx.x * kThr + threadIdx.x; li < A.npts; li += stride) {
    int x, y, z;
    box_coords(A, li, x, y, z);
    double t = 0.0;
    for (int c = 0; c < A.nch; ++c) t += cf[c] * (double)vals[A.off + (long)c * A.npts + li];
    const long idx = ((long)x * N + y) * N + z;
    dst[idx] = (T)((double)dst[idx] + t);
  }
}

constexpr int kFusedThr = 1024;

// Whole-box apply of one atom in one 1024-thread block per vector, for boxes
// of at most kFusedThr * PER points: o_c = dV sum_box p_c psi (fixed-order
// block sums, also stored in o), cf_c = sum_q h_cq o_q, then out[box] +=
// sum_c cf_c p_c.  Every point of a thread is touched once: its grid index
// (32-bit arithmetic), psi and out are loaded together in the first phase, so
// the kernel waits on memory once before the reduction instead of again for
// the update of out.  The box is read completely before it is written, so psi
// may alias out.
template <typename T, int PER>
__global__ __launch_bounds__(kFusedThr) void nl_apply_fused_small_kernel(
    int N, NLAtom A, const T* __restrict__ vals, const T* psi, long ld, int n_proj, double dV,
    double* __restrict__ o, T* out, long ldo) {
  __shared__ double red[16];
  __shared__ double ov[kMaxCh], cf[kMaxCh];
  const int v = blockIdx.y;
  const T* p = psi + (long)v * ld;
  T* dst = out + (long)v * ldo;
  const int npts = (int)A.npts, bz = A.dims[2], by = A.dims[1];
  int idx[PER];
  double ps[PER], cur[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int li = threadIdx.x + u * kFusedThr;
    idx[u] = -1;
    ps[u] = cur[u] = 0.0;
    if (li < npts) {
      const int z = A.lo[2] + li % bz, q = li / bz;
      const int y = A.lo[1] + q % by, x = A.lo[0] + q / by;
      idx[u] = (x * N + y) * N + z;
      ps[u] = (double)p[idx[u]];
      cur[u] = (double)dst[idx[u]];
    }
  }
  KSum acc[kMaxCh];
#pragma unroll
  for (int c = 0; c < kMaxCh; ++c) acc[c].init();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int li = threadIdx.x + u * kFusedThr;
    if (li < npts) {
#pragma unroll
      for (int c = 0; c < kMaxCh; ++c)
        if (c < A.nch) acc[c].add((double)vals[A.off + (long)c * A.npts + li] * ps[u]);
    }
  }
#pragma unroll
  for (int c = 0; c < kMaxCh; ++c) {
    if (c < A.nch) {
      const double b = block_sum(acc[c].value(), red);
      if (threadIdx.x == 0) {
        ov[c] = b * dV;
        o[(long)v * n_proj + A.ch0 + c] = b * dV;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < A.nch) {
    double t = 0.0;
    for (int q = 0; q < A.nch; ++q) t += A.h[threadIdx.x * kMaxCh + q] * ov[q];
    cf[threadIdx.x] = t;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int li = threadIdx.x + u * kFusedThr;
    if (li < npts) {
      double t = 0.0;
      for (int c = 0; c < A.nch; ++c) t += cf[c] * (double)vals[A.off + (long)c * A.npts + li];
      dst[idx[u]] = (T)(cur[u] + t);
    }
  }
}

// E[v] = sum_atoms sum_cq o_vc h_cq o_vq ; grad[a*3+k] = occ * sum_v 2 sum_cq h_cq do_vck o_vq
__global__ void nl_contract_kernel(int nv, int n_proj, const double* __restrict__ o,
                                   const double* __restrict__ dov, const double* __restrict__ H,
                                   const int* __restrict__ ch_atom, int n_atoms, double occ,
                                   double* __restrict__ E, double* __restrict__ grad) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (E) {
    for (int v = 0; v < nv; ++v) {
      double e = 0.0;
      for (int c = 0; c < n_proj; ++c)
        for (int q = 0; q < n_proj; ++q)
          e += o[(long)v * n_proj + c] * H[c * n_proj + q] * o[(long)v * n_proj + q];
      E[v] = e;
    }
  }
  if (grad) {
    for (int a = 0; a < 3 * n_atoms; ++a) grad[a] = 0.0;
    for (int v = 0; v < nv; ++v)
      for (int c = 0; c < n_proj; ++c)
        for (int q = 0; q < n_proj; ++q) {
          const double hq = H[c * n_proj + q];
          if (hq == 0.0) continue;
          const int at = ch_atom[c];
          for (int k = 0; k < 3; ++k)
            grad[at * 3 + k] += occ * 2.0 * hq * dov[((long)v * n_proj + c) * 3 + k] *
                                o[(long)v * n_proj + q];
        }
  }
}

inline int nblocks(long n) {
  long b = (n + kThr * 4 - 1) / (kThr * 4);
  if (b < 1) b = 1;
  if (b > kReduceBlocks) b = kReduceBlocks;
  return (int)b;
}

}  // namespace
}  // namespace gato

struct gato_nl_plan {
  int dtype, N, n_atoms, n_proj;
  double L, h, dV;
  std::vector<gato::NLAtom> atoms;  // atoms that carry projectors
  std::vector<int> ch_atom;         // global channel -> atom index
  void* vals = nullptr;
  double* H = nullptr;       // device n_proj^2
  int* d_ch_atom = nullptr;  // device
  double* o = nullptr;       // device overlaps, capacity cap_nv * n_proj
  double* dov = nullptr;     // device derivative overlaps, capacity cap_nv * n_proj * 3
  int cap_nv = 0;
  long total_pts = 0;
  bool has_empty_box = false;
};

namespace gato {
namespace {

int nl_ensure_capacity(gato_nl_plan* pl, int nv) {
  if (nv <= pl->cap_nv) return GATO_OK;
  if (pl->o) (void)hipFree(pl->o);
  if (pl->dov) (void)hipFree(pl->dov);
  pl->o = nullptr;
  pl->dov = nullptr;
  const int cap = nv < 16 ? 16 : nv;
  GATO_HIP_CHECK(hipMalloc(&pl->o, sizeof(double) * cap * (pl->n_proj > 0 ? pl->n_proj : 1)));
  GATO_HIP_CHECK(hipMalloc(&pl->dov, sizeof(double) * cap * 3 * (pl->n_proj > 0 ? pl->n_proj : 1)));
  pl->cap_nv = cap;
  return GATO_OK;
}

// radius beyond which |p(r)| < tol * max|p|, by an outward scan in r
double cutoff_radius(const NLChan& c, double tol) {
  auto f = [&](double r) {
    double v = std::exp(-(r * r) / (2.0 * c.r * c.r));
    for (int k = 0; k < c.power; ++k) v *= r;
    return std::fabs(v);
  };
  const double dr = c.r * 1e-3;
  double peak = 0.0, r = 0.0;
  for (; r < 60.0 * c.r; r += dr) peak = std::fmax(peak, f(r));
  double rc = 60.0 * c.r;
  for (r = 60.0 * c.r; r > 0; r -= dr) {
    if (f(r) >= tol * peak) { rc = r + dr; break; }
  }
  return rc;
}

// Zero n doubles with a kernel rather than hipMemsetAsync, so that calls
// recorded into a HIP graph replay correctly.
__global__ void zero_doubles_kernel(double* p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

int zero_doubles(double* p, long n, hipStream_t s) {
  if (n <= 0) return GATO_OK;
  const long b = (n + 255) / 256;
  hipLaunchKernelGGL(zero_doubles_kernel, dim3((unsigned)(b < 1024 ? b : 1024)), dim3(256), 0, s, p,
                     n);
  GATO_LAUNCH_CHECK();
  return GATO_OK;
}

template <typename T>
int nl_apply_t(gato_nl_plan* pl, int nv, const T* psi, long ld, T* out, long ldo, void* ws,
               hipStream_t s, bool add, bool energy, double* E, double occ, double* grad) {
  if (pl->n_proj == 0) {
    int st = E ? zero_doubles(E, nv, s) : GATO_OK;
    if (!st && grad) st = zero_doubles(grad, 3L * pl->n_atoms, s);
    return st;
  }
  int st = nl_ensure_capacity(pl, nv);
  if (st) return st;
  const bool want_grad = grad != nullptr;
  if (pl->has_empty_box) {  // channels of atoms outside the grid overlap nothing
    st = zero_doubles(pl->o, (long)nv * pl->n_proj, s);
    if (!st) st = zero_doubles(pl->dov, (long)nv * 3 * pl->n_proj, s);
    if (st) return st;
  }
  for (const NLAtom& A : pl->atoms) {
    if (add && !want_grad && A.npts <= 8L * kFusedThr && (long)pl->N * pl->N * pl->N < (1L << 31)) {
      hipLaunchKernelGGL((nl_apply_fused_small_kernel<T, 8>), dim3(1, nv), dim3(kFusedThr), 0, s,
                         pl->N, A, (const T*)pl->vals, psi, ld, pl->n_proj, pl->dV, pl->o, out,
                         ldo);
      GATO_LAUNCH_CHECK();
      continue;
    }
    if (add && !want_grad) {
      // overlap partials over many blocks, then the add kernel reduces them
      long ob = (A.npts + kThr * 4 - 1) / (kThr * 4);
      const int oblk = (int)(ob < 1 ? 1 : (ob > 512 ? 512 : ob));
      GATO_REQUIRE((long)oblk * A.nch * nv <= kPartialsCap, GATO_ERR_UNSUPPORTED,
                   "too many vectors for the reduction workspace (%d)", nv);
      hipLaunchKernelGGL((nl_overlap_kernel<T, 0>), dim3(oblk, nv), dim3(kThr), 0, s, pl->N,
                         pl->L, A, (const T*)pl->vals, psi, ld, ws_partials(ws));
      GATO_LAUNCH_CHECK();
      hipLaunchKernelGGL((nl_add_reduce_kernel<T>), dim3(nblocks(A.npts), nv), dim3(kThr), 0, s,
                         pl->N, A, (const T*)pl->vals, ws_partials(ws), oblk, pl->dV, pl->n_proj,
                         pl->o, out, ldo);
      GATO_LAUNCH_CHECK();
      continue;
    }
    const int nblk = nblocks(A.npts);
    const int nslot = want_grad ? 4 * A.nch : A.nch;
    GATO_REQUIRE((long)nblk * nslot * nv <= kPartialsCap, GATO_ERR_UNSUPPORTED,
                 "too many vectors for the reduction workspace (%d)", nv);
    dim3 grid(nblk, nv);
    if (want_grad)
      hipLaunchKernelGGL((nl_overlap_kernel<T, 1>), grid, dim3(kThr), 0, s, pl->N, pl->L, A,
                         (const T*)pl->vals, psi, ld, ws_partials(ws));
    else
      hipLaunchKernelGGL((nl_overlap_kernel<T, 0>), grid, dim3(kThr), 0, s, pl->N, pl->L, A,
                         (const T*)pl->vals, psi, ld, ws_partials(ws));
    GATO_LAUNCH_CHECK();
    hipLaunchKernelGGL(nl_finalize_kernel, dim3(nv * nslot), dim3(64), 0, s, ws_partials(ws),
                       nblk, nv, nslot, A.nch, A.ch0, pl->n_proj, pl->dV, pl->o, pl->dov);
    GATO_LAUNCH_CHECK();
    if (add) {
      hipLaunchKernelGGL((nl_add_kernel<T>), dim3(nblocks(A.npts), nv), dim3(kThr), 0, s, pl->N, A,
                         (const T*)pl->vals, pl->o, pl->n_proj, out, ldo);
      GATO_LAUNCH_CHECK();
    }
  }
  if (energy || want_grad) {
    hipLaunchKernelGGL(nl_contract_kernel, dim3(1), dim3(64), 0, s, nv, pl->n_proj, pl->o,
                       pl->dov, pl->H, pl->d_ch_atom, pl->n_atoms, occ, E, grad);
    GATO_LAUNCH_CHECK();
  }
  return GATO_OK;
}

}  // namespace

int nl_apply_internal(gato_nl_plan* pl, int nv, const void* psi, long ld, void* out, long ldo,
                      void* ws, hipStream_t s) {
  if (pl->dtype == GATO_F32)
    return nl_apply_t<float>(pl, nv, (const float*)psi, ld, (float*)out, ldo, ws, s, true, false,
                             nullptr, 0.0, nullptr);
  return nl_apply_t<double>(pl, nv, (const double*)psi, ld, (double*)out, ldo, ws, s, true, false,
                            nullptr, 0.0, nullptr);
}

int nl_plan_dtype(const gato_nl_plan* pl) { return pl->dtype; }

}  // namespace gato

using namespace gato;

extern "C" int gato_local_potential(int dtype, int N, double L, int n_centers,
                                    const double* centers, int accumulate, void* V, void* stream) {
  clear_ok();
  GATO_REQUIRE(N >= 1 && L > 0 && V && (n_centers == 0 || centers), GATO_ERR_VALUE,
               "bad arguments");
  hipStream_t s = as_stream(stream);
  int done = 0;
  bool acc = accumulate != 0;
  do {
    const int n = (n_centers - done) > kMaxCenters ? kMaxCenters : (n_centers - done);
    Centers cs;
    int st = load_centers(cs, n, centers + (size_t)done * GATO_CENTER_PARAMS);
    if (st) return st;
    const long tot = (long)N * N * N;
    long nb = (tot + 1023) / 1024;
    if (nb > 65535) nb = 65535;
    if (dtype == GATO_F32)
      hipLaunchKernelGGL((local_potential_kernel<float>), dim3(nb), dim3(256), 0, s, N, L, cs,
                         acc ? 1 : 0, (float*)V);
    else
      hipLaunchKernelGGL((local_potential_kernel<double>), dim3(nb), dim3(256), 0, s, N, L, cs,
                         acc ? 1 : 0, (double*)V);
    GATO_LAUNCH_CHECK();
    done += n;
    acc = true;
  } while (done < n_centers);
  return GATO_OK;
}

extern "C" int gato_local_potential_grad(int dtype, int N, double L, int n_centers,
                                         const double* centers, const void* rho, double* grad,
                                         void* ws, void* stream) {
  clear_ok();
  GATO_REQUIRE(N >= 1 && L > 0 && rho && grad && ws && n_centers >= 1, GATO_ERR_VALUE,
               "bad arguments");
  Centers cs;
  int st = load_centers(cs, n_centers, centers);
  if (st) return st;
  hipStream_t s = as_stream(stream);
  const long tot = (long)N * N * N;
  int nblk = nblocks(tot);
  while ((long)nblk * 3 * n_centers > kPartialsCap) nblk /= 2;
  dim3 grid(nblk, n_centers);
  if (dtype == GATO_F32)
    hipLaunchKernelGGL((local_grad_kernel<float>), grid, dim3(kThr), 0, s, N, L, cs,
                       (const float*)rho, ws_partials(ws));
  else
    hipLaunchKernelGGL((local_grad_kernel<double>), grid, dim3(kThr), 0, s, N, L, cs,
                       (const double*)rho, ws_partials(ws));
  GATO_LAUNCH_CHECK();
  const double h = L / N;
  return finalize_partials(ws_partials(ws), nblk, 3 * n_centers, h * h * h, false, grad, s);
}

extern "C" int gato_nl_plan_create(int dtype, int N, double L, int n_atoms, const double* atom_pos,
                                   int n_proj, const int* proj_atom, const int* proj_ell,
                                   const int* proj_m, const int* proj_i, const double* proj_r,
                                   const double* hmat, double cutoff_tol, gato_nl_plan** plan,
                                   void* stream) {
  clear_ok();
  GATO_REQUIRE(plan != nullptr, GATO_ERR_VALUE, "null plan pointer");
  *plan = nullptr;
  GATO_REQUIRE(dtype == GATO_F32 || dtype == GATO_F64, GATO_ERR_VALUE, "unknown dtype %d", dtype);
  GATO_REQUIRE(N >= 1 && L > 0 && n_atoms >= 1 && n_proj >= 0 && atom_pos, GATO_ERR_VALUE,
               "bad arguments");
  GATO_REQUIRE(cutoff_tol > 0 && cutoff_tol < 1, GATO_ERR_VALUE, "cutoff_tol must be in (0,1)");
  hipStream_t s = as_stream(stream);
  auto* pl = new gato_nl_plan();
  pl->dtype = dtype;
  pl->N = N;
  pl->L = L;
  pl->h = L / N;
  pl->dV = pl->h * pl->h * pl->h;
  pl->n_atoms = n_atoms;
  pl->n_proj = n_proj;
  pl->ch_atom.assign(n_proj, 0);
  // group channels by atom (channels of one atom must be contiguous)
  long off = 0;
  for (int p = 0; p < n_proj;) {
    const int a = proj_atom[p];
    if (a < 0 || a >= n_atoms) {
      delete pl;
      set_error(GATO_ERR_VALUE, "projector %d names atom %d of %d", p, a, n_atoms);
      return GATO_ERR_VALUE;
    }
    NLAtom A{};
    A.ch0 = p;
    A.pos[0] = atom_pos[3 * a];
    A.pos[1] = atom_pos[3 * a + 1];
    A.pos[2] = atom_pos[3 * a + 2];
    double rc = 0.0;
    int q = p;
    for (; q < n_proj && proj_atom[q] == a; ++q) {
      if (q - p >= kMaxCh || proj_ell[q] < 0 || proj_ell[q] > 1 || proj_i[q] < 1 ||
          proj_r[q] <= 0 || proj_m[q] < -proj_ell[q] || proj_m[q] > proj_ell[q]) {
        delete pl;
        set_error(GATO_ERR_UNSUPPORTED,
                  "projector %d: need l <= 1, |m| <= l, i >= 1, r > 0 and <= %d channels per atom",
                  q, kMaxCh);
        return GATO_ERR_UNSUPPORTED;
      }
      NLChan& c = A.ch[q - p];
      c.ell = proj_ell[q];
      c.m = proj_m[q];
      c.r = proj_r[q];
      c.power = proj_ell[q] + 2 * (proj_i[q] - 1);
      const double nexp = proj_ell[q] + (4.0 * proj_i[q] - 1.0) / 2.0;
      c.norm = std::sqrt(2.0) / (std::pow(c.r, nexp) * std::sqrt(std::tgamma(nexp)));
      rc = std::fmax(rc, cutoff_radius(c, cutoff_tol));
      pl->ch_atom[q] = a;
    }
    A.nch = q - p;
    for (int i = 0; i < A.nch; ++i)
      for (int j = 0; j < A.nch; ++j) A.h[i * kMaxCh + j] = hmat[(p + i) * n_proj + (p + j)];
    for (int ax = 0; ax < 3; ++ax) {
      // grid index i has coordinate -L/2 + (i + 1/2) h
      int lo = (int)std::ceil((A.pos[ax] - rc + L / 2) / pl->h - 0.5);
      int hi = (int)std::floor((A.pos[ax] + rc + L / 2) / pl->h - 0.5);
      lo = lo < 0 ? 0 : lo;
      hi = hi > N - 1 ? N - 1 : hi;
      if (hi < lo) hi = lo - 1;
      A.lo[ax] = lo;
      A.dims[ax] = hi - lo + 1;
    }
    A.npts = (long)A.dims[0] * A.dims[1] * A.dims[2];
    A.off = off;
    off += A.npts * A.nch;
    if (A.npts > 0) pl->atoms.push_back(A);
    else pl->has_empty_box = true;
    p = q;
  }
  pl->total_pts = off;
  const size_t es = dtype == GATO_F32 ? 4 : 8;
  int st = GATO_OK;
  if (n_proj > 0) {
    if (hipMalloc(&pl->vals, es * (off > 0 ? off : 1)) != hipSuccess ||
        hipMalloc(&pl->H, sizeof(double) * n_proj * n_proj) != hipSuccess ||
        hipMalloc(&pl->d_ch_atom, sizeof(int) * n_proj) != hipSuccess) {
      gato_nl_plan_destroy(pl);
      set_error(GATO_ERR_HIP, "hipMalloc failed for the nonlocal plan");
      return GATO_ERR_HIP;
    }
    if (hipMemcpyAsync(pl->H, hmat, sizeof(double) * n_proj * n_proj, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemcpyAsync(pl->d_ch_atom, pl->ch_atom.data(), sizeof(int) * n_proj,
                       hipMemcpyHostToDevice, s) != hipSuccess) {
      gato_nl_plan_destroy(pl);
      set_error(GATO_ERR_HIP, "hipMemcpyAsync failed for the nonlocal plan");
      return GATO_ERR_HIP;
    }
    for (const NLAtom& A : pl->atoms) {
      long nb = (A.npts * A.nch + 255) / 256;
      if (nb > 65535) nb = 65535;
      if (dtype == GATO_F32)
        hipLaunchKernelGGL((nl_values_kernel<float>), dim3(nb), dim3(256), 0, s, N, L, A,
                           (float*)pl->vals);
      else
        hipLaunchKernelGGL((nl_values_kernel<double>), dim3(nb), dim3(256), 0, s, N, L, A,
                           (double*)pl->vals);
      if (hipGetLastError() != hipSuccess) {
        gato_nl_plan_destroy(pl);
        set_error(GATO_ERR_HIP, "projector sampling launch failed");
        return GATO_ERR_HIP;
      }
    }
    st = nl_ensure_capacity(pl, 16);
    if (st) {
      gato_nl_plan_destroy(pl);
      return st;
    }
    if (hipStreamSynchronize(s) != hipSuccess) {
      gato_nl_plan_destroy(pl);
      set_error(GATO_ERR_HIP, "nonlocal plan setup failed");
      return GATO_ERR_HIP;
    }
  }
  *plan = pl;
  return GATO_OK;
}

extern "C" int gato_nl_plan_destroy(gato_nl_plan* pl) {
  if (!pl) return GATO_OK;
  (void)hipDeviceSynchronize();
  if (pl->vals) (void)hipFree(pl->vals);
  if (pl->H) (void)hipFree(pl->H);
  if (pl->d_ch_atom) (void)hipFree(pl->d_ch_atom);
  if (pl->o) (void)hipFree(pl->o);
  if (pl->dov) (void)hipFree(pl->dov);
  delete pl;
  return GATO_OK;
}

extern "C" long gato_nl_plan_points(const gato_nl_plan* pl) { return pl ? pl->total_pts : 0; }

extern "C" int gato_nl_apply(gato_nl_plan* pl, int nv, const void* psi, long ld, void* out,
                             long ldo, void* ws, void* stream) {
  clear_ok();
  GATO_REQUIRE(pl && psi && out && ws && nv >= 1, GATO_ERR_VALUE, "bad arguments");
  return nl_apply_internal(pl, nv, psi, ld, out, ldo, ws, as_stream(stream));
}

extern "C" int gato_nl_expectation(gato_nl_plan* pl, int nv, const void* psi, long ld, double* E,
                                   void* ws, void* stream) {
  clear_ok();
  GATO_REQUIRE(pl && psi && E && ws && nv >= 1, GATO_ERR_VALUE, "bad arguments");
  hipStream_t s = as_stream(stream);
  if (pl->dtype == GATO_F32)
    return nl_apply_t<float>(pl, nv, (const float*)psi, ld, nullptr, 0, ws, s, false, true, E, 0.0,
                             nullptr);
  return nl_apply_t<double>(pl, nv, (const double*)psi, ld, nullptr, 0, ws, s, false, true, E, 0.0,
                            nullptr);
}

extern "C" int gato_nl_grad(gato_nl_plan* pl, int nv, const void* psi, long ld, double occ,
                            double* grad, void* ws, void* stream) {
  clear_ok();
  GATO_REQUIRE(pl && psi && grad && ws && nv >= 1, GATO_ERR_VALUE, "bad arguments");
  hipStream_t s = as_stream(stream);
  if (pl->dtype == GATO_F32)
    return nl_apply_t<float>(pl, nv, (const float*)psi, ld, nullptr, 0, ws, s, false, false,
                             nullptr, occ, grad);
  return nl_apply_t<double>(pl, nv, (const double*)psi, ld, nullptr, 0, ws, s, false, false,
                            nullptr, occ, grad);
}
