// Fock operator, native Lanczos loop, host tridiagonal eigensolver and the
// library's error channel.
//
// Reference: scf.py:88-158 (Fock matvec: kinetic + V_ext + V_nl + V_H - K),
// solvers/lanczos.py:110-206 (Lanczos with full reorthogonalisation,
// tridiagonal eigh, Ritz reconstruction).
//
// One Lanczos step j issues, all on one stream with no host round trip:
//   w  = F v_j - beta_{j-1} v_{j-1}   exchange (5 FFT passes), stencil+V_eff,
//                                     nonlocal overlap+update
//   alpha_j = <v_j, w>                the reference's alpha (lanczos.py:150)
//   c  = V[0..j]^T (w - alpha_j v_j)  batched dot with the alpha term removed
//                                     on the fly; these coefficients are
//                                     rounding-level, so one classical sweep
//                                     equals the reference's modified
//                                     Gram-Schmidt to second order.  (Taking
//                                     alpha_j inside the same sweep instead
//                                     is unstable: the large alpha_j times the
//                                     basis' loss of orthogonality grows
//                                     step by step.)
//   w -= alpha_j v_j + V c             basis walked downwards
//   beta_j = ||w||, v_{j+1} = w / beta_j   one launch from the wave partials
// The K wanted eigenpairs of the m x m tridiagonal are found on the host in
// double (Sturm bisection, then inverse iteration), then the K Ritz vectors
// are formed on the device.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "common.hpp"

struct gato_poisson_plan;
struct gato_nl_plan;

namespace gato {

namespace {
thread_local char g_err[1024] = "";
thread_local int g_code = 0;
}  // namespace

void set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  g_code = code;
}

int clear_ok() {
  g_err[0] = '\0';
  g_code = 0;
  return 0;
}

namespace {
struct Probe {
  int kind = PROBE_NONE;
  int elem = 0;                // 0: any element size
  std::vector<hipEvent_t> ev;  // pairs (begin, end), reused across resets
  size_t used = 0;             // events recorded since the last reset
  long launches = 0;
  double bytes = 0.0;
};
Probe g_probe;
}  // namespace

bool probe_on(int kind, int elem_bytes) {
  return g_probe.kind != PROBE_NONE && g_probe.kind == kind &&
         (g_probe.elem == 0 || g_probe.elem == elem_bytes);
}

namespace {
// Launches recorded into a graph are not probed: the probe's events would be
// recorded once and replayed without its bookkeeping.
bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return st != hipStreamCaptureStatusNone;
}
}  // namespace

void probe_begin(hipStream_t s) {
  Probe& p = g_probe;
  if (capturing(s)) return;
  if (p.used + 2 > p.ev.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    p.ev.push_back(a);
    p.ev.push_back(b);
  }
  (void)hipEventRecord(p.ev[p.used], s);
}

void probe_end(hipStream_t s, double bytes) {
  Probe& p = g_probe;
  if (capturing(s) || p.used + 2 > p.ev.size()) return;
  (void)hipEventRecord(p.ev[p.used + 1], s);
  p.used += 2;
  p.launches += 1;
  p.bytes += bytes;
}

int stencil_apply(int dtype, int N, int batch, double h, int order, int boundary, double c_lap,
                  const void* psi, const void* V, int accumulate, const void* sub,
                  const double* sub_coef, void* out, hipStream_t s);
int dot_batched(int dtype, long n, int nk, const void* A, long lda, const void* B, long ldb,
                const void* C, const void* S, const double* s_coef, double scale, double* out,
                void* ws, hipStream_t s, long cap = kPartialsCap);
int cgs_update(int dtype, long n, int nk, const void* V, long ldv, const double* coef,
               const void* S, const double* s_coef, const void* w, void* w_out, double scale,
               double* beta, void* ws, hipStream_t s);
int cgs_coefficients(int dtype, long n, int nk, const void* V, long ldv, const void* w,
                     double scale, double* coef, double* alpha, void* ws, hipStream_t s);
int scale_by_inverse(int dtype, long n, const void* w, const double* beta, void* v, hipStream_t s);
int cgs_update_partials(int dtype, long n, int nk, const void* V, long ldv, const double* coef,
                        const void* S, const double* s_coef, const void* w, void* w_out,
                        long* np, void* ws, hipStream_t s);
int scale_by_partials(int dtype, long n, const void* w, const void* ws, long np, double scale,
                      double* beta, void* v, hipStream_t s);
int ritz_combine(int dtype, long n, int m, int K, const void* V, long ldv, const double* S,
                 void* out, long ldo, hipStream_t s);
int poisson_exchange(gato_poisson_plan* pl, int nrhs, const void* psi, long ldpsi,
                     const void* orb, long ld, int n_orb, double coef, int accumulate, void* out,
                     long ldo, hipStream_t s);
int poisson_plan_dtype(const gato_poisson_plan* pl);
int poisson_plan_N(const gato_poisson_plan* pl);
int nl_apply_internal(gato_nl_plan* pl, int nv, const void* psi, long ld, void* out, long ldo,
                      void* ws, hipStream_t s);
int nl_plan_dtype(const gato_nl_plan* pl);

// ---------------------------------------------------------------------------
// Symmetric tridiagonal eigensolver for the K wanted pairs.
// ---------------------------------------------------------------------------
// The K lowest eigenvalues (ascending) of the symmetric tridiagonal T (diagonal
// a, off-diagonal b[0..m-1)) by bisection on Sturm counts: the number of
// eigenvalues below x is the number of negative pivots of the LDL^T
// factorisation of T - x I, q_i = (a_i - x) - b_{i-1}^2 / q_{i-1}, with a zero
// pivot replaced by -pivmin (Barth, Martin and Wilkinson, 1967).  Each value is
// bisected from the Gershgorin interval until the bracket is at the rounding
// level of ||T||: O(K m log(1/eps)) instead of the O(m^2) of a full QR/QL sweep.
void tridiag_lowest_eigvals(int m, const double* a, const double* b, int K, double* out) {
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
  auto count_below = [&](double x) {
    int c = 0;
    double q = a[0] - x;
    if (std::fabs(q) < pivmin) q = -pivmin;
    if (q < 0) ++c;
    for (int i = 1; i < m; ++i) {
      q = (a[i] - x) - b[i - 1] * b[i - 1] / q;
      if (std::fabs(q) < pivmin) q = -pivmin;
      if (q < 0) ++c;
    }
    return c;
  };
  double left = lo;  // eigenvalue k lies above the previous one
  for (int k = 0; k < K; ++k) {
    double l = left, h = hi;
    for (int it = 0; it < 200; ++it) {
      const double mid = 0.5 * (l + h);
      if (h - l <= 2.0 * eps * std::fmax(std::fabs(l), std::fabs(h)) + pivmin || mid == l || mid == h)
        break;
      if (count_below(mid) > k) h = mid;
      else l = mid;
    }
    out[k] = 0.5 * (l + h);
    left = l;
  }
}

// Eigenvectors of the unreduced symmetric tridiagonal T (diagonal a, off-diagonal
// b[0..m-1)) for the ascending eigenvalues lam[0..K), by inverse iteration:
// T - sigma I is factored by Gaussian elimination with partial pivoting (U then
// has two superdiagonals; a zero pivot is replaced by eps ||T||), three solves
// from a fixed start vector, and after each solve the vector is orthogonalised
// against the earlier vectors of its cluster (eigenvalues within 1e-3 ||T||)
// -- the scheme of Wilkinson / Peters and Wilkinson (1971).  Z is m x K,
// Z[j*K + k]; each column has unit norm and its largest-magnitude component
// positive.
void tridiag_inverse_iteration(int m, const double* a, const double* b, int K, const double* lam,
                               double* Z) {
  const double eps = 2.220446049250313e-16;
  double tn = 0.0;
  for (int i = 0; i < m; ++i) {
    double r = std::fabs(a[i]);
    if (i > 0) r += std::fabs(b[i - 1]);
    if (i + 1 < m) r += std::fabs(b[i]);
    tn = std::fmax(tn, r);
  }
  if (tn == 0.0) tn = 1.0;
  const double tiny = eps * tn;
  std::vector<double> D(m), U1(m), U2(m), L(m), x(m), y(m);
  std::vector<char> piv(m);
  double sigma_prev = 0.0;
  for (int k = 0; k < K; ++k) {
    double sigma = lam[k];
    if (k > 0 && sigma - sigma_prev < 10.0 * tiny) sigma = sigma_prev + 10.0 * tiny;
    sigma_prev = sigma;
    // factor T - sigma I: row i holds D[i] (col i), U1[i] (i+1), U2[i] (i+2)
    for (int i = 0; i < m; ++i) {
      D[i] = a[i] - sigma;
      U1[i] = i + 1 < m ? b[i] : 0.0;
      U2[i] = 0.0;
    }
    double sub = m > 1 ? b[0] : 0.0;  // element (i+1, i) of the active rows
    for (int i = 0; i + 1 < m; ++i) {
      // rows i: (D[i], U1[i], U2[i]); i+1: (sub, D[i+1], U1[i+1]) before elimination
      double r1d = D[i + 1], r1u = U1[i + 1];
      if (std::fabs(sub) > std::fabs(D[i])) {
        piv[i] = 1;
        const double t0 = D[i], t1 = U1[i], t2 = U2[i];
        D[i] = sub;
        U1[i] = r1d;
        U2[i] = r1u;
        // the old row i becomes row i+1: (t0, t1, t2)
        const double l = t0 / D[i];
        L[i] = l;
        D[i + 1] = t1 - l * U1[i];
        U1[i + 1] = t2 - l * U2[i];
      } else {
        piv[i] = 0;
        if (D[i] == 0.0) D[i] = tiny;
        const double l = sub / D[i];
        L[i] = l;
        D[i + 1] = r1d - l * U1[i];
        U1[i + 1] = r1u - l * U2[i];
      }
      U2[i + 1] = 0.0;
      sub = i + 2 < m ? b[i + 1] : 0.0;
    }
    if (D[m - 1] == 0.0) D[m - 1] = tiny;
    for (int i = 0; i < m; ++i) x[i] = 1.0 + 0.5 * std::sin(1.0 + 0.7 * i);
    for (int it = 0; it < 3; ++it) {
      // forward: apply the row exchanges and multipliers to y = x
      for (int i = 0; i < m; ++i) y[i] = x[i];
      for (int i = 0; i + 1 < m; ++i) {
        if (piv[i]) std::swap(y[i], y[i + 1]);
        y[i + 1] -= L[i] * y[i];
      }
      // back substitution with U (two superdiagonals)
      for (int i = m - 1; i >= 0; --i) {
        double t = y[i];
        if (i + 1 < m) t -= U1[i] * x[i + 1];
        if (i + 2 < m) t -= U2[i] * x[i + 2];
        x[i] = t / (std::fabs(D[i]) < tiny ? (D[i] < 0 ? -tiny : tiny) : D[i]);
      }
      for (int j = 0; j < k; ++j) {
        if (lam[k] - lam[j] > 1e-3 * tn) continue;
        double dot = 0.0;
        for (int i = 0; i < m; ++i) dot += Z[(size_t)i * K + j] * x[i];
        for (int i = 0; i < m; ++i) x[i] -= dot * Z[(size_t)i * K + j];
      }
      double nrm = 0.0;
      for (int i = 0; i < m; ++i) nrm += x[i] * x[i];
      nrm = std::sqrt(nrm);
      if (nrm == 0.0) nrm = 1.0;
      for (int i = 0; i < m; ++i) x[i] /= nrm;
    }
    double big = 0.0;
    int at = 0;
    for (int i = 0; i < m; ++i)
      if (std::fabs(x[i]) > big) { big = std::fabs(x[i]); at = i; }
    const double sg = x[at] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < m; ++i) Z[(size_t)i * K + k] = sg * x[i];
  }
}

}  // namespace gato

using namespace gato;

struct gato_fock {
  int dtype, N, order, boundary, n_orb;
  double h;
  long ld, n;
  const void* V_eff;
  gato_poisson_plan* pp;
  const void* orbitals;
  gato_nl_plan* nl;
  // compressed exchange (gato_fock_set_ace): when set, the exchange term is
  // -xi (dV xi^T psi) instead of the Poisson solves
  const void* ace_xi = nullptr;
  int ace_n = 0;
  double* ace_coef = nullptr;
};

extern "C" const char* gato_last_error(void) { return g_err; }

extern "C" void gato_probe_reset(int kind) {
  g_probe.kind = kind & 0xff;
  g_probe.elem = (kind >> 8) & 0xff;
  g_probe.used = 0;
  g_probe.launches = 0;
  g_probe.bytes = 0.0;
}

extern "C" int gato_probe_read(long* launches, double* ms, double* bytes) {
  clear_ok();
  GATO_REQUIRE(launches && ms && bytes, GATO_ERR_VALUE, "null output pointer");
  double tot = 0.0;
  for (size_t i = 0; i + 1 < g_probe.used; i += 2) {
    GATO_HIP_CHECK(hipEventSynchronize(g_probe.ev[i + 1]));
    float t = 0.f;
    GATO_HIP_CHECK(hipEventElapsedTime(&t, g_probe.ev[i], g_probe.ev[i + 1]));
    tot += t;
  }
  *launches = g_probe.launches;
  *ms = tot;
  *bytes = g_probe.bytes;
  return GATO_OK;
}
extern "C" int gato_abi_version(void) { return GATO_ABI_VERSION; }

extern "C" int gato_tridiag_eigh(int m, const double* alpha, const double* beta, int K,
                                 double* evals, double* evecs) {
  clear_ok();
  GATO_REQUIRE(m >= 1 && K >= 1 && K <= m && alpha && evals && evecs && (m == 1 || beta),
               GATO_ERR_VALUE, "bad arguments m=%d K=%d", m, K);
  // the K lowest eigenvalues by bisection, their vectors by inverse iteration:
  // O(K m) work instead of the O(m^3) of accumulating all m eigenvectors
  std::vector<double> bb(m > 1 ? m - 1 : 1, 0.0);
  for (int i = 0; i + 1 < m; ++i) bb[i] = beta[i];
  tridiag_lowest_eigvals(m, alpha, bb.data(), K, evals);
  tridiag_inverse_iteration(m, alpha, bb.data(), K, evals, evecs);
  return GATO_OK;
}

extern "C" int gato_fock_create(int dtype, int N, double h, int order, int boundary,
                                const void* V_eff, gato_poisson_plan* pplan, const void* orbitals,
                                long ld, int n_orb, gato_nl_plan* nl, gato_fock** fock) {
  clear_ok();
  GATO_REQUIRE(fock != nullptr, GATO_ERR_VALUE, "null fock pointer");
  *fock = nullptr;
  GATO_REQUIRE(dtype == GATO_F32 || dtype == GATO_F64, GATO_ERR_VALUE, "unknown dtype %d", dtype);
  GATO_REQUIRE(order == 2 || order == 4, GATO_ERR_VALUE, "Unknown order %d; expected 2 or 4.",
               order);
  GATO_REQUIRE(boundary == GATO_DIRICHLET || boundary == GATO_PERIODIC, GATO_ERR_VALUE,
               "Unknown boundary %d; expected 'dirichlet' or 'periodic'.", boundary);
  GATO_REQUIRE(N >= 1 && h > 0, GATO_ERR_VALUE, "bad grid");
  if (pplan) {
    GATO_REQUIRE(orbitals && n_orb >= 1, GATO_ERR_VALUE, "exchange needs orbitals");
    GATO_REQUIRE(poisson_plan_dtype(pplan) == dtype && poisson_plan_N(pplan) == N, GATO_ERR_VALUE,
                 "Poisson plan does not match the operator's dtype/grid");
  }
  if (nl) GATO_REQUIRE(nl_plan_dtype(nl) == dtype, GATO_ERR_VALUE, "nonlocal plan dtype mismatch");
  auto* f = new gato_fock();
  f->dtype = dtype;
  f->N = N;
  f->h = h;
  f->order = order;
  f->boundary = boundary;
  f->V_eff = V_eff;
  f->pp = pplan;
  f->orbitals = orbitals;
  f->ld = ld;
  f->n_orb = pplan ? n_orb : 0;
  f->nl = nl;
  f->n = (long)N * N * N;
  *fock = f;
  return GATO_OK;
}

extern "C" int gato_fock_destroy(gato_fock* f) {
  delete f;
  return GATO_OK;
}

namespace gato {
int block_dot(int dtype, long n, int nk, const void* V, long ldv, const void* X, long ldx, int nb,
              double scale, double* out, void* ws, hipStream_t s);
int block_update(int dtype, long n, int nk, const void* V, long ldv, const double* C,
                 const void* X, long ldx, int nb, void* Y, long ldy, double scale, double* gram,
                 void* ws, hipStream_t s);
int block_scale(int dtype, long n, const void* X, long ldx, int nb, const double* Rinv, void* Y,
                long ldy, double scale, double* gram, void* ws, hipStream_t s);
int ace_factor(int dtype, long len, int n, void* xi, double* coef, void* ws, hipStream_t s);

int fock_dtype(const gato_fock* f) { return f->dtype; }
long fock_n(const gato_fock* f) { return f->n; }

// out -= xi C, C = dV xi^T psi (nb fields): the compressed exchange, two
// passes over the n_orb xi fields and the nb inputs.
int ace_subtract(gato_fock* f, int nb, const void* psi, void* out, void* ws, hipStream_t s) {
  const double dV = f->h * f->h * f->h;
  int st = block_dot(f->dtype, f->n, f->ace_n, f->ace_xi, f->n, psi, f->n, nb, dV, f->ace_coef,
                     ws, s);
  if (st) return st;
  return block_update(f->dtype, f->n, f->ace_n, f->ace_xi, f->n, f->ace_coef, out, f->n, nb, out,
                      f->n, 0.0, nullptr, ws, s);
}

int fock_apply(gato_fock* f, const void* psi, void* out, const void* sub, const double* sub_coef,
               void* ws, hipStream_t s) {
  const bool ace = f->ace_xi != nullptr;
  int acc = 0;
  if (f->pp && !ace) {
    int st = poisson_exchange(f->pp, 1, psi, 0, f->orbitals, f->ld, f->n_orb, -1.0, 0, out, 0, s);
    if (st) return st;
    acc = 1;
  }
  int st = stencil_apply(f->dtype, f->N, 1, f->h, f->order, f->boundary, -0.5, psi, f->V_eff, acc,
                         sub, sub_coef, out, s);
  if (st) return st;
  if (f->nl) {
    st = nl_apply_internal(f->nl, 1, psi, f->n, out, f->n, ws, s);
    if (st) return st;
  }
  return ace ? ace_subtract(f, 1, psi, out, ws, s) : GATO_OK;
}

// F on nb contiguous fields (psi + c*n -> out + c*n): the exchange solves of
// all nb fields share launches (one kernel-transform read per launch instead
// of one per field), the stencil and the nonlocal term run once over the
// batch.  Each output equals fock_apply of its field; with a compressed
// exchange and with_ace false, its subtraction is left to the caller.
int fock_apply_block(gato_fock* f, int nb, const void* psi, void* out, void* ws, hipStream_t s,
                     bool with_ace) {
  const bool ace = f->ace_xi != nullptr;
  int acc = 0;
  if (f->pp && !ace) {
    int st = poisson_exchange(f->pp, nb, psi, f->n, f->orbitals, f->ld, f->n_orb, -1.0, 0, out,
                              f->n, s);
    if (st) return st;
    acc = 1;
  }
  int st = stencil_apply(f->dtype, f->N, nb, f->h, f->order, f->boundary, -0.5, psi, f->V_eff,
                         acc, nullptr, nullptr, out, s);
  if (st) return st;
  if (f->nl) {
    st = nl_apply_internal(f->nl, nb, psi, f->n, out, f->n, ws, s);
    if (st) return st;
  }
  return ace && with_ace ? ace_subtract(f, nb, psi, out, ws, s) : GATO_OK;
}

// The compressed exchange's field count (0 without one) and fields.
int fock_ace(const gato_fock* f, const void** xi) {
  *xi = f->ace_xi;
  return f->ace_xi ? f->ace_n : 0;
}
}  // namespace gato

// Adaptively compressed exchange (L. Lin, J. Chem. Theory Comput. 12, 2242
// (2016)): W = K phi over the operator's n orbitals, M = dV phi^T W = R^T R
// (M is the exchange operator restricted to the orbitals: symmetric positive
// definite), xi = W R^{-1}.  Then xi (dV xi^T psi) = W M^{-1} (dV W^T psi)
// equals K psi for every psi in the span of the orbitals.  Everything stays
// on the stream (the n x n Cholesky in one wave); coef: device scratch of
// 4 n^2 doubles.
extern "C" int gato_ace_build(gato_fock* f, void* xi, double* coef, void* ws, void* stream) {
  clear_ok();
  GATO_REQUIRE(f && xi && coef && ws, GATO_ERR_VALUE, "bad arguments");
  GATO_REQUIRE(f->pp && f->orbitals && f->n_orb >= 1 && f->n_orb <= 8, GATO_ERR_VALUE,
               "compressed exchange needs the operator's exchange orbitals (1..8 of them)");
  hipStream_t s = as_stream(stream);
  const int n = f->n_orb;
  const double dV = f->h * f->h * f->h;
  // W = +K phi into xi (the fields of xi are contiguous, ld = f->n)
  int st = poisson_exchange(f->pp, n, f->orbitals, f->ld, f->orbitals, f->ld, n, 1.0, 0, xi, f->n,
                            s);
  if (st) return st;
  st = block_dot(f->dtype, f->n, n, f->orbitals, f->ld, xi, f->n, n, dV, coef, ws, s);
  if (st) return st;
  return ace_factor(f->dtype, f->n, n, xi, coef, ws, s);
}

extern "C" int gato_fock_set_ace(gato_fock* f, const void* xi, int n_xi, double* coef) {
  clear_ok();
  GATO_REQUIRE(f, GATO_ERR_VALUE, "null operator");
  if (!xi) {
    f->ace_xi = nullptr;
    f->ace_n = 0;
    f->ace_coef = nullptr;
    return GATO_OK;
  }
  GATO_REQUIRE(n_xi >= 1 && coef, GATO_ERR_VALUE, "compressed exchange needs n_xi >= 1 and a "
               "coefficient buffer");
  f->ace_xi = xi;
  f->ace_n = n_xi;
  f->ace_coef = coef;
  return GATO_OK;
}

extern "C" int gato_fock_apply(gato_fock* f, const void* psi, void* out, const void* sub,
                               const double* sub_coef, void* ws, void* stream) {
  clear_ok();
  GATO_REQUIRE(f && psi && out, GATO_ERR_VALUE, "bad arguments");
  GATO_REQUIRE(!f->nl || ws, GATO_ERR_VALUE, "nonlocal apply needs a workspace");
  return fock_apply(f, psi, out, sub, sub_coef, ws, as_stream(stream));
}

// scratch layout (doubles): coef[m+1] (this step's Gram-Schmidt coefficients) |
// alpha[m] | beta[m] | S[m*K] | norm0[1]
extern "C" size_t gato_lanczos_scratch_bytes(int m, int K) {
  return sizeof(double) * ((size_t)(m + 1) + 2 * (size_t)m + (size_t)m * K + 8);
}

namespace {
// The device part of gato_lanczos: v_0 = psi0 / ||psi0||, then m steps, with
// alpha and beta left in the scratch.  No host synchronisation and no
// allocation, so the whole sequence can be recorded into a graph.
int lanczos_steps(gato_fock* f, const void* psi0, int m, int K, double dV, void* V_buf,
                  void* scratch, void* ws, hipStream_t s) {
  const int dt = f->dtype;
  const long n = f->n;
  const size_t es = dt == GATO_F32 ? 4 : 8;
  char* Vb = (char*)V_buf;
  auto vec = [&](int j) { return (void*)(Vb + (size_t)j * n * es); };
  double* coef = (double*)scratch;
  double* d_alpha = coef + (m + 1);
  double* d_beta = d_alpha + m;
  double* d_S = d_beta + m;
  double* d_norm0 = d_S + (size_t)m * K;
  int st;
  // v0 = psi0 / ||psi0||  (lanczos.py:134): copy + norm, then scale
  st = cgs_update(dt, n, 0, nullptr, 0, nullptr, nullptr, nullptr, psi0, vec(0), dV, d_norm0, ws,
                  s);
  if (st) return st;
  st = scale_by_inverse(dt, n, vec(0), d_norm0, vec(0), s);
  if (st) return st;
  for (int j = 0; j < m; ++j) {
    void* w = vec(j + 1);
    // w = F v_j - beta_{j-1} v_{j-1}   (lanczos.py:144-148)
    st = fock_apply(f, vec(j), w, j > 0 ? vec(j - 1) : nullptr, j > 0 ? d_beta + (j - 1) : nullptr,
                    ws, s);
    if (st) return st;
    // alpha_j = <v_j, w> (lanczos.py:150) and c_k = <v_k, w - alpha_j v_j>,
    // k <= j, in one pass over the basis (c_k = <v_k,w> - alpha_j <v_k,v_j>).
    // After the three-term recurrence the c_k are rounding-level, so one
    // sweep equals the reference's modified Gram-Schmidt to second order
    // (lanczos.py:151-159)
    st = cgs_coefficients(dt, n, j + 1, V_buf, n, w, dV, coef, d_alpha + j, ws, s);
    if (st) return st;
    // w -= alpha_j v_j + sum_k c_k v_k, leaving the wave partials of ||w||^2
    // (lanczos.py:151-163); then beta_j = ||w|| and v_{j+1} = w / beta_j (zero
    // if beta_j == 0, lanczos.py:165-167) in one launch
    long np = 0;
    st = cgs_update_partials(dt, n, j + 1, V_buf, n, coef, vec(j), d_alpha + j, w, w, &np, ws, s);
    if (st) return st;
    st = scale_by_partials(dt, n, w, ws, np, dV, d_beta + j, w, s);
    if (st) return st;
  }
  return GATO_OK;
}
}  // namespace

extern "C" int gato_lanczos_steps(gato_fock* f, const void* psi0, int m, int K, double dV,
                                  void* V_buf, void* scratch, void* ws, void* stream) {
  clear_ok();
  GATO_REQUIRE(f && psi0 && V_buf && scratch && ws, GATO_ERR_VALUE, "bad arguments");
  GATO_REQUIRE(m >= 1 && K >= 1, GATO_ERR_VALUE, "bad sizes m=%d K=%d", m, K);
  return lanczos_steps(f, psi0, m, K, dV, V_buf, scratch, ws, as_stream(stream));
}

extern "C" int gato_lanczos(gato_fock* f, const void* psi0, int m, int K, double dV, void* V_buf,
                            void* scratch, double* alphas, double* betas, double* evals,
                            void* eigenstates, void* ws, void* stream) {
  clear_ok();
  GATO_REQUIRE(f && psi0 && V_buf && scratch && alphas && betas && evals && eigenstates && ws,
               GATO_ERR_VALUE, "bad arguments");
  GATO_REQUIRE(m >= 1 && K >= 1, GATO_ERR_VALUE, "bad sizes m=%d K=%d", m, K);
  hipStream_t s = as_stream(stream);
  const int dt = f->dtype;
  const long n = f->n;
  double* d_alpha = (double*)scratch + (m + 1);
  double* d_beta = d_alpha + m;
  double* d_S = d_beta + m;
  int st = lanczos_steps(f, psi0, m, K, dV, V_buf, scratch, ws, s);
  if (st) return st;
  GATO_HIP_CHECK(hipMemcpyAsync(alphas, d_alpha, sizeof(double) * m, hipMemcpyDeviceToHost, s));
  GATO_HIP_CHECK(hipMemcpyAsync(betas, d_beta, sizeof(double) * m, hipMemcpyDeviceToHost, s));
  GATO_HIP_CHECK(hipStreamSynchronize(s));
  const int KK = K < m ? K : m;
  std::vector<double> S((size_t)m * KK);
  st = gato_tridiag_eigh(m, alphas, betas, KK, evals, S.data());
  if (st) return st;
  GATO_HIP_CHECK(hipMemcpyAsync(d_S, S.data(), sizeof(double) * S.size(), hipMemcpyHostToDevice, s));
  st = ritz_combine(dt, n, m, KK, V_buf, n, d_S, eigenstates, n, s);
  if (st) return st;
  // the host vector S must outlive the copy
  GATO_HIP_CHECK(hipStreamSynchronize(s));
  return GATO_OK;
}
