/*
 * gato_amd — C ABI of the MI355X (gfx950) real-space RHF hot path.
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t
 * (passed as void*), is stream-ordered, and returns 0 (GATO_OK) or a nonzero
 * status; gato_last_error() returns the message of the last failure on the
 * calling host thread.  Python callers map GATO_ERR_VALUE to ValueError,
 * GATO_ERR_KEY to KeyError and the rest to RuntimeError, which is how the
 * reference package reports the same conditions.
 *
 * Fields on an N^3 grid are C-ordered [x][y][z] (z contiguous), the layout of
 * meshgrid(indexing="ij") in the reference (src/gato/grid.py:46-49).  Stacks
 * of fields are [k][x][y][z] with a leading-dimension `ld` (elements).
 * dtype selects float (GATO_F32) or double (GATO_F64) storage; reductions
 * always accumulate in double.
 *
 * The reference is pure JAX and has no FFI; each declaration below names the
 * reference Python function whose work it carries (file:line under the
 * reference's src/gato/).
 */
#ifndef GATO_AMD_H
#define GATO_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GATO_ABI_VERSION 1

/* status codes */
#define GATO_OK 0
#define GATO_ERR_VALUE 1       /* bad argument: the reference raises ValueError */
#define GATO_ERR_KEY 2         /* unknown element: the reference raises KeyError */
#define GATO_ERR_HIP 3         /* HIP runtime failure */
#define GATO_ERR_UNSUPPORTED 4 /* size or option outside what the kernels support */

/* dtype */
#define GATO_F32 0
#define GATO_F64 1

/* boundary (operators.py:40-64) */
#define GATO_DIRICHLET 0
#define GATO_PERIODIC 1

/* local-potential centre kinds */
#define GATO_CENTER_SOFT_COULOMB 0 /* -Z/sqrt(|r-R|^2+eps^2): potentials.py:16-61 */
#define GATO_CENTER_HGH_LOCAL 1    /* HGH local part: pseudopotentials.py:201-227 */
#define GATO_CENTER_HARMONIC 2     /* 0.5 w^2 |r-R|^2: potentials.py:64-71 */
#define GATO_CENTER_SLATER_1S 3    /* Z^1.5 exp(-Z sqrt(|r-R|^2+s^2)): init_lcao ansatz/grid.py:53-76 */
#define GATO_CENTER_PARAMS 10      /* doubles per centre, see gato_local_potential */

typedef struct gato_poisson_plan gato_poisson_plan;
typedef struct gato_nl_plan gato_nl_plan;
typedef struct gato_fock gato_fock;

const char* gato_last_error(void);
int gato_abi_version(void);

/* Measurement hooks (no reference counterpart): gato_probe_reset(kind) clears
 * the counters and, for kind != 0, brackets every later launch of that kernel
 * kind with HIP events on its stream (1 = reorthogonalisation update,
 * 2 = reorthogonalisation coefficients, 3 = strided FFT passes, 4 = z FFT
 * passes, 5 = block-Lanczos projection on the matrix cores, 6 = block
 * update).  kind | (4 << 8) or kind | (8 << 8) restricts the probe to the
 * launches on float or on double fields.  gato_probe_read() returns the
 * launch count, summed GPU milliseconds and summed algorithmic bytes of the
 * bracketed launches. */
void gato_probe_reset(int kind);
int gato_probe_read(long* launches, double* ms, double* bytes);

/* Stencil variant choice (no reference counterpart).  The vector stencil path
 * has a one-plane kernel and an x-marching kernel with XC-plane chunks
 * (XC = 4, 8, 16, 32); all evaluate the same expression in the same order.
 * With autotuning on (the default), the first gato_stencil_apply of a shape
 * (dtype, N, batch, order, boundary, and which of V / accumulate / sub are
 * given) times the candidates on scratch fields on the caller's stream and
 * keeps the fastest for the process.  Calls under stream capture, and all
 * calls with gato_set_autotune(0), take a measured static rule.
 * gato_stencil_force_variant(v) forces variant v for every later call; v < 0
 * restores the choice.  v = 0 is the one-plane kernel; v = XC + 100 * (PF - 1)
 * + 1000 * O + 10000 * Y is the marching kernel with XC-plane chunks, PF = 1..3
 * planes of register prefetch, block order O (0: each XCD takes a contiguous
 * range of (chunk, tile, field); 1: each XCD takes an eighth of the spatial
 * units for every field, so the shared potential is read once per unit) and
 * Y = 1 for one y row per thread in 512-thread blocks (0: two rows, 256).
 * Adding 100000 * K picks the kernel: K = 0 plain loads, K = 1 buffer-resource
 * loads (same tiles), K = 2 the barrier-free kernel in which every wave
 * streams its own strip of a plane (no LDS; Y = 0, 1, 2 for strips of 256,
 * 512, 128 vectors per block; PF = 1..2). */
void gato_set_autotune(int on);
void gato_stencil_force_variant(int variant);

/* FFT pencils per block (no reference counterpart).  The four-step middle
 * passes of a float Poisson plan with 2N = 128, 160, 192, 256 or 320 run
 * with 32 or 16 pencils per block; gato_poisson_plan_create times both on the
 * plan's scratch (unless gato_set_autotune(0)) and keeps the faster.  Both
 * give the same bits.  gato_poisson_force_pencils(16 or 32) forces one for
 * every later launch; any other value restores the choice. */
void gato_poisson_force_pencils(int pencils);

/* Bytes of device workspace that every reducing entry point needs. */
size_t gato_reduce_workspace_bytes(void);

/* out[b] = (accumulate ? out[b] : 0) + c_lap * lap(psi[b]) + (V ? V*psi[b] : 0)
 *          - (sub ? sub_coef[0] * sub[b] : 0)
 * 7-point (order 2) or 13-point (order 4) Laplacian, Dirichlet zero padding or
 * periodic wrap.  Replaces operators.laplacian / kinetic (operators.py:40-130),
 * Hamiltonian.apply (hamiltonian.py:25-27) and the local part of
 * _fock_apply_kernel (scf.py:106-109).  sub_coef is a device pointer (the
 * Lanczos beta_{j-1} of lanczos.py:146-148). */
int gato_stencil_apply(int dtype, int N, int batch, double h, int order, int boundary,
                       double c_lap, const void* psi, const void* V, int accumulate,
                       const void* sub, const double* sub_coef, void* out, void* stream);

/* One field: out = c_lap * lap(psi) + (V ? V*psi : 0) - (sub ? sub_coef[0] * sub : 0),
 * then out /= sqrt(dV <out, out>) (out = 0 when the norm is 0); norm_out (a
 * device double, or NULL) receives the norm.  The Euler step and
 * renormalisation of imaginary time (imag_time.py:47-55) in two launches: the
 * stencil writes block partials of <out, out> to the workspace and one
 * launch sums them (fixed order) and scales.  Fields that do not take the
 * vector stencil path run the stencil and gato_normalize separately. */
int gato_stencil_normalize(int dtype, int N, double h, int order, int boundary, double c_lap,
                           const void* psi, const void* V, const void* sub,
                           const double* sub_coef, void* out, double dV, double* norm_out,
                           void* ws, void* stream);

/* n_steps >= 0 imaginary-time steps psi <- normalise(psi + (dt/2) lap(psi) +
 * Vdt psi) (the loop of imag_time.py:47-62 with H psi = -lap/2 psi + V psi;
 * the caller passes Vdt = -dt V).  a holds the normalised input; the steps
 * alternate a -> b -> a ..., so the result is in b when n_steps is odd, in a
 * when it is even, normalised; norm_out (a device double, or NULL) receives
 * the norm of the last update before its normalisation.  One stencil launch
 * per step: each scales its result by the inverse norm of its input, summed
 * from the block partials the previous launch wrote, so the renormalisation
 * costs no pass of its own; one final pass normalises.  Uses the whole
 * partials area of ws.  Fields that do not take the vector stencil path run
 * the stencil and gato_normalize per step. */
int gato_imag_time_steps(int dtype, int N, double h, int order, int boundary, double dt,
                         const void* Vdt, void* a, void* b, int n_steps, double dV,
                         double* norm_out, void* ws, void* stream);

/* out[k] = scale * sum_i A[k*lda + i] * B'[k*ldb + i] * (C ? C[i] : 1), k < nk,
 * with B' = B - (S ? s_coef[0] * S : 0) (S only with ldb = 0, which broadcasts
 * one B over all k); double accumulation (compensated), out a device pointer.
 * Replaces grid.integrate / inner_product / norm_sq (grid.py:57-69), alpha_j
 * of lanczos.py:150 and the reorthogonalisation coefficients
 * <v_k, w - alpha_j v_j> of lanczos.py:151-159. */
int gato_dot_batched(int dtype, long n, int nk, const void* A, long lda, const void* B, long ldb,
                     const void* C, const void* S, const double* s_coef, double scale,
                     double* out, void* ws, void* stream);

/* gato_dot_batched without the workspace argument.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It allocates nothing: its reduction partials live in a static device array
 * of 2 MiB (262144 doubles, plus a 256-byte header) that each GPU the library
 * runs on holds until the process ends.  Same kernels, same launch choices and
 * same summation order as gato_dot_batched, so the results are identical.
 * Calls whose partials would exceed the array fail with GATO_ERR_UNSUPPORTED
 * before any launch: more than 256 vectors of 128^3 floats (128 of 128^3
 * doubles) against one broadcast vector, for example.  All calls on one GPU
 * share its array: calls that overlap on one GPU (two streams, two host
 * threads) race.  The stream must belong to the current device. */
int gato_dot_batched_noalloc(int dtype, long n, int nk, const void* A, long lda, const void* B,
                             long ldb, const void* C, const void* S, const double* s_coef,
                             double scale, double* out, void* stream);

/* w_out = w - (S ? s_coef[0] * S : 0) - sum_{k<nk} coef[k] * V[k*ldv + .];
 * beta_out[0] = sqrt(scale*||w_out||^2) (w_out may alias w).  The update
 * w -= alpha_j v_j of lanczos.py:151 fused with the reorthogonalisation sweep
 * of lanczos.py:155-159 and beta_j = ||w|| of lanczos.py:163. */
int gato_cgs_update(int dtype, long n, int nk, const void* V, long ldv, const double* coef,
                    const void* S, const double* s_coef, const void* w, void* w_out,
                    double scale, double* beta_out, void* ws, void* stream);

/* Lanczos projections of step j = nk - 1 in one pass over V (v_j = V[(nk-1)*ldv]):
 *   alpha[0] = scale * <v_j, w>                               (lanczos.py:150)
 *   coef[k]  = scale * <v_k, w> - alpha[0] * scale * <v_k, v_j>, k < nk
 *            = scale * <v_k, w - alpha v_j>    (the coefficients of the
 *              reorthogonalisation sweep of lanczos.py:151-159)
 * coef and alpha are device pointers; feed them to gato_cgs_update with
 * S = v_j, s_coef = alpha. */
int gato_cgs_coefficients(int dtype, long n, int nk, const void* V, long ldv, const void* w,
                          double scale, double* coef, double* alpha, void* ws, void* stream);

/* v = beta[0] > 0 ? w / beta[0] : 0  (lanczos.py:165-167) */
int gato_scale_by_inverse(int dtype, long n, const void* w, const double* beta, void* v,
                          void* stream);

/* v = w / sqrt(scale * <w, w>) (v = 0 when the norm is 0); norm_out (a device
 * double, or NULL) receives sqrt(scale * <w, w>).  v may alias w.  Two
 * launches: wave partials of <w, w>, then one kernel that sums them (every
 * block in the same fixed order) and scales.  Replaces grid.normalize
 * (grid.py:72-73); it is also the renormalisation of imaginary time
 * (imag_time.py:51, :55) and of the Lanczos start vector (lanczos.py:134). */
int gato_normalize(int dtype, long n, const void* w, double scale, double* norm_out, void* v,
                   void* ws, void* stream);

/* out[k] = sum_{j<m} S[j*K + k] * V[j*ldv + .], k < K <= 16  (lanczos.py:193-194) */
int gato_ritz_combine(int dtype, long n, int m, int K, const void* V, long ldv, const double* S,
                      void* out, long ldo, void* stream);

/* Host routine: the K lowest eigenpairs of the m x m symmetric tridiagonal with
 * diagonal alpha[0..m) and off-diagonal beta[0..m-1) (Sturm-count bisection for
 * the eigenvalues, inverse iteration for the vectors, reorthogonalised within
 * clusters).  evals ascending; evecs row-major m x K (evecs[j*K+k]), each
 * column's largest component positive.  Replaces the jnp.linalg.eigh of
 * lanczos.py:176-180. */
int gato_tridiag_eigh(int m, const double* alpha, const double* beta, int K, double* evals,
                      double* evecs);

/* rho_out = mixing * 2 * sum_k |phi_k|^2 + (1 - mixing) * rho_old   (rho_old may be
 * NULL when mixing == 1).  scf.py:61-63 and :422-423. */
int gato_density_mix(int dtype, long n, int n_orb, const void* orbitals, long ld, double mixing,
                     const void* rho_old, void* rho_out, void* stream);

/* y = a*x + b*y (elementwise helper used by the SCF driver). */
int gato_axpby(int dtype, long n, double a, const void* x, double b, void* y, void* stream);

/* Converts between float and double fields (in_dtype -> out_dtype). */
int gato_convert(int in_dtype, const void* x, int out_dtype, void* y, long n, void* stream);

/* V = (accumulate ? V : 0) + sum_c f_c(|r - R_c|) on the cell-centred grid of
 * grid.py:42-49.  centers holds GATO_CENTER_PARAMS doubles per centre (host):
 * [kind, Rx, Ry, Rz, a0, a1, a2, a3, a4, a5] with
 *   SOFT_COULOMB: a0 = Z, a1 = eps               (potentials.py:16-61)
 *   HGH_LOCAL:    a0 = Z_ion, a1 = r_loc, a2..a5 = C1..C4 (pseudopotentials.py:201-253)
 *   HARMONIC:     a0 = omega                     (potentials.py:64-71)
 *   SLATER_1S:    a0 = Z, a1 = softening s       (ansatz/grid.py:53-76)
 * Evaluated in double, stored as dtype. */
int gato_local_potential(int dtype, int N, double L, int n_centers, const double* centers,
                         int accumulate, void* V, void* stream);

/* grad[c*3 + a] = dV_c/dR_{c,a} integrated against rho:  sum_r rho(r) dV_c(r)/dR * dV.
 * The Hellmann-Feynman local term that jax.grad(bo_energy) produces through
 * hgh_local_potential / multi_center_softened_coulomb (water.py:181-224). */
int gato_local_potential_grad(int dtype, int N, double L, int n_centers, const double* centers,
                              const void* rho, double* grad, void* ws, void* stream);

/* --- Hockney free-space Poisson (solvers/poisson.py) --------------------- */

/* Plan for one grid: N, spacing h, kernel softening epsilon, and room for up to
 * max_pairs complex transforms per launch (each carries two real solves).
 * Builds the transform of the softened kernel of poisson.py:29-44 once, in
 * double, on the device. */
int gato_poisson_plan_create(int dtype, int N, double h, double epsilon, int max_pairs,
                             gato_poisson_plan** plan, void* stream);
int gato_poisson_plan_destroy(gato_poisson_plan* plan);

/* J[k] = hartree_potential(rho[k]) for k < n_rho (poisson.py:47-79). */
int gato_hartree_potential(gato_poisson_plan* plan, int n_rho, const void* rho, long ld, void* J,
                           long ldj, void* stream);

/* E[k] = 0.5 * int rho[k] J[rho[k]] dV  (poisson.py:82-92); E device pointer. */
int gato_hartree_energy(gato_poisson_plan* plan, int n_rho, const void* rho, long ld, double* E,
                        void* ws, void* stream);

/* out = (accumulate ? out : 0) + coef * sum_j phi_j * J[phi_j psi]   (scf.py:66-85;
 * coef = -1 gives the -K psi term of scf.py:110). */
int gato_exchange_apply(gato_poisson_plan* plan, const void* psi, const void* orbitals, long ld,
                        int n_orb, double coef, int accumulate, void* out, void* stream);

/* out[p] = (ij|ij) = int phi_i phi_j J[phi_i phi_j] dV over the n(n+1)/2 pairs i <= j
 * in row-major upper-triangle order.  E_x = -(sum_i (ii|ii) + 2 sum_{i<j} (ij|ij))
 * equals -sum_i <phi_i|K phi_i> of scf.py:225-229. */
int gato_exchange_pair_integrals(gato_poisson_plan* plan, const void* orbitals, long ld, int n_orb,
                                 double* out, void* ws, void* stream);

/* --- HGH separable nonlocal part (pseudopotentials.py:265-383) ----------- */

/* n_proj projector channels; channel p sits on atom proj_atom[p] with angular
 * momentum proj_ell[p] (<= 1), real-harmonic index proj_m[p], radial index
 * proj_i[p] (1-based) and width proj_r[p]; hmat is the n_proj x n_proj coupling
 * (row-major; nonzero only inside one (atom, l, m) block).  Values are sampled
 * on a per-atom box holding every point where a projector exceeds
 * cutoff_tol * its peak (the reference samples the full grid). */
int gato_nl_plan_create(int dtype, int N, double L, int n_atoms, const double* atom_pos,
                        int n_proj, const int* proj_atom, const int* proj_ell, const int* proj_m,
                        const int* proj_i, const double* proj_r, const double* hmat,
                        double cutoff_tol, gato_nl_plan** plan, void* stream);
int gato_nl_plan_destroy(gato_nl_plan* plan);
/* Number of grid points in the plan's boxes (diagnostics). */
long gato_nl_plan_points(const gato_nl_plan* plan);

/* out[v] += V_nl psi[v] for v < nv  (hgh_nonlocal_apply, pseudopotentials.py:360-383). */
int gato_nl_apply(gato_nl_plan* plan, int nv, const void* psi, long ld, void* out, long ldo,
                  void* ws, void* stream);

/* E[v] = <psi_v | V_nl | psi_v>, device pointer (scf.py:217-218). */
int gato_nl_expectation(gato_nl_plan* plan, int nv, const void* psi, long ld, double* E, void* ws,
                        void* stream);

/* grad[a*3 + c] = d/dR_a sum_v occ * <psi_v | V_nl(R) | psi_v>  (device pointer; the
 * nonlocal Hellmann-Feynman term of water.py:221-241). */
int gato_nl_grad(gato_nl_plan* plan, int nv, const void* psi, long ld, double occ, double* grad,
                 void* ws, void* stream);

/* --- Fock operator and Lanczos (scf.py:88-158, solvers/lanczos.py) ------- */

/* F psi = c_lap lap psi + V_eff psi + V_nl psi - sum_j phi_j J[phi_j psi].
 * V_eff = V_ext + V_H is supplied pre-summed; pplan/orbitals may be NULL (no
 * exchange, i.e. Hamiltonian.apply), nl may be NULL (all-electron). */
int gato_fock_create(int dtype, int N, double h, int order, int boundary, const void* V_eff,
                     gato_poisson_plan* pplan, const void* orbitals, long ld, int n_orb,
                     gato_nl_plan* nl, gato_fock** fock);
int gato_fock_destroy(gato_fock* fock);

/* out = F psi - (sub ? sub_coef[0] * sub : 0)  (scf.py:146-158). */
int gato_fock_apply(gato_fock* fock, const void* psi, void* out, const void* sub,
                    const double* sub_coef, void* ws, void* stream);

/* Adaptively compressed exchange (L. Lin, J. Chem. Theory Comput. 12, 2242
 * (2016)) of the operator's exchange: W = K phi over its n_orb (<= 8)
 * orbitals (the n_orb x n_orb exchange solves of scf.py:66-85), M = dV
 * phi^T W = R^T R, xi = W R^{-1} into xi (n_orb contiguous fields).  Then
 * xi (dV xi^T psi) equals K psi for psi in the span of the orbitals.  coef:
 * device scratch of >= 4 n_orb^2 doubles.  Stream-ordered (the n x n
 * Cholesky runs in one wave on the device).  No reference counterpart: the
 * reference applies the exact exchange in every Lanczos step
 * (scf.py:146-158). */
int gato_ace_build(gato_fock* fock, void* xi, double* coef, void* ws, void* stream);

/* Later applies of fock (single, block, and the native Lanczos loops) take the
 * exchange term as -xi (dV xi^T psi) over the n_xi fields xi instead of the
 * Poisson solves; xi = NULL restores the exact exchange.  coef: device
 * scratch of >= 8 n_xi doubles, used by every apply (applies on one operator
 * are stream-ordered).  The mixed-precision SCF's float iterations use it. */
int gato_fock_set_ace(gato_fock* fock, const void* xi, int n_xi, double* coef);

/* Bytes of device scratch gato_lanczos needs for m steps and K Ritz vectors
 * (besides the (m+1) x n Krylov buffer). */
size_t gato_lanczos_scratch_bytes(int m, int K);

/* Single-vector Lanczos with full reorthogonalisation (lanczos.py:110-206):
 * m steps from psi0, K lowest Ritz pairs.  V_buf holds (m+1) fields; scratch
 * holds gato_lanczos_scratch_bytes(m, K) bytes.  alphas/betas/evals are host
 * arrays (m, m, K); eigenstates is a device stack of K fields (ld = n). */
int gato_lanczos(gato_fock* fock, const void* psi0, int m, int K, double dV, void* V_buf,
                 void* scratch, double* alphas, double* betas, double* evals, void* eigenstates,
                 void* ws, void* stream);

/* The device part of gato_lanczos (lanczos.py:134-167): v_0 = psi0 / ||psi0||
 * and m steps into V_buf; alpha_j and beta_j stay on the device in the
 * scratch, at doubles (m+1) .. (2m) and (2m+1) .. (3m) of it.  No host
 * synchronisation and no allocation, so the sequence can be recorded into a
 * HIP graph and replayed while the operator's inputs change in place.  The
 * caller then solves the tridiagonal problem (gato_tridiag_eigh) and forms
 * the Ritz vectors (gato_ritz_combine), as gato_lanczos does. */
int gato_lanczos_steps(gato_fock* fock, const void* psi0, int m, int K, double dV, void* V_buf,
                       void* scratch, void* ws, void* stream);

/* ---- Block Lanczos ------------------------------------------------------
 * The reference has no block eigensolver: solvers/lanczos.py:83-107 names
 * block Lanczos as the fix for degenerate manifolds (a single start vector
 * cannot split an exactly degenerate level) and README.md:431 lists it as
 * open work.  These entry points extend the Lanczos recurrence of
 * lanczos.py:140-194 to blocks of nb <= 8 vectors.  Matrices of
 * coefficients are row-major doubles on the device unless marked host. */

/* out[k * nb + c] = scale * <V_k, X_c>, k < nk, c < nb <= 16 (one pass over V). */
int gato_block_dot(int dtype, long n, int nk, const void* V, long ldv, const void* X, long ldx,
                   int nb, double scale, double* out, void* ws, void* stream);

/* Y_c = X_c - sum_{k<nk} C[k * nb + c] V_k (Y may alias X; nb <= 8); when gram
 * is not NULL it receives scale * <Y_c, Y_d> for c <= d, packed row by row
 * (nb (nb+1)/2 doubles). */
int gato_block_update(int dtype, long n, int nk, const void* V, long ldv, const double* C,
                      const void* X, long ldx, int nb, void* Y, long ldy, double scale,
                      double* gram, void* ws, void* stream);

/* Device scratch of gato_block_orthonormalize / gato_block_lanczos_step for
 * steps with up to nk_max basis vectors. */
size_t gato_block_step_scratch_bytes(int nb, int nk_max);

/* Y R = X with Y orthonormal (scale <Y_c, Y_d> = delta_cd) by Cholesky QR
 * applied twice; R (nb x nb, upper) receives the factor.  A column whose
 * remaining Gram pivot is not above rel_tol times its squared norm is
 * dependent on the earlier ones: dropped[c] = 1 and Y_c = 0. */
int gato_block_orthonormalize(int dtype, long n, int nb, const void* X, long ldx, void* Y,
                              long ldy, double scale, double rel_tol, double* R, int* dropped,
                              void* scratch, void* ws, void* stream);

/* Step i of block Lanczos on the basis V (blocks of nb fields, ld = n): on
 * entry blocks 0..i hold Q_0..Q_i and block i+1 holds W = H Q_i; on return
 * block i+1 holds Q_{i+1}.  A receives the symmetrised diagonal block
 * Q_i^T H Q_i, B the factor B_i of
 * W - Q_{i-1} B_{i-1}^T - Q_i A_i - (reorthogonalisation) = Q_{i+1} B_i, all
 * components removed by one projection and one update pass over the basis
 * (Bprev, B_{i-1}, is not read: the projection measures that component). */
int gato_block_lanczos_step(int dtype, long n, int i, int nb, void* V, double scale,
                            double rel_tol, const double* Bprev, double* A, double* B,
                            int* dropped, void* scratch, void* ws, void* stream);

/* Rayleigh-Ritz on the block-tridiagonal T of s steps (host arrays: A and B
 * s blocks of nb x nb, each A_i symmetric with its lower triangle read,
 * dropped (s+1) x nb flags): K lowest eigenvalues and
 * their coefficient vectors S (s*nb x K, host; zero rows for dropped
 * columns). */
int gato_block_ritz(int nb, int s, const double* A, const double* B, const int* dropped, int K,
                    double* evals, double* S);

/* K lowest eigenpairs of a dense symmetric m x m host matrix (Householder
 * tridiagonalisation, bisection and inverse iteration on the tridiagonal, then
 * the reflectors applied back); evecs is m x K, each column's largest
 * component positive.  The dense counterpart of jnp.linalg.eigh
 * (lanczos.py:180) for the block-tridiagonal T. */
int gato_sym_eigh(int m, const double* A, int K, double* evals, double* evecs);

/* Bytes of device scratch gato_block_lanczos needs for nb and s steps. */
size_t gato_block_scratch_bytes(int nb, int s);

/* Block Lanczos on a Fock operator: s steps from the nb fields X0 (ld = ldx),
 * K <= 16 lowest Ritz pairs.  V_buf holds (s+1)*nb fields.  A_out, B_out
 * (s nb x nb blocks), dropped_out ((s+1) x nb) and evals (K) are host
 * arrays; eigenstates is a device stack of K fields. */
int gato_block_lanczos(gato_fock* fock, const void* X0, long ldx, int nb, int s, int K,
                       double dV, double rel_tol, void* V_buf, void* scratch, double* A_out,
                       double* B_out, int* dropped_out, double* evals, void* eigenstates,
                       void* ws, void* stream);

/* gato_block_lanczos with a residual-controlled Krylov dimension: at most
 * s_max steps; from s_min steps on, the K Ritz pairs of T are formed on the
 * host after each step (after every other step beyond 96 vectors) while the
 * next step runs, and the run stops once every pair's residual
 * ||H y - theta y|| = ||B_last S_last|| is below res_tol (res_tol = 0: all
 * s_max steps, the same result as gato_block_lanczos).  The stop keeps the
 * step that was running, so s_used (out) is the check's step count + 1;
 * res_out gets the last check's largest residual (-1 if none ran).  The
 * reference always runs its full Krylov dimension (lanczos.py:123-126); the
 * FP32 phase of the mixed-precision SCF uses this variant. */
int gato_block_lanczos_adaptive(gato_fock* fock, const void* X0, long ldx, int nb, int s_max,
                                int s_min, double res_tol, int K, double dV, double rel_tol,
                                void* V_buf, void* scratch, double* A_out, double* B_out,
                                int* dropped_out, double* evals, void* eigenstates, int* s_used,
                                double* res_out, void* ws, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GATO_AMD_H */
