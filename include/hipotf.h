/*
 * libhipotf -- MI355X (gfx950) on-the-fly FCI / VQE operator library.
 *
 * C ABI: plain pointers and sizes only.  Every vector argument of the device entry points is a
 * DEVICE pointer to float64[dim] in the reference determinant order (|a,b> at ia*Nb + ib,
 * colex-ordered strings, interleaved Jordan-Wigner signs).  Small host-side arrays (Ritz
 * coefficients, heff, thetas, gradients of the fused VQE call) are host pointers, as in the
 * Julia callers.  Nothing here exits the process: failures return NULL / NaN / nonzero and
 * leave a message in hipotf_last_error().  All work is issued on the default stream; entry
 * points that return a value synchronise.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   libcuotf.so      src/extern/cuotf.cu:7-162          (callers: jl/cuda/cunetwork.jl)
 *   libbasis.so      src/include/basis/declare.hpp:4-81  (callers: jl/network.jl:1-116)
 *   libotf_native.so src/include/op/declare.hpp:13-38    (build/destroy of the network handle)
 *   libdiag.so       src/include/diag/declare.hpp:4-75  (callers: jl/davidson.jl)
 */
#ifndef HIPOTF_H
#define HIPOTF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- errors (no reference counterpart: the reference exits via CUDA_CHECK,
 *      src/include/cuda/common.cuh:42-51) ---------------------------------------------------- */
const char *hipotf_last_error(void);
void hipotf_clear_error(void);

/* ---- host basis: replaces libbasis.so (src/include/basis/declare.hpp:4-81) ---------------- */
/* create_basis_manager: basis/declare.hpp:24-33 -> basis.hpp:118-296.  Only the single-block
 * case (orbsym all 0) is accepted; total_sym != 0 yields an empty space as the reference does. */
void *create_basis_manager(int64_t norb, int64_t na, int64_t nb, int64_t total_sym,
                           const int64_t *orbsym, int64_t num_irreps);
int64_t get_subspace_dim(void *basis);          /* basis/declare.hpp:6-9 */
int64_t get_num_symmetry_blocks(void *basis);   /* basis/declare.hpp:11-14 */
void destroy_basis_manager(void *basis);        /* basis/declare.hpp:16-21 */
/* set_det_coeff_f64: basis/declare.hpp:35-45 -> basis.hpp:298-330 (vec is a HOST array) */
void set_det_coeff_f64(void *basis, uint32_t astr, uint32_t bstr, double coeff, double *vec);
/* complex float64, the layout of the reference's complexf64 (real, imaginary) */
typedef struct {
    double re, im;
} HipotfC64;
/* set_det_coeff_c64: basis/declare.hpp:46-56 -> basis.hpp:298-330 on a complex HOST vector
 * (adds coeff at the determinant's index; determinants outside the basis are ignored) */
void set_det_coeff_c64(void *basis, uint32_t astr, uint32_t bstr, HipotfC64 coeff, HipotfC64 *vec);

/* ---- host operator network: replaces build/destroy_network_otf_f64 of libotf_native.so
 *      (src/include/op/declare.hpp:15-38, ham/otf.hpp:223-349).  Same flat SVD-group arrays
 *      as jl/network.jl:313-339 passes. ---------------------------------------------------- */
void *build_network_otf_f64(const int64_t *orbsym, int64_t norb, int64_t ngs, const uint32_t *axs,
                            const uint32_t *bxs, const int64_t *ranks, const int64_t *num_zas,
                            const int64_t *num_zbs, const uint32_t *flat_zas,
                            const uint32_t *flat_zbs, const double *flat_wa,
                            const double *flat_wb);
void destroy_network_otf_f64(void *net);

/* ---- device handles and operators: replaces libcuotf.so (src/extern/cuotf.cu:7-162) ------- */
void sync_device_cuda(void);                              /* cuotf.cu:9-12 */
void *build_basisdev_f64(void *host_basis);               /* cuotf.cu:25-30 */
void destroy_basisdev_f64(void *basis);                   /* cuotf.cu:14-23 */
/* build_networkdev_f64 (cuotf.cu:43-48): decodes a host network from build_network_otf_f64.
 * A network with a diagonal group is read as a two-body Hamiltonian (integrals recovered by
 * probing matrix elements), the fast path; otherwise every group must be +-(T - T^dagger) of
 * one excitation (a generator pool).  The decode is re-checked against the network's elements
 * on probe determinants.  A network with a diagonal group that it does not represent exactly
 * (spin models, S^2 penalties, non-two-body terms) is kept as its Pauli groups and evaluated
 * group by group (hvec.cuh:6-446's algorithm; hvec_cuda and get_diags_elements_cuda only); a
 * pool network that fails returns NULL. */
void *build_networkdev_f64(void *host_net);
void destroy_networkdev_f64(void *net);                   /* cuotf.cu:32-41 */
/* which form a device operator handle took (no reference counterpart; for callers that want to
 * know whether the fast path applies): 1 two-body Hamiltonian in integral form, 2 generator
 * pool, 3 generic Pauli-group network; 0 on error */
int hipotf_network_kind(void *net);
/* hvec_cuda (cuotf.cu:50-60): dst = H src (dst overwritten, hvec.cuh:439) */
void hvec_cuda(void *basis, void *net, const double *src, double *dst);
/* get_diags_elements_cuda (cuotf.cu:62-71): diags[i] += H_ii (accumulates) */
void get_diags_elements_cuda(void *basis, void *net, double *diags);
/* expm_cuda (cuotf.cu:73-83): vec <- exp(theta G_idx) vec, idx 0-based pool index */
void expm_cuda(void *basis, void *net, int64_t idx, double theta, double *vec);
/* grad_cuda (cuotf.cu:85-96): sum over pairs of grad_update (core/math.hpp:148-162) */
double grad_cuda(void *basis, void *net, int64_t idx, double theta, const double *lp,
                 const double *rp);
/* backgrad_cuda (cuotf.cu:98-109): lp <- R(-theta) lp; g; rp <- R(-theta) rp (math.hpp:187-215) */
double backgrad_cuda(void *basis, void *net, int64_t idx, double theta, double *lp, double *rp);
/* batchgrad_cuda (cuotf.cu:111-122): grads[g] = grad(g, thetas[g], lp, rp); device arrays */
void batchgrad_cuda(void *basis, void *net, const double *thetas, const double *lp,
                    const double *rp, double *grads);
/* The pool operators the reference has on the CPU only (libotf_native,
 * src/include/op/declare.hpp:76-130; jl/network.jl:572-596 binds them as tvec, backtran,
 * batchexpm, batchtran), here on device vectors with a device basis / pool handle.
 * tvec_contract_otf (op/tvec.hpp:373-445): dst = G_idx src; dst overwritten (zero outside the
 * generator's pairs); src != dst */
void tvec_cuda(void *basis, void *net, int64_t idx, const double *src, double *dst);
/* backtran_contract_otf (op/backtran.hpp:383-452): lp, rp <- exp(theta G_idx) lp, rp, then
 * bp = G_idx lp (bp overwritten); three different vectors */
void backtran_cuda(void *basis, void *net, int64_t idx, double theta, double *lp, double *rp,
                   double *bp);
/* batch_expm_contract_otf (op/batchexpm.hpp:376-444): exp(theta G_idx) applied to columns
 * 0..nvec-1 of the det-major matrix m[i * ld + k] (dim rows of ld entries; nvec <= ld).  The
 * reference takes int ld / num_vecs; the widths here are 64-bit. */
void batchexpm_cuda(void *basis, void *net, int64_t idx, double theta, double *matrix,
                    int64_t ld, int64_t nvec);
/* batch_tran_contract_otf (op/batchtran.hpp:517-528): trans[g] = sum over the pairs of G_g of
 * tran_update (math.hpp:181-185) = -<lp| G_g |rp>, for every generator; device arrays */
void batchtran_cuda(void *basis, void *net, const double *lp, const double *rp, double *trans);
void expm_cuda_2d(void *basis, void *net, int64_t idx, double theta, double *vec); /* cuotf.cu:124-133 */
double grad_cuda_2d(void *basis, void *net, int64_t idx, double theta, const double *lp,
                    const double *rp);                                            /* cuotf.cu:135-146 */
double backgrad_cuda_2d(void *basis, void *net, int64_t idx, double theta, double *lp,
                        double *rp);                                              /* cuotf.cu:148-159 */

/* ---- constructors added beside the network path (no reference counterpart: the reference
 *      only accepts SVD networks, jl/cuda/cunetwork.jl:38-53) ------------------------------- */
/* Hamiltonian from the same integral buffers generate_hamiltonian_real_64_128_f64 takes
 * (src/include/ham/declare.hpp:7-27): one_body_mo[p + q n] = h_pq,
 * two_body_mo[p + q n + r n^2 + s n^3] = (ps|qr) (column-major), energy_nuc = E_core. */
void *hipotf_build_ham_f64(int64_t norb, const double *one_body_mo, const double *two_body_mo,
                           double energy_nuc);
/* Generator pool from spin-orbital lists in the format of generate_ci_spin_orbitals
 * (jl/ansatz.jl:67-167): op g uses orbs[offsets[g] .. offsets[g+1]), creators first, then
 * annihilators (spin-orbital 2p = alpha_p, 2p+1 = beta_p); G_g = T_g - T_g^dagger. */
void *hipotf_build_pool_f64(int64_t norb, int64_t nops, const int32_t *offsets,
                            const int32_t *orbs);
int64_t hipotf_num_generators(void *net);
/* builds the basis-dependent device tables of an operator now instead of on first use (the
 * integral-form Hamiltonian and pools; a generic network builds its tables in hvec_cuda) */
int hipotf_prepare(void *basis, void *net);
/* launch choices of the sigma kernels, process-wide (no reference counterpart; the
 * reference's tiles are compile-time, src/include/ham/otf.hpp:6-9 and cuda/common.cuh:7-12):
 * v[8] = {alpha-beta kernel 0 auto (matrix-core at 14-16 orbitals, else class-split FP32 pairs,
 * source-lane, register, chunked) | 1 source-lane | 2 register | 3 chunked | 4 source-lane with
 * FP32 vector pairs on the class-split kernel | 5 matrix-core (16 orbitals or fewer), threads per
 * source string 1|2, integral-row buffers 1|2, FP32 vectors per alpha-beta launch 1|2,
 * row-SpMM source-range passes 1|2|4|8 (0: as many as keep a column tile's gathered slice
 * within an XCD's L2), deterministic 0|1, same-spin destination rows per
 * gathered source row 4|8|16 (0: one row per wave), VQE rotation sweeps 0 fused (as many rows
 * in LDS as fit, the others in global memory) | 1 fused with rows in global memory | 2 one
 * launch per generator | 3 fused with at most one row and half of the next in LDS, +4 (4..7)
 * the same with the adjoint sweep's alpha-pair segments on interleaved (l, r) rows (default 4;
 * FP64 table only)},
 * for the FP64
 * (precision 0) or the
 * FP32 (precision 1) kernels.  binsim_amd/autotune.py measures the candidates and persists its
 * choice as JSON.  Return 0 on success. */
int hipotf_set_tuning(int precision, const int32_t *v, int n);
int hipotf_get_tuning(int precision, int32_t *v, int n);
/* Timing of the alpha-beta part of the sigma builds (no counterpart in the reference; for
 * benchmarks and roofline figures).  enable = 1 zeroes the totals and starts, 0 stops, -1 only
 * reads.  While on, every sigma build records HIP events on its stream around its alpha-beta
 * launches and waits for them (one synchronization per build; not for use under graph
 * capture).  Writes through the non-null pointers: milliseconds summed over the timed sections,
 * the number of sections, and how many of the launches in them were the matrix-core kernel
 * (ab_kernel 5).  Process-wide, not thread-safe.  Return 0 on success. */
int hipotf_ab_timing(int enable, double *ms, int64_t *sections, int64_t *mma_launches);
/* hvec_cuda without device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  dst = H src in FP64 as hvec_cuda computes it, with the four sigma
 * work buffers in the library's static device scratch array (64 MiB per GPU, held until the
 * process ends), so at most 2,097,152 determinants (larger bases are refused).  The operator's
 * device tables must already exist for this basis (hipotf_prepare or one hvec_cuda call);
 * otherwise the call is refused rather than building them.  Calls on one GPU share the
 * scratch array: calls that overlap on one GPU (several host threads or streams) race.
 * Returns 0 on success, 1 on error (message in hipotf_last_error()). */
int hipotf_hvec_noalloc(void *basis, void *net, const double *src, double *dst);
/* block sigma: dst[v] = H src[v], v < nvec (host arrays of device pointers).
 * precision 0: float64 arithmetic (as hvec_cuda); 1: float32 coefficients and integrals with
 * float64 input/output vectors (the mixed-precision Davidson bulk). */
int hipotf_hvec_block(void *basis, void *net, int nvec, const double *const *src,
                      double *const *dst, int precision);
/* fused VQE objective of jl/vqe.jl:1-21 (energy_objective): forward expm sweep over
 * idxs/thetas (host arrays), lam = H psi, energy, variance, reverse backgrad sweep.
 * psi: device, in = reference state, out = reference state after the uncomputing sweep;
 * lam: device workspace; grads: host [nparams] = 2 Re(backgrad)/|psi|^2;
 * out3: host {energy, max(0, |H psi|^2/|psi|^2 - E^2), |psi|^2}.  Returns 0 on success. */
int hipotf_vqe_energy_grad(void *basis, void *ham, void *pool, int64_t nparams, const int64_t *idxs,
                           const double *thetas, double *psi, double *lam, double *grads,
                           double *out3);
/* energy only (the evaluations of jl/vqe.jl:1-21 without the gradient: line searches,
 * finite-difference checks): forward sweep, lam = H psi, out3 as above; psi: in = reference
 * state, out = U(theta)|ref> (not uncomputed).  With ham == NULL only the forward sweep runs
 * (lam is workspace, out3 untouched and may be NULL). */
int hipotf_vqe_energy(void *basis, void *ham, void *pool, int64_t nparams, const int64_t *idxs,
                      const double *thetas, double *psi, double *lam, double *out3);

/* Davidson solve on the device (outer loop of jl/davidson.jl:1-185 extended to nroots roots;
 * no reference counterpart runs on the GPU).  mode 0: FP64, one FP64 sigma per correction
 * vector (the drop-in arithmetic, restart as jl/davidson.jl); mode 1: mixed-precision block
 * -- the `block` (<= 0: max(4, nroots)) lowest Ritz pairs are corrected each iteration in one
 * FP32 block sigma, FP64 subspace / Rayleigh-Ritz / residuals, thick restart to half the
 * subspace, then FP64 block refinement once the wanted residuals < switch_tol (or the FP32
 * residuals stagnate) until tol.  guess: device [nroots*dim] or NULL (unit vectors on the
 * lowest diagonals); energies: host [nroots] (the latest estimates also when not
 * converged); vecs: device [nroots*dim] or NULL; info: host [7] = {iterations, fp64 sigma
 * vectors, fp32 sigma vectors, max residual, seconds, switch iteration, converged}.
 * maxspace <= 0 selects 23 for one root (ncv + 20) and 8*nroots in mode 0, 16*block in
 * mode 1 (at most 64).  The subspace, residuals and sigma work space are checked against the
 * free device memory before the loop (error if they do not fit); the loop allocates nothing.
 * Returns 0 when converged, 2 when maxiter was reached, 3 on breakdown (no new direction while
 * a residual is still above tol; info[6] = 2), 1 on error; with 2 and 3, energies and vecs
 * hold the latest Ritz pairs. */
int hipotf_davidson_f64(void *basis, void *ham, int nroots, int mode, int block, double tol,
                        double switch_tol, int maxiter, int maxspace, int verbose, const double *guess,
                        double *energies, double *vecs, double *info);

/* ---- Davidson inner loop on the device: replaces libdiag.so (src/include/diag/declare.hpp).
 *      V / AV are HOST arrays of DEVICE vector pointers (<= 64); coeffs, out_col and heff are
 *      host arrays. ---------------------------------------------------------------------- */
double build_ritz_and_residual_f64(int64_t N, int dim, const double *coeffs, double min_e,
                                   const double **V, const double **AV, double *axt); /* :7-15 */
void build_ritz_vector_f64(int64_t N, int dim, const double *coeffs, const double **V,
                           double *xt);                                                  /* :17-24 */
void apply_preconditioner_inplace_f64(int64_t N, double *axt, const double *diags, double min_e,
                                      double shift);                                     /* :26-33 */
double mgs_orthogonalize_f64(int64_t N, int dim, const double **V, double *vt);          /* :35-41 */
double inplace_normalize_f64(int64_t N, double *vec);                                    /* :43-48 */
double fast_dot_f64(int64_t N, const double *v1, const double *v2);                      /* :50-56 */
void compute_heff_col_f64(int64_t N, int dim, const double **V, const double *AV_dim,
                          double *out_col);                                              /* :58-65 */
void rebuild_heff_f64(int64_t N, int dim, const double **V, const double **AV, double *heff,
                      int ldh);                                                          /* :67-75 */

#ifdef __cplusplus
}
#endif

#endif /* HIPOTF_H */
