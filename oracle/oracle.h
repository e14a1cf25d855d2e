/*
 * ORACLE -- test infrastructure only.  Clean-room C restatement of the reference CPU algorithms
 * (libotf_native / libdiag).  Only tests/, the golden-fixture script, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it, and only as the checker; the product never does.
 *
 * The operator is given as the flat SVD-group network of build_network_otf
 * (src/include/ham/otf.hpp:223-349, arrays as jl/network.jl:313-339 passes them).  Every
 * function follows the reference's arithmetic (src/include/core/math.hpp) with a plain loop
 * order; summation order differs from the reference, so float64 results agree to rounding.
 */
#ifndef ORACLE_H
#define ORACLE_H
#include <stdint.h>

typedef struct {
    int64_t ngs;
    const uint32_t *axs, *bxs;
    const int64_t *ranks, *num_zas, *num_zbs;
    const uint32_t *flat_zas, *flat_zbs;
    const double *flat_wa, *flat_wb;
} oracle_net;

/* colex (ascending numeric) strings with k of n bits: basis.hpp:141-155 */
int64_t oracle_strings(int n, int k, uint32_t *out);
/* hvec: dst = sum_g f_g(src) x[src]  (op/hvec.hpp:481-496) */
void oracle_hvec(int n, int na, int nb, const oracle_net *net, const double *x, double *y);
/* diag[i] += f_diag(a_i, b_i)  (op/hvec.hpp:498-554) */
void oracle_diags(int n, int na, int nb, const oracle_net *net, double *diag);
/* one group g (original index) of a pool network: op/expm.hpp:374-438, grad.hpp:397-465,
 * backgrad.hpp:406-474 */
void oracle_expm(int n, int na, int nb, const oracle_net *net, int64_t g, double theta, double *v);
double oracle_grad(int n, int na, int nb, const oracle_net *net, int64_t g, double theta,
                   const double *l, const double *r);
double oracle_backgrad(int n, int na, int nb, const oracle_net *net, int64_t g, double theta,
                       double *l, double *r);
/* dst = G_g src (dst zeroed first): op/tvec.hpp:373-445 */
void oracle_tvec(int n, int na, int nb, const oracle_net *net, int64_t g, const double *src, double *dst);
/* l, r <- exp(theta G_g) l, r; b = G_g l (b zeroed first): op/backtran.hpp:383-452 */
void oracle_backtran(int n, int na, int nb, const oracle_net *net, int64_t g, double theta, double *l,
                     double *r, double *b);
/* sum over the pairs of G_g of tran_update (math.hpp:181-185): one entry of batch_tran
 * (op/batchtran.hpp:517-528) */
double oracle_tran(int n, int na, int nb, const oracle_net *net, int64_t g, const double *l, const double *r);
/* exp(theta G_g) on vectors 0..nvec-1 of m[i * ld + k]: op/batchexpm.hpp:376-444 */
void oracle_batch_expm(int n, int na, int nb, const oracle_net *net, int64_t g, double theta, double *m,
                       int64_t ld, int nvec);
/* libdiag (diag/davidson.hpp:31-272) */
double oracle_ritz_and_residual(int64_t N, int dim, const double *c, double e, const double *const *V,
                                const double *const *AV, double *axt);
void oracle_ritz_vector(int64_t N, int dim, const double *c, const double *const *V, double *xt);
void oracle_precondition(int64_t N, double *axt, const double *diags, double e, double shift);
double oracle_mgs(int64_t N, int dim, const double *const *V, double *vt);
double oracle_normalize(int64_t N, double *v);
double oracle_dot(int64_t N, const double *a, const double *b);
void oracle_heff_col(int64_t N, int dim, const double *const *V, const double *av, double *out);
#endif
