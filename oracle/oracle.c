/*
 * ORACLE -- test infrastructure only (see oracle.h).  Restates, in plain loops:
 *   basis enumeration     src/include/basis/basis.hpp:141-155, 271-285, 327
 *   network phases        src/include/core/math.hpp:37-124 (precompute_phase, compute_coeff)
 *   hvec                  src/include/op/hvec.hpp:481-496 (gather; dst zeroed first)
 *   get_diags_elements    src/include/op/hvec.hpp:498-554 (accumulates)
 *   expm / grad / backgrad src/include/op/expm.hpp, grad.hpp, backgrad.hpp with the update
 *                         formulas of core/math.hpp:132-215; each pair (src,dst), dst = src^x,
 *                         is visited once, from the source with the larger determinant index
 *                         (expm.hpp:104-131, 317-318)
 *   tvec / backtran       op/tvec.hpp:373-445, op/backtran.hpp:383-452 (tvec_update,
 *                         backtran_update: math.hpp:164-225; the output vector zeroed first)
 *   batch expm / tran     op/batchexpm.hpp:376-444 (expm_batch_update_matrix: the pair's rows
 *                         of a det-major matrix with leading dimension ld), op/batchtran.hpp:
 *                         517-528 (tran_update summed per generator)
 *   libdiag               src/include/diag/davidson.hpp:31-272
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

static inline int par32(uint32_t v) { return __builtin_popcount(v) & 1; }

int64_t oracle_strings(int n, int k, uint32_t *out) {
    int64_t cnt = 0;
    uint32_t s = (k == 0) ? 0u : ((1u << k) - 1u);
    const uint64_t lim = 1ull << n;
    for (;;) {
        if (out) out[cnt] = s;
        cnt++;
        if (k == 0) break;
        uint32_t c = s & (~s + 1u);
        uint32_t r = s + c;
        uint32_t t = (((r ^ s) >> 2) / c) | r;
        if ((uint64_t)t >= lim || t <= s) break;
        s = t;
    }
    return cnt;
}

typedef struct {
    int n, na, nb;
    int64_t NA, NB;
    uint32_t *A, *B;
    int32_t *amap, *bmap;  /* 2^n tables, -1 outside the basis (basis.hpp:271-285) */
} basis_t;

static void basis_init(basis_t *b, int n, int na, int nb) {
    b->n = n;
    b->na = na;
    b->nb = nb;
    b->NA = oracle_strings(n, na, NULL);
    b->NB = oracle_strings(n, nb, NULL);
    b->A = (uint32_t *)malloc(sizeof(uint32_t) * b->NA);
    b->B = (uint32_t *)malloc(sizeof(uint32_t) * b->NB);
    oracle_strings(n, na, b->A);
    oracle_strings(n, nb, b->B);
    const size_t m = (size_t)1 << n;
    b->amap = (int32_t *)malloc(sizeof(int32_t) * m);
    b->bmap = (int32_t *)malloc(sizeof(int32_t) * m);
    for (size_t i = 0; i < m; ++i) b->amap[i] = b->bmap[i] = -1;
    for (int64_t i = 0; i < b->NA; ++i) b->amap[b->A[i]] = (int32_t)i;
    for (int64_t i = 0; i < b->NB; ++i) b->bmap[b->B[i]] = (int32_t)i;
}

static void basis_free(basis_t *b) {
    free(b->A);
    free(b->B);
    free(b->amap);
    free(b->bmap);
}

typedef struct {
    uint32_t ax, bx;
    int rank, nza, nzb;
    const uint32_t *za, *zb;
    const double *wa, *wb; /* column r contiguous: w[r * nz + k] */
} group_t;

static group_t *groups_of(const oracle_net *net) {
    group_t *g = (group_t *)malloc(sizeof(group_t) * (net->ngs ? net->ngs : 1));
    size_t oza = 0, ozb = 0, owa = 0, owb = 0;
    for (int64_t i = 0; i < net->ngs; ++i) {
        g[i].ax = net->axs[i];
        g[i].bx = net->bxs[i];
        g[i].rank = (int)net->ranks[i];
        g[i].nza = (int)net->num_zas[i];
        g[i].nzb = (int)net->num_zbs[i];
        g[i].za = net->flat_zas + oza;
        g[i].zb = net->flat_zbs + ozb;
        g[i].wa = net->flat_wa + owa;
        g[i].wb = net->flat_wb + owb;
        oza += g[i].nza;
        ozb += g[i].nzb;
        owa += (size_t)g[i].nza * g[i].rank;
        owb += (size_t)g[i].nzb * g[i].rank;
    }
    return g;
}

/* P_r(s) = sum_k w[r,k] (-1)^popc(s & z_k)   (math.hpp:37-78), for every string of a list */
static double *phases(const group_t *g, int side, const uint32_t *strs, int64_t ns) {
    const int rank = g->rank, nz = side ? g->nzb : g->nza;
    const uint32_t *z = side ? g->zb : g->za;
    const double *w = side ? g->wb : g->wa;
    double *P = (double *)malloc(sizeof(double) * (size_t)(rank ? rank : 1) * (ns ? ns : 1));
    for (int64_t i = 0; i < ns; ++i)
        for (int r = 0; r < rank; ++r) {
            double v = 0.0;
            for (int k = 0; k < nz; ++k) v += par32(strs[i] & z[k]) ? -w[r * nz + k] : w[r * nz + k];
            P[(size_t)r * ns + i] = v;
        }
    return P;
}

static inline double coeff(const double *PA, int64_t NA, int64_t ia, const double *PB, int64_t NB,
                           int64_t ib, int rank) {
    double v = 0.0;
    for (int r = 0; r < rank; ++r) v += PA[(size_t)r * NA + ia] * PB[(size_t)r * NB + ib];
    return v;
}

void oracle_hvec(int n, int na, int nb, const oracle_net *net, const double *x, double *y) {
    basis_t b;
    basis_init(&b, n, na, nb);
    group_t *G = groups_of(net);
    const int64_t NA = b.NA, NB = b.NB;
    memset(y, 0, sizeof(double) * (size_t)NA * NB);
    for (int64_t g = 0; g < net->ngs; ++g) {
        /* phases are evaluated at the SOURCE strings: src = dst ^ x (hvec.hpp:133-145) */
        double *PA = phases(&G[g], 0, b.A, NA);
        double *PB = phases(&G[g], 1, b.B, NB);
        const uint32_t ax = G[g].ax, bx = G[g].bx;
        const int rank = G[g].rank;
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t a = 0; a < NA; ++a) {
            const int32_t sa = b.amap[b.A[a] ^ ax];
            if (sa < 0) continue;
            for (int64_t ib = 0; ib < NB; ++ib) {
                const int32_t sb = b.bmap[b.B[ib] ^ bx];
                if (sb < 0) continue;
                y[a * NB + ib] += coeff(PA, NA, sa, PB, NB, sb, rank) * x[(int64_t)sa * NB + sb];
            }
        }
        free(PA);
        free(PB);
    }
    free(G);
    basis_free(&b);
}

void oracle_diags(int n, int na, int nb, const oracle_net *net, double *diag) {
    basis_t b;
    basis_init(&b, n, na, nb);
    group_t *G = groups_of(net);
    for (int64_t g = 0; g < net->ngs; ++g) {
        if (G[g].ax || G[g].bx) continue;
        double *PA = phases(&G[g], 0, b.A, b.NA);
        double *PB = phases(&G[g], 1, b.B, b.NB);
#pragma omp parallel for
        for (int64_t a = 0; a < b.NA; ++a)
            for (int64_t ib = 0; ib < b.NB; ++ib)
                diag[a * b.NB + ib] += coeff(PA, b.NA, a, PB, b.NB, ib, G[g].rank);
        free(PA);
        free(PB);
        break; /* only diag_groups[0] (hvec.hpp:501) */
    }
    free(G);
    basis_free(&b);
}

/* mode 0 expm, 1 grad, 2 backgrad, 3 tvec (v -> l), 4 backtran (l, r; b = v), 5 tran (cl, cr),
 * 6 batch expm (v = matrix with leading dimension ld, nvec vectors) */
static double rot(int n, int na, int nb, const oracle_net *net, int64_t gi, double theta, double *v,
                  double *l, double *r, const double *cl, const double *cr, int mode, int64_t ld,
                  int nvec) {
    basis_t b;
    basis_init(&b, n, na, nb);
    group_t *G = groups_of(net);
    const group_t *g = &G[gi];
    double res = 0.0;
    if (g->ax == 0 && g->bx == 0) { /* real diag group: identity / zero gradient (math.hpp:8-35) */
        free(G);
        basis_free(&b);
        return 0.0;
    }
    double *PA = phases(g, 0, b.A, b.NA);
    double *PB = phases(g, 1, b.B, b.NB);
    const int64_t NA = b.NA, NB = b.NB;
    const double cth = cos(theta), sth = sin(theta);
#pragma omp parallel for reduction(+ : res) schedule(dynamic, 4)
    for (int64_t a = 0; a < NA; ++a) {
        const int32_t sa = b.amap[b.A[a] ^ g->ax];
        if (sa < 0 || sa < a) continue;
        for (int64_t ib = 0; ib < NB; ++ib) {
            const int32_t sb = b.bmap[b.B[ib] ^ g->bx];
            if (sb < 0) continue;
            const int64_t di = a * NB + ib, si = (int64_t)sa * NB + sb;
            if (si <= di) continue;
            const double vt = coeff(PA, NA, sa, PB, NB, sb, g->rank);
            if (mode == 0) { /* expm_update, cd = cos-1, co = sin */
                const double vd = 1.0 + (cth - 1.0) * vt * vt, fw = sth * vt;
                const double vi = v[si], vj = v[di];
                v[si] = vi * vd - vj * fw;
                v[di] = vj * vd + vi * fw;
            } else if (mode == 1) { /* grad_update, cd = -sin, co = cos */
                const double vd = -sth * vt * vt, fw = cth * vt;
                res += cl[si] * (cr[si] * vd + cr[di] * fw) + cl[di] * (cr[di] * vd - cr[si] * fw);
            } else if (mode == 3) { /* tvec_update: dst_s = -vt src_d, dst_d = vt src_s */
                l[si] = -vt * v[di];
                l[di] = vt * v[si];
            } else if (mode == 4) { /* backtran_update: expm on l and r, then b = tvec(l) */
                const double vd = 1.0 + (cth - 1.0) * vt * vt, fw = sth * vt;
                const double lvi = l[si], lvj = l[di], rvi = r[si], rvj = r[di];
                l[si] = lvi * vd - lvj * fw;
                l[di] = lvj * vd + lvi * fw;
                r[si] = rvi * vd - rvj * fw;
                r[di] = rvj * vd + rvi * fw;
                v[si] = -vt * l[di];
                v[di] = vt * l[si];
            } else if (mode == 5) { /* tran_update: (l_s vt) r_d - l_d vt r_s */
                res += cl[si] * vt * cr[di] - cl[di] * vt * cr[si];
            } else if (mode == 6) { /* expm_batch_update_matrix */
                const double vd = 1.0 + (cth - 1.0) * vt * vt, fw = sth * vt;
                for (int k = 0; k < nvec; ++k) {
                    const double vi = v[si * ld + k], vj = v[di * ld + k];
                    v[si * ld + k] = vi * vd - vj * fw;
                    v[di * ld + k] = vj * vd + vi * fw;
                }
            } else { /* backgrad_update: ecd = cos-1, eco = -sin, gcd = -sin, gco = cos */
                const double evd = 1.0 + (cth - 1.0) * vt * vt, efw = -sth * vt;
                const double gvd = -sth * vt * vt, gfw = cth * vt;
                const double lvi = l[si], lvj = l[di], rvi = r[si], rvj = r[di];
                const double els = lvi * evd - lvj * efw, eld = lvj * evd + lvi * efw;
                res += els * (rvi * gvd + rvj * gfw) + eld * (rvj * gvd - rvi * gfw);
                l[si] = els;
                l[di] = eld;
                r[si] = rvi * evd - rvj * efw;
                r[di] = rvj * evd + rvi * efw;
            }
        }
    }
    free(PA);
    free(PB);
    free(G);
    basis_free(&b);
    return res;
}

void oracle_expm(int n, int na, int nb, const oracle_net *net, int64_t g, double theta, double *v) {
    rot(n, na, nb, net, g, theta, v, NULL, NULL, NULL, NULL, 0, 0, 0);
}
double oracle_grad(int n, int na, int nb, const oracle_net *net, int64_t g, double theta,
                   const double *l, const double *r) {
    return rot(n, na, nb, net, g, theta, NULL, NULL, NULL, l, r, 1, 0, 0);
}
double oracle_backgrad(int n, int na, int nb, const oracle_net *net, int64_t g, double theta,
                       double *l, double *r) {
    return rot(n, na, nb, net, g, theta, NULL, l, r, NULL, NULL, 2, 0, 0);
}
static int64_t dim_of(int n, int na, int nb) {
    basis_t b;
    basis_init(&b, n, na, nb);
    const int64_t d = b.NA * b.NB;
    basis_free(&b);
    return d;
}
void oracle_tvec(int n, int na, int nb, const oracle_net *net, int64_t g, const double *src, double *dst) {
    memset(dst, 0, sizeof(double) * (size_t)dim_of(n, na, nb));
    rot(n, na, nb, net, g, 0.0, (double *)src, dst, NULL, NULL, NULL, 3, 0, 0);
}
void oracle_backtran(int n, int na, int nb, const oracle_net *net, int64_t g, double theta, double *l,
                     double *r, double *b) {
    memset(b, 0, sizeof(double) * (size_t)dim_of(n, na, nb));
    rot(n, na, nb, net, g, theta, b, l, r, NULL, NULL, 4, 0, 0);
}
double oracle_tran(int n, int na, int nb, const oracle_net *net, int64_t g, const double *l, const double *r) {
    return rot(n, na, nb, net, g, 0.0, NULL, NULL, NULL, l, r, 5, 0, 0);
}
void oracle_batch_expm(int n, int na, int nb, const oracle_net *net, int64_t g, double theta, double *m,
                       int64_t ld, int nvec) {
    rot(n, na, nb, net, g, theta, m, NULL, NULL, NULL, NULL, 6, ld, nvec);
}

/* ---- libdiag --------------------------------------------------------------------------- */
double oracle_ritz_and_residual(int64_t N, int dim, const double *c, double e, const double *const *V,
                                const double *const *AV, double *axt) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s)
    for (int64_t i = 0; i < N; ++i) {
        double x = 0.0, ax = 0.0;
        for (int j = 0; j < dim; ++j) {
            x += c[j] * V[j][i];
            ax += c[j] * AV[j][i];
        }
        const double rr = ax - e * x;
        axt[i] = rr;
        s += rr * rr;
    }
    return sqrt(s);
}

void oracle_ritz_vector(int64_t N, int dim, const double *c, const double *const *V, double *xt) {
#pragma omp parallel for
    for (int64_t i = 0; i < N; ++i) {
        double x = 0.0;
        for (int j = 0; j < dim; ++j) x += c[j] * V[j][i];
        xt[i] = x;
    }
}

void oracle_precondition(int64_t N, double *axt, const double *diags, double e, double shift) {
    const double as = fabs(shift);
#pragma omp parallel for
    for (int64_t i = 0; i < N; ++i) {
        double d = diags[i] - e;
        if (fabs(d) < as) d = d >= 0 ? as : -as;
        axt[i] /= d;
    }
}

double oracle_dot(int64_t N, const double *a, const double *b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s)
    for (int64_t i = 0; i < N; ++i) s += a[i] * b[i];
    return s;
}

double oracle_mgs(int64_t N, int dim, const double *const *V, double *vt) {
    for (int j = 0; j < dim; ++j) {
        const double p = oracle_dot(N, V[j], vt);
#pragma omp parallel for
        for (int64_t i = 0; i < N; ++i) vt[i] -= p * V[j][i];
    }
    return sqrt(oracle_dot(N, vt, vt));
}

double oracle_normalize(int64_t N, double *v) {
    const double nrm = sqrt(oracle_dot(N, v, v)), inv = 1.0 / nrm;
#pragma omp parallel for
    for (int64_t i = 0; i < N; ++i) v[i] *= inv;
    return nrm;
}

void oracle_heff_col(int64_t N, int dim, const double *const *V, const double *av, double *out) {
    for (int j = 0; j < dim; ++j) out[j] = oracle_dot(N, V[j], av);
}
