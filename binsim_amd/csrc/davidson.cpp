// Device-resident Davidson driver.
//
// mode 0 ("plain FP64"): the outer loop of jl/davidson.jl:1-185 -- Ritz pair from the small
//   eigenproblem, residual, restart to the lowest Ritz vectors when the subspace is full,
//   diagonal preconditioner with shift max(|r| 1e-2, 1e-4), modified Gram-Schmidt repeated
//   when the norm drops below 0.717, normalisation -- extended to nroots roots (one correction
//   vector per unconverged root per iteration, Davidson-Liu).  Every sigma is one FP64 vector.
// mode 1 ("mixed block"): block Davidson -- the k = max(block, nroots) lowest Ritz pairs are
//   corrected every iteration (the extra pairs accelerate the wanted ones) and all correction
//   vectors of an iteration go through one block sigma launch in FP32 (coefficients,
//   integrals and work vectors); the subspace, Rayleigh-Ritz, residuals and energies stay
//   FP64; restarts keep half the subspace (thick restart).  When the wanted residuals are
//   below switch_tol (or the FP32 residuals stop improving) the k Ritz vectors seed an FP64
//   phase (FP64 block sigma) that runs to tol.
// The small symmetric eigenproblem (m <= 64) is solved on the host by cyclic Jacobi.
#include "hipotf.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <numeric>

namespace hipotf {
double ritz_and_residual(i64 N, int dim, const double *coeffs, double min_e, const double *const *V,
                         const double *const *AV, double *axt, hipStream_t st);
void ritz_vector(i64 N, int dim, const double *coeffs, const double *const *V, double *xt, hipStream_t st);
void precondition(i64 N, double *axt, const double *diags, double min_e, double shift, hipStream_t st);
double mgs(i64 N, int dim, const double *const *V, double *vt, hipStream_t st);
double normalize(i64 N, double *vec, hipStream_t st);
double dot(i64 N, const double *a, const double *b, hipStream_t st);
void multidot(i64 N, int dim, const double *const *V, const double *a, double *out_host, hipStream_t st);
void block_combine(i64 N, int m, int k, const double *C, const double *w, const double *const *X,
                   const double *const *Y, double *const *out, double *norms, hipStream_t st);
void block_dots(i64 N, int na, const double *const *A, int nbv, const double *const *B, double *G,
                hipStream_t st);
void block_subtract(i64 N, int m, const double *const *V, const double *P, int nt, double *const *T,
                    hipStream_t st);
void rotate_inplace(i64 N, int m, int keep, const double *C, double *const *X, hipStream_t st);
void scatter_values(double *v, const i64 *idx, const double *val, int count, hipStream_t st);

// determinants of the P space of the initial guess (lowest diagonals)
constexpr int PSPACE = 256;

// eigen-decomposition of a symmetric m x m matrix (row-major a), ascending eigenvalues,
// eigenvectors in the columns of v (row-major)
void jacobi_eigh(int m, std::vector<double> a, std::vector<double> &w, std::vector<double> &v) {
    v.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) v[i * m + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                tot += a[i * m + j] * a[i * m + j];
                if (i != j) off += a[i * m + j] * a[i * m + j];
            }
        // converged when the off-diagonal part is at rounding level (|off| <= 1e-15 |A|), or
        // when a sweep finds nothing above it to rotate away
        if (off <= 1e-30 * tot || off == 0.0) break;
        const double small = 1e-17 * std::sqrt(tot);
        int rotated = 0;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = a[p * m + q];
                if (std::fabs(apq) <= small) continue;
                ++rotated;
                const double app = a[p * m + p], aqq = a[q * m + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < m; ++k) {
                    const double akp = a[k * m + p], akq = a[k * m + q];
                    a[k * m + p] = c * akp - s * akq;
                    a[k * m + q] = s * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {
                    const double apk = a[p * m + k], aqk = a[q * m + k];
                    a[p * m + k] = c * apk - s * aqk;
                    a[q * m + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; ++k) {
                    const double vkp = v[k * m + p], vkq = v[k * m + q];
                    v[k * m + p] = c * vkp - s * vkq;
                    v[k * m + q] = s * vkp + c * vkq;
                }
            }
        if (rotated == 0) break;
    }
    std::vector<int> order(m);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int i, int j) { return a[i * m + i] < a[j * m + j]; });
    w.resize(m);
    std::vector<double> vs((size_t)m * m);
    for (int c = 0; c < m; ++c) {
        w[c] = a[order[c] * m + order[c]];
        for (int r = 0; r < m; ++r) vs[r * m + c] = v[r * m + order[c]];
    }
    v.swap(vs);
}

struct DavidsonOut {
    int iterations = 0;
    i64 sigma64 = 0, sigma32 = 0;
    double max_residual = 0.0;
    double seconds = 0.0;
    int refine_iter = -1;
    int converged = 0;
};

DavidsonOut davidson(DevHam *H, int nroots, int mode, int block, double tol, double switch_tol, int maxiter,
                     int maxspace, int verbose, const double *guess, double *energies, double *vecs) {
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = (hipStream_t)0;
    const i64 N = H->basis->host.dim;
    HIPOTF_REQUIRE(N > 0, "empty basis");
    HIPOTF_REQUIRE(nroots >= 1 && nroots <= 16 && nroots <= N, "nroots out of range");
    // k = Ritz pairs corrected per iteration: nroots in FP64 mode (jl/davidson.jl), at least 4
    // in the mixed block mode (the extra pairs speed up the wanted ones)
    const int k = mode == 0 ? nroots : (int)std::min<i64>(std::max(block > 0 ? block : 4, nroots), 16);
    if (maxspace <= 0) maxspace = mode == 0 ? (nroots == 1 ? 23 : 8 * nroots) : 16 * k;
    maxspace = (int)std::min<i64>(std::min(maxspace, 64), N);
    // vectors kept on restart: ncv = 3 for one root (jl/davidson.jl), 4 per root otherwise;
    // the block mode keeps half the subspace (thick restart)
    const int nkeep = mode == 1 ? std::min(maxspace - k, std::max(2 * k, maxspace / 2))
                      : nroots == 1 ? std::min(3, maxspace - 1)
                                    : std::max(nroots, std::min(4 * nroots, maxspace / 2));
    HIPOTF_REQUIRE(maxspace >= 2 * k || maxspace == N, "maxspace must be >= 2 * block");
    DavidsonOut out;

    // device memory of the whole solve, checked before anything is allocated: 2 * maxspace
    // subspace vectors (V, AV), the diagonal, k residuals (plain mode; the block mode forms its
    // residuals in free subspace slots), and the sigma work buffers (4 per basis, dim *
    // vectors per pass * word size; grown on first use).  The block mode passes as many of its
    // up to 8 vectors through one sigma as the memory left allows (nv_sigma).  The loop itself
    // allocates nothing: restarts and the FP32 -> FP64 switch combine the subspace in place.
    int nv_sigma = 1;
    {
        const size_t vec = (size_t)N * sizeof(double);
        const size_t need_vec = vec * (2 * (size_t)maxspace + (mode == 0 ? k : 0) + 1);
        const size_t have_sig = std::min(std::min(H->basis->xt.bytes, H->basis->xtT.bytes),
                                         std::min(H->basis->s1.bytes, H->basis->s2T.bytes));
        // FP64 phase: 4 buffers of dim * nv doubles (the FP32 phase needs half)
        auto grow = [&](int nv) { return vec * nv > have_sig ? 4 * (vec * nv - have_sig) : (size_t)0; };
        size_t free_b = 0, total_b = 0;
        HIPOTF_CHECK(hipMemGetInfo(&free_b, &total_b));
        if (need_vec + grow(1) > free_b) {
            char msg[320];
            std::snprintf(msg, sizeof msg,
                          "davidson: needs %.1f GB of device memory (%d + %d subspace vectors, %d residuals, "
                          "diagonal, %.1f GB more sigma work space) but %.1f GB are free; lower maxspace or block",
                          (need_vec + grow(1)) / 1e9, maxspace, maxspace, mode == 0 ? k : 0, grow(1) / 1e9,
                          free_b / 1e9);
            throw Error(msg);
        }
        if (mode == 1)
            for (int nv = std::min(k, 8); nv >= 1; --nv)
                if (need_vec + grow(nv) <= free_b) {
                    nv_sigma = nv;
                    break;
                }
    }
    std::vector<DevBuf> Vb(maxspace), AVb(maxspace), R(mode == 0 ? k : 0);
    for (auto &b : Vb) b.alloc(N * sizeof(double));
    for (auto &b : AVb) b.alloc(N * sizeof(double));
    for (auto &b : R) b.alloc(N * sizeof(double));
    DevBuf diag(N * sizeof(double));
    HIPOTF_CHECK(hipMemsetAsync(diag.p, 0, N * sizeof(double), st));
    diag_elements(H, diag.as<double>(), st);
    std::vector<double *> V(maxspace), AV(maxspace);
    for (int i = 0; i < maxspace; ++i) {
        V[i] = Vb[i].as<double>();
        AV[i] = AVb[i].as<double>();
    }

    // initial subspace: the caller's nroots vectors, else the lowest eigenvectors of H in the
    // "P space" of the NP lowest-diagonal determinants (H_PP built on the host by Slater-Condon
    // rules, the FCI codes' usual start), else unit vectors on the lowest diagonals; the block
    // mode tops it up to k vectors
    std::vector<i64> low;
    {
        std::vector<double> dh(N);
        HIPOTF_CHECK(hipMemcpy(dh.data(), diag.p, N * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<i64> idx(N);
        std::iota(idx.begin(), idx.end(), 0);
        const i64 nl = std::min<i64>(N, std::max<i64>((i64)k + nroots, PSPACE));
        std::partial_sort(idx.begin(), idx.begin() + nl, idx.end(),
                          [&](i64 a, i64 b) { return dh[a] < dh[b] || (dh[a] == dh[b] && a < b); });
        low.assign(idx.begin(), idx.begin() + nl);
    }
    std::vector<std::vector<double>> pvec;  // P-space eigenvectors over low[0..np)
    const int np = (int)std::min<i64>(PSPACE, (i64)low.size());
    if (!guess && np >= std::max(k, nroots) + 1 && H->basis->host.norb <= 32) {
        const HostBasis &hb = H->basis->host;
        const i64 NB = hb.B.N;
        std::vector<u64> occ(np);
        for (int i = 0; i < np; ++i)
            occ[i] = interleave_occ((u32)hb.A.strs[low[i] / NB], (u32)hb.B.strs[low[i] % NB]);
        std::vector<double> hp((size_t)np * np), wp, cp;
        for (int i = 0; i < np; ++i)
            for (int j = 0; j <= i; ++j)
                hp[(size_t)i * np + j] = hp[(size_t)j * np + i] =
                    slater_condon(H->n, H->h.data(), H->eri.data(), H->ecore, occ[i], occ[j]);
        jacobi_eigh(np, hp, wp, cp);
        const int nv0 = mode == 1 ? k : nroots;
        for (int i = 0; i < nv0; ++i) {
            pvec.emplace_back(np);
            for (int j = 0; j < np; ++j) pvec.back()[j] = cp[(size_t)j * np + i];
        }
    }
    const double one = 1.0;
    int m = 0;
    auto add_unit = [&](i64 at) {
        HIPOTF_CHECK(hipMemsetAsync(V[m], 0, N * sizeof(double), st));
        HIPOTF_CHECK(hipMemcpyAsync(V[m] + at, &one, sizeof(double), hipMemcpyHostToDevice, st));
        HIPOTF_CHECK(hipStreamSynchronize(st));
    };
    auto accept = [&](bool required) {  // orthonormalise V[m] against V[0..m); keep if independent
        const double nb = std::sqrt(dot(N, V[m], V[m], st));
        double na = mgs(N, m, (const double *const *)V.data(), V[m], st);
        if (na < 0.717 * nb) na = mgs(N, m, (const double *const *)V.data(), V[m], st);
        if (na <= 1e-12 * std::max(1.0, nb)) {
            HIPOTF_REQUIRE(!required, "initial guesses are linearly dependent");
            return;
        }
        normalize(N, V[m], st);
        ++m;
    };
    size_t next_low = 0;
    if (!pvec.empty()) {
        std::vector<i64> pidx(low.begin(), low.begin() + np);
        for (const auto &pv : pvec) {
            HIPOTF_CHECK(hipMemsetAsync(V[m], 0, N * sizeof(double), st));
            scatter_values(V[m], pidx.data(), pv.data(), np, st);
            accept(true);
        }
        next_low = low.size();  // nothing to top up with
    }
    for (int i = 0; i < nroots && pvec.empty(); ++i) {
        if (guess)
            HIPOTF_CHECK(hipMemcpyAsync(V[m], guess + (size_t)i * N, N * sizeof(double),
                                        hipMemcpyDeviceToDevice, st));
        else
            add_unit(low[next_low++]);
        accept(true);
    }
    while (m < (mode == 1 ? k : nroots) && next_low < low.size()) {
        add_unit(low[next_low++]);
        accept(false);
    }
    int phase = mode == 1 ? 1 : 0;  // 1: fp32 block sigma, 0: fp64
    auto apply = [&](int first, int count) {
        if (count == 0) return;
        if (mode == 0) {
            for (int i = first; i < first + count; ++i) {
                const double *s = V[i];
                double *d = AV[i];
                sigma(H, 1, &s, &d, 0, st);
                out.sigma64++;
            }
        } else {
            sigma(H, count, (const double *const *)&V[first], &AV[first], phase, st, nv_sigma);
            if (phase == 1) out.sigma32 += count;
            else out.sigma64 += count;
        }
    };
    apply(0, m);
    std::vector<double> heff((size_t)maxspace * maxspace, 0.0);
    auto set_cols = [&](int first, int count) {
        if (mode == 1) {
            // columns first..first+count of V^T AV, up to 8 at a time, one pass over V each
            for (int c0 = first; c0 < first + count; c0 += 8) {
                const int nc = std::min(8, first + count - c0), nr = c0 + nc;
                std::vector<double> G((size_t)nr * nc);
                block_dots(N, nr, (const double *const *)V.data(), nc, (const double *const *)&AV[c0], G.data(), st);
                for (int j = 0; j < nc; ++j) {
                    const int c = c0 + j;
                    for (int r = 0; r <= c; ++r)
                        heff[(size_t)r * maxspace + c] = heff[(size_t)c * maxspace + r] = G[(size_t)r * nc + j];
                }
            }
            return;
        }
        std::vector<double> col(maxspace);
        for (int c = first; c < first + count; ++c) {
            multidot(N, c + 1, (const double *const *)V.data(), AV[c], col.data(), st);
            for (int r = 0; r <= c; ++r) heff[(size_t)r * maxspace + c] = heff[(size_t)c * maxspace + r] = col[r];
        }
    };
    set_cols(0, m);

    std::vector<double> w, C, best(3, 1e300);
    std::vector<double> rn(k);
    int stall = 0;
    // the latest Ritz vectors when the loop ends unconverged: V[0..nroots) themselves after a
    // restart or the FP32 -> FP64 switch in the last iteration, else the last coefficients over
    // V[0..coef_m) (vectors added after them sit at higher positions)
    std::vector<std::vector<double>> coef_last;
    bool ritz_in_V = false;
    auto write_current = [&]() {
        if (!vecs) return;
        for (int i = 0; i < nroots; ++i) {
            if (ritz_in_V || coef_last.empty())
                HIPOTF_CHECK(hipMemcpyAsync(vecs + (size_t)i * N, V[std::min(i, m - 1)], N * sizeof(double),
                                            hipMemcpyDeviceToDevice, st));
            else
                ritz_vector(N, (int)coef_last[std::min<size_t>(i, coef_last.size() - 1)].size(),
                            coef_last[std::min<size_t>(i, coef_last.size() - 1)].data(),
                            (const double *const *)V.data(), vecs + (size_t)i * N, st);
        }
    };
    auto rayleigh_ritz = [&]() {
        std::vector<double> a((size_t)m * m);
        for (int r = 0; r < m; ++r)
            for (int c = 0; c < m; ++c) a[(size_t)r * m + c] = heff[(size_t)r * maxspace + c];
        jacobi_eigh(m, a, w, C);
    };
    // restart to the nkeep lowest Ritz vectors of the last Rayleigh-Ritz step (jl/davidson.jl:82-114):
    // V_i, AV_i <- the Ritz combinations, in place (no second copy of the kept vectors)
    auto restart = [&]() {
        const int keep = std::min(nkeep, m);
        std::vector<double> Cm((size_t)m * keep);
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < keep; ++i) Cm[(size_t)j * keep + i] = C[(size_t)j * m + i];
        rotate_inplace(N, m, keep, Cm.data(), V.data(), st);
        rotate_inplace(N, m, keep, Cm.data(), AV.data(), st);
        m = keep;
        ritz_in_V = true;
        std::fill(heff.begin(), heff.end(), 0.0);
        set_cols(0, m);
    };
    for (int it = 1; it <= maxiter; ++it) {
        out.iterations = it;
        rayleigh_ritz();
        if (mode == 1 && m + std::min(k, m) > maxspace) {
            // the block mode forms its residuals in the free slots V[m..m+kk): restart first
            restart();
            rayleigh_ritz();
        }
        const int kk = std::min(k, m);  // Ritz pairs with residuals this iteration
        std::vector<std::vector<double>> coef(kk, std::vector<double>(m));
        double rmax = 0.0;  // over the wanted roots
        for (int i = 0; i < kk; ++i)
            for (int j = 0; j < m; ++j) coef[i][j] = C[(size_t)j * m + i];
        coef_last = coef;
        ritz_in_V = false;
        if (mode == 1) {
            // all kk residuals in one pass over the subspace, into the free slots V[m..m+kk)
            std::vector<double> Cm((size_t)m * kk);
            std::vector<double *> Rp(kk);
            for (int j = 0; j < m; ++j)
                for (int i = 0; i < kk; ++i) Cm[(size_t)j * kk + i] = coef[i][j];
            for (int i = 0; i < kk; ++i) Rp[i] = V[m + i];
            block_combine(N, m, kk, Cm.data(), w.data(), (const double *const *)V.data(),
                          (const double *const *)AV.data(), Rp.data(), rn.data(), st);
        } else {
            for (int i = 0; i < kk; ++i)
                rn[i] = ritz_and_residual(N, m, coef[i].data(), w[i], (const double *const *)V.data(),
                                          (const double *const *)AV.data(), R[i].as<double>(), st);
        }
        for (int i = 0; i < std::min(nroots, kk); ++i) rmax = std::max(rmax, rn[i]);
        for (int i = 0; i < nroots; ++i) energies[i] = w[i];  // current estimates
        out.max_residual = rmax;
        if (verbose) {  // progress goes to stderr: stdout belongs to the caller
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::fprintf(stderr, "  Step %03d [%s] m=%2d  max residual: %.4e  e0: %.14f  t=%.1fs\n", it,
                         phase == 1 ? "fp32" : "fp64", m, rmax, w[0], el);
            std::fflush(stderr);
        }
        const double cur_tol = phase == 1 ? std::max(switch_tol, tol) : tol;
        // FP32 phase: leave when converged to switch_tol or when the residual stagnates: no 10%
        // gain over 8 iterations (the maximum over several roots moves unevenly: with 4 roots at
        // 16 orbitals it regularly stays above its best for 4 iterations while still falling
        // ~7% per iteration, and a 4-iteration window handed over to FP64 at 1e-3)
        if (phase == 1) {
            if (rmax < 0.9 * best[0]) {
                best[0] = rmax;
                stall = 0;
            } else if (++stall >= 8 && rmax < 1e-3) {
                stall = 0;
                rmax = 0.0;  // force the switch
            }
        }
        if (rmax < cur_tol) {
            if (phase == 1) {
                // the kk lowest Ritz vectors -> new subspace (combined in place), FP64 sigma from
                // here on
                {
                    std::vector<double> Cm((size_t)m * kk);
                    for (int j = 0; j < m; ++j)
                        for (int i = 0; i < kk; ++i) Cm[(size_t)j * kk + i] = coef[i][j];
                    rotate_inplace(N, m, kk, Cm.data(), V.data(), st);
                }
                m = 0;
                for (int i = 0; i < kk; ++i) {
                    mgs(N, m, (const double *const *)V.data(), V[i], st);
                    normalize(N, V[i], st);
                    ++m;
                }
                phase = 0;
                ritz_in_V = true;
                out.refine_iter = it;
                apply(0, m);
                set_cols(0, m);
                continue;
            }
            for (int i = 0; i < nroots; ++i) energies[i] = w[i];
            if (vecs)
                for (int i = 0; i < nroots; ++i)
                    ritz_vector(N, m, coef[i].data(), (const double *const *)V.data(), vecs + (size_t)i * N, st);
            out.converged = 1;
            break;
        }
        // correction vectors for the unconverged Ritz pairs (the wanted roots, plus the extra
        // pairs of the block mode)
        std::vector<int> todo;
        for (int i = 0; i < kk; ++i)
            if (rn[i] >= cur_tol) todo.push_back(i);
        if (mode == 0 && m + (int)todo.size() > maxspace) restart();  // (the block mode restarted above)
        const int first = m;
        if (mode == 1) {
            // preconditioned residuals of up to 8 pairs at once (residual i sits in V[m + i];
            // the slot pointers are permuted so the b-th pair of todo moves to V[m + b]),
            // orthogonalised against the subspace by two passes of block classical Gram-Schmidt,
            // then against each other
            int b = 0;
            std::vector<double> nb0;
            for (int i : todo) {
                if (m + b >= maxspace || b >= 8) break;
                std::swap(V[m + b], V[m + i]);  // i >= b: earlier swaps only touched lower slots
                double *t = V[m + b];
                precondition(N, t, diag.as<double>(), w[i], std::max(rn[i] * 1e-2, 1e-4), st);
                nb0.push_back(std::sqrt(dot(N, t, t, st)));
                ++b;
            }
            if (b > 0 && m > 0) {
                std::vector<double> G((size_t)m * b);
                for (int pass = 0; pass < 2; ++pass) {
                    block_dots(N, m, (const double *const *)V.data(), b, (const double *const *)&V[m], G.data(), st);
                    block_subtract(N, m, (const double *const *)V.data(), G.data(), b, &V[m], st);
                }
            }
            int kept = 0;
            for (int j = 0; j < b; ++j) {
                std::swap(V[m + kept], V[m + j]);  // candidate into the next free slot
                double *t = V[m + kept];
                double na = mgs(N, kept, (const double *const *)&V[m], t, st);
                if (kept > 0 && na < 0.717 * nb0[j]) na = mgs(N, kept, (const double *const *)&V[m], t, st);
                if (na < 1e-12 * std::max(1.0, nb0[j])) continue;  // lies in the subspace
                normalize(N, t, st);
                ++kept;
            }
            m += kept;
        }
        for (int i : (mode == 1 ? std::vector<int>() : todo)) {
            if (m >= maxspace) break;
            const double shift = std::max(rn[i] * 1e-2, 1e-4);
            double *t = V[m];
            HIPOTF_CHECK(hipMemcpyAsync(t, R[i].p, N * sizeof(double), hipMemcpyDeviceToDevice, st));
            precondition(N, t, diag.as<double>(), w[i], shift, st);
            const double nb = std::sqrt(dot(N, t, t, st));
            double na = mgs(N, m, (const double *const *)V.data(), t, st);
            if (na < 0.717 * nb) na = mgs(N, m, (const double *const *)V.data(), t, st);
            if (na < 1e-12 * std::max(1.0, nb)) continue;  // lies in the subspace
            normalize(N, t, st);
            ++m;
        }
        if (m == first) {
            // breakdown: no new direction although the residual is above tol
            // (jl/davidson.jl:135-143 returns the current Ritz vector)
            write_current();
            out.converged = 2;
            break;
        }
        apply(first, m - first);
        set_cols(first, m - first);
    }
    if (out.converged == 0) write_current();  // maxiter: the latest estimates
    HIPOTF_CHECK(hipDeviceSynchronize());
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

}  // namespace hipotf
