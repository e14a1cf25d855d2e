// Two-body Hamiltonian in integral form, bound to a device basis.
//
// The operator is the one the reference builds as Pauli terms in
// generate_hamiltonian_real_tmpl (src/include/ham/ham.hpp:998-1069):
//   H = E_core + sum_pq h_pq sum_s a+_ps a_qs + 1/2 sum_pqrs (ps|qr) sum_st a+_ps a+_qt a_rt a_ss
// with the interleaved Jordan-Wigner order.  Here it is split as
//   H = E ( G_alpha (x) 1 + 1 (x) G_beta + H_alpha-beta + E_core ) E
// where E is the diagonal +-1 gauge from interleaved to alpha-major determinant order and
//   G[I,I]   = sum_{p in I} h_pp + 1/2 sum_{p,q in I} [(pp|qq) - (pq|qp)]
//   G[I,J]   = sgn [h_rs + sum_{j in I, j != r} ((rs|jj) - (rj|js))]        J = I - r + s
//   G[I,J]   = sgn1 sgn2 [(r1 s1|r2 s2) - (r1 s2|r2 s1)]                    J = I - r1 - r2 + s1 + s2
//   H_ab     = sum_pqrs (pq|rs) E^a_pq E^b_rs     (p = q and r = s terms included)
// Integrals must carry the 8-fold permutational symmetry of real orbitals (checked).
#include "hipotf.hpp"

#include <algorithm>
#include <cmath>
#include <thread>

namespace hipotf {

i64 host_find_str(const SpinSpace &S, u64 s);

static inline int popc64(u64 v) { return __builtin_popcountll(v); }
static inline u64 between_mask(int r, int s) {
    int lo = r < s ? r : s, hi = r < s ? s : r;
    return ((1ull << hi) - 1) & ~((1ull << (lo + 1)) - 1);
}

template <typename F>
static void par_for(i64 n, F f) {
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (n < 1024 || nt == 1) {
        for (i64 i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([=, &f]() {
            for (i64 i = t; i < n; i += nt) f(i);
        });
    for (auto &x : th) x.join();
}

static void build_same_spin(SameSpinTable &T, const SpinSpace &S, const std::vector<double> &h,
                            const std::vector<double> &eri, int n) {
    const int k = S.k, nv = n - k;
    T.L = 1 + S.nsx + S.ndx;
    const size_t L = T.L;
    std::vector<int32_t> J((size_t)S.N * L);
    std::vector<double> C((size_t)S.N * L);
    std::vector<double> gd(S.N);
    // P[I][g] = first entry of row I whose source index is >= range_start(g)
    std::vector<int32_t> P((size_t)S.N * (SS_RANGES + 1));
    auto range_start = [&](int g) { return (i64)g * S.N / SS_RANGES; };
    auto E = [&](int p, int q, int r, int s) { return eri[(((size_t)p * n + q) * n + r) * n + s]; };
    par_for(S.N, [&](i64 I) {
        const u64 s = S.strs[I];
        int occ[64], vir[64], no = 0, nvv = 0;
        for (int p = 0; p < n; ++p) {
            if ((s >> p) & 1) occ[no++] = p;
            else vir[nvv++] = p;
        }
        int32_t *Jr = &J[(size_t)I * L];
        double *Cr = &C[(size_t)I * L];
        double d = 0.0;
        for (int a = 0; a < k; ++a) {
            int p = occ[a];
            d += h[p * n + p];
            for (int b = 0; b < k; ++b) {
                int q = occ[b];
                d += 0.5 * (E(p, p, q, q) - E(p, q, q, p));
            }
        }
        gd[I] = d;
        Jr[0] = (int32_t)I;
        Cr[0] = d;
        size_t l = 1;
        for (int io = 0; io < k; ++io)
            for (int iv = 0; iv < nv; ++iv, ++l) {
                int r = occ[io], t = vir[iv];
                u64 js = s ^ (1ull << r) ^ (1ull << t);
                double v = h[r * n + t];
                for (int a = 0; a < k; ++a) {
                    int j = occ[a];
                    if (j == r) continue;
                    v += E(r, t, j, j) - E(r, j, j, t);
                }
                bool neg = popc64(s & between_mask(r, t)) & 1;
                Jr[l] = (int32_t)host_find_str(S, js);
                Cr[l] = neg ? -v : v;
            }
        for (int a1 = 0; a1 < k; ++a1)
            for (int a2 = a1 + 1; a2 < k; ++a2)
                for (int b1 = 0; b1 < nv; ++b1)
                    for (int b2 = b1 + 1; b2 < nv; ++b2, ++l) {
                        int r1 = occ[a1], r2 = occ[a2], s1 = vir[b1], s2 = vir[b2];
                        u64 js = s ^ (1ull << r1) ^ (1ull << r2) ^ (1ull << s1) ^ (1ull << s2);
                        // apply E_{r1 s1} then E_{r2 s2} to J to reach I
                        int g1 = popc64(js & between_mask(r1, s1)) & 1;
                        u64 j1 = js ^ (1ull << s1) ^ (1ull << r1);
                        int g2 = popc64(j1 & between_mask(r2, s2)) & 1;
                        double v = E(r1, s1, r2, s2) - E(r1, s2, r2, s1);
                        Jr[l] = (int32_t)host_find_str(S, js);
                        Cr[l] = (g1 ^ g2) ? -v : v;
                    }
        // ascending source index, so a pass over one source range reads a contiguous piece
        std::vector<std::pair<int32_t, double>> row(L);
        for (size_t e = 0; e < L; ++e) row[e] = {Jr[e], Cr[e]};
        std::sort(row.begin(), row.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        int32_t *Pr = &P[(size_t)I * (SS_RANGES + 1)];
        int rg = 0;
        for (size_t e = 0; e < L; ++e) {
            Jr[e] = row[e].first;
            Cr[e] = row[e].second;
            while (rg <= SS_RANGES && Jr[e] >= range_start(rg)) Pr[rg++] = (int32_t)e;
        }
        while (rg <= SS_RANGES) Pr[rg++] = (int32_t)L;
    });
    for (size_t i = 0; i < J.size(); ++i)
        HIPOTF_REQUIRE(J[i] >= 0 && J[i] < S.N, "internal: same-spin table index out of range");
    T.J.upload(J);
    T.P.upload(P);
    T.c64.upload(C);
    std::vector<float> C32(C.begin(), C.end());
    T.c32.upload(C32);
    T.gdiag.upload(gd);
    T.N = S.N;
    T.hJ = std::move(J);
    T.hC = std::move(C);
}

const RowBlockTable &row_blocks(SameSpinTable &T, int rb) {
    HIPOTF_REQUIRE(rb == 4 || rb == 8 || rb == 16, "row block size must be 4, 8 or 16");
    const int slot = rb == 4 ? 0 : (rb == 8 ? 1 : 2);
    if (T.blocks[slot]) return *T.blocks[slot];
    if (T.dev_built) {
        auto R = std::make_unique<RowBlockTable>();
        row_blocks_dev(T, *R, rb);
        T.blocks[slot] = std::move(R);
        return *T.blocks[slot];
    }
    const i64 N = T.N;
    const size_t L = T.L;
    HIPOTF_REQUIRE(T.hJ.size() == (size_t)N * L, "internal: same-spin host table missing");
    auto R = std::make_unique<RowBlockTable>();
    R->rb = rb;
    R->nblk = (N + rb - 1) / rb;
    const i64 nblk = R->nblk;
    // the ascending, duplicate-free union of the sources of rows [g*rb, g*rb + rb)
    auto block_union = [&](i64 g, std::vector<int32_t> &u) {
        const i64 r0 = g * rb, r1 = std::min<i64>(N, r0 + rb);
        u.assign(T.hJ.begin() + r0 * L, T.hJ.begin() + r1 * L);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
    };
    std::vector<i64> off(nblk + 1, 0);
    par_for(nblk, [&](i64 g) {
        std::vector<int32_t> u;
        block_union(g, u);
        off[g + 1] = (i64)u.size();
    });
    for (i64 g = 0; g < nblk; ++g) off[g + 1] += off[g];
    const i64 E = off[nblk];
    HIPOTF_REQUIRE(E < (1ll << 31), "row-block table too large for 32-bit offsets");
    R->entries = E;
    std::vector<int32_t> J((size_t)E);
    std::vector<double> C((size_t)E * rb, 0.0);
    std::vector<int32_t> P((size_t)nblk * (SS_RANGES + 1));
    auto range_start = [&](int g) { return (i64)g * N / SS_RANGES; };
    par_for(nblk, [&](i64 g) {
        std::vector<int32_t> u;
        block_union(g, u);
        std::copy(u.begin(), u.end(), J.begin() + off[g]);
        const i64 r0 = g * rb, r1 = std::min<i64>(N, r0 + rb);
        for (i64 I = r0; I < r1; ++I)
            for (size_t e = 0; e < L; ++e) {
                const int32_t j = T.hJ[I * L + e];
                const size_t pos = std::lower_bound(u.begin(), u.end(), j) - u.begin();
                C[((size_t)off[g] + pos) * rb + (size_t)(I - r0)] += T.hC[I * L + e];
            }
        for (int rg = 0; rg <= SS_RANGES; ++rg) {
            const size_t pos = rg == SS_RANGES ? u.size()
                                               : std::lower_bound(u.begin(), u.end(), (int32_t)range_start(rg)) - u.begin();
            P[(size_t)g * (SS_RANGES + 1) + rg] = (int32_t)(off[g] + (i64)pos);
        }
    });
    R->J.upload(J);
    R->P.upload(P);
    R->c64.upload(C);
    std::vector<float> C32(C.begin(), C.end());
    R->c32.upload(C32);
    T.blocks[slot] = std::move(R);
    return *T.blocks[slot];
}

// Determinants as interleaved spin-orbital occupations: bit 2p = alpha p, bit 2p+1 = beta p
// (src/include/ham/ham.hpp:519-551).
u64 interleave_occ(u32 a, u32 b) {
    u64 x = 0;
    for (int p = 0; p < 32; ++p) x |= (u64)((a >> p) & 1u) << (2 * p) | (u64)((b >> p) & 1u) << (2 * p + 1);
    return x;
}

// apply one ladder operator at spin-orbital pos to occupation x (Jordan-Wigner sign: parity of
// the occupied spin-orbitals below pos); false if it annihilates the state
bool ladder_apply(u64 &x, int &par, int pos, bool create) {
    const u64 bit = 1ull << pos;
    if (((x & bit) != 0) == create) return false;
    par += __builtin_popcountll(x & (bit - 1));
    x ^= bit;
    return true;
}

// <bra| H |ket> of the two-body operator (ecore, h[p*n+q], eri[((p*n+q)*n+r)*n+s] = (pq|rs)
// chemist order), by the Slater-Condon rules in the interleaved spin-orbital order, with the
// signs of the ladder operators applied to ket:
//   H = ecore + sum h_PQ a+_P a_Q + 1/2 sum (PQ|RS) a+_P a+_R a_S a_Q
double slater_condon(int n, const double *hh, const double *eri, double ecore, u64 bra, u64 ket) {
    auto h1 = [&](int P, int Q) { return ((P ^ Q) & 1) ? 0.0 : hh[(size_t)(P >> 1) * n + (Q >> 1)]; };
    auto g2 = [&](int P, int Q, int R, int S) {
        if (((P ^ Q) & 1) || ((R ^ S) & 1)) return 0.0;
        return eri[((((size_t)(P >> 1) * n + (Q >> 1)) * n + (R >> 1)) * n) + (S >> 1)];
    };
    const u64 diff = bra ^ ket;
    const int nd = __builtin_popcountll(diff);
    std::vector<int> occ;
    for (u64 m = ket; m; m &= m - 1) occ.push_back(__builtin_ctzll(m));
    if (nd == 0) {
        double e = ecore;
        for (int P : occ) e += h1(P, P);
        for (int P : occ)
            for (int R : occ)
                if (P != R) e += 0.5 * (g2(P, P, R, R) - g2(P, R, R, P));
        return e;
    }
    if (nd == 2) {
        const int Q = __builtin_ctzll(ket & diff), P = __builtin_ctzll(bra & diff);
        if (!(ket & diff) || !(bra & diff)) return 0.0;
        u64 x = ket;
        int par = 0;
        ladder_apply(x, par, Q, false);
        ladder_apply(x, par, P, true);
        double v = h1(P, Q);
        for (int R : occ)
            if (R != Q) v += g2(P, Q, R, R) - g2(P, R, R, Q);
        return (par & 1) ? -v : v;
    }
    if (nd == 4) {
        const u64 an = ket & diff, cr = bra & diff;
        if (__builtin_popcountll(an) != 2 || __builtin_popcountll(cr) != 2) return 0.0;
        const int Q1 = __builtin_ctzll(an), Q2 = 63 - __builtin_clzll(an);
        const int P1 = __builtin_ctzll(cr), P2 = 63 - __builtin_clzll(cr);
        u64 x = ket;
        int par = 0;
        ladder_apply(x, par, Q1, false);
        ladder_apply(x, par, Q2, false);
        ladder_apply(x, par, P2, true);
        ladder_apply(x, par, P1, true);
        const double v = g2(P1, Q1, P2, Q2) - g2(P1, Q2, P2, Q1);
        return (par & 1) ? -v : v;
    }
    return 0.0;
}

void integrals_from_reference_layout(int n, const double *h_cm, const double *g_ref,
                                     std::vector<double> &h, std::vector<double> &eri) {
    HIPOTF_REQUIRE(n >= 1 && n <= 32, "norb must be in [1, 32]");
    HIPOTF_REQUIRE(h_cm && g_ref, "integral pointers are NULL");
    const size_t n2 = (size_t)n * n, n3 = n2 * n, n4 = n3 * n;
    h.resize(n2);
    eri.resize(n4);
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) h[p * n + q] = h_cm[p + (size_t)q * n];
    // (pq|rs) = two_body_mo[p, r, s, q]  (two_body_mo[p,q,r,s] = (ps|qr), column-major)
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q)
            for (int r = 0; r < n; ++r)
                for (int s = 0; s < n; ++s)
                    eri[((size_t)(p * n + q) * n + r) * n + s] =
                        g_ref[p + (size_t)r * n + (size_t)s * n2 + (size_t)q * n3];
    double scale = 0.0;
    for (double v : eri) scale = std::max(scale, std::fabs(v));
    for (double v : h) scale = std::max(scale, std::fabs(v));
    const double tol = 1e-10 * std::max(1.0, scale);
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) {
            HIPOTF_REQUIRE(std::fabs(h[p * n + q] - h[q * n + p]) <= tol,
                           "one-electron integrals are not symmetric");
            for (int r = 0; r < n; ++r)
                for (int s = 0; s < n; ++s) {
                    const double v = eri[((size_t)(p * n + q) * n + r) * n + s];
                    const double v1 = eri[((size_t)(q * n + p) * n + r) * n + s];
                    const double v2 = eri[((size_t)(r * n + s) * n + p) * n + q];
                    HIPOTF_REQUIRE(std::fabs(v - v1) <= tol && std::fabs(v - v2) <= tol,
                                   "two-electron integrals lack 8-fold symmetry");
                }
        }
}

DevHam *make_dev_ham(DevBasis *b, int n, const std::vector<double> &h, const std::vector<double> &eri,
                     double ecore) {
    HIPOTF_REQUIRE(b != nullptr, "basis handle is NULL");
    HIPOTF_REQUIRE(n == b->host.norb, "orbital count of the operator does not match the basis");
    auto H = std::make_unique<DevHam>();
    H->basis = b;
    H->n = n;
    H->ecore = ecore;
    H->h = h;
    H->eri = eri;
    // ERI_PAD zero entries after the table: kernels that fill fixed-size LDS tiles from one
    // (pq|..) row may read up to that far past the last row
    std::vector<double> e64(H->eri);
    e64.resize(H->eri.size() + ERI_PAD, 0.0);
    H->eri64.upload(e64);
    std::vector<float> e32(e64.begin(), e64.end());
    H->eri32.upload(e32);
    const HostBasis &hb = b->host;
    H->h64.upload(H->h);
    // same-spin tables on the device where the rows fit its in-LDS sort, else on the host
    if (!build_same_spin_dev(H->tA, *b->pA, H->h64, H->eri64, n)) build_same_spin(H->tA, hb.A, H->h, H->eri, n);
    H->ptA = &H->tA;
    if (b->pB == b->pA) {
        H->ptB = &H->tA;
    } else {
        if (!build_same_spin_dev(H->tB, *b->pB, H->h64, H->eri64, n))
            build_same_spin(H->tB, hb.B, H->h, H->eri, n);
        H->ptB = &H->tB;
    }
    // the default row blocks now, so that the first sigma does not pay for them
    for (int prec = 0; prec < 2; ++prec) {
        const int rb = tuning(prec).rs_block;
        if (rb == 4 || rb == 8 || rb == 16) {
            row_blocks(*H->ptA, rb);
            row_blocks(*H->ptB, rb);
        }
    }
    build_jab_dev(H->jab, *b->pA, n, H->eri64);
    return H.release();
}

}  // namespace hipotf
