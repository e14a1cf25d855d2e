// extern "C" boundary of libhipotf.  Declarations and reference citations: include/hipotf.h.
//
// Error convention: no entry point exits or throws across the boundary.  Failures record a
// thread-local message (hipotf_last_error) and return NULL / NaN / nonzero status; void entry
// points only record the message.
#include "hipotf.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <mutex>
#include <unordered_map>

namespace hipotf {
const char *last_error_cstr();
double ritz_and_residual(i64 N, int dim, const double *coeffs, double min_e, const double *const *V,
                         const double *const *AV, double *axt, hipStream_t st);
void ritz_vector(i64 N, int dim, const double *coeffs, const double *const *V, double *xt, hipStream_t st);
void precondition(i64 N, double *axt, const double *diags, double min_e, double shift, hipStream_t st);
double mgs(i64 N, int dim, const double *const *V, double *vt, hipStream_t st);
double normalize(i64 N, double *vec, hipStream_t st);
double dot(i64 N, const double *a, const double *b, hipStream_t st);
void dot_dev(i64 N, const double *a, const double *b, double *dres, hipStream_t st);
void multidot(i64 N, int dim, const double *const *V, const double *a, double *out_host, hipStream_t st);
void validate_pool(int n, i64 nops, const int32_t *offsets, const int32_t *orbs);
struct DavidsonOut {
    int iterations = 0;
    i64 sigma64 = 0, sigma32 = 0;
    double max_residual = 0.0;
    double seconds = 0.0;
    int refine_iter = -1;
    int converged = 0;
};
DavidsonOut davidson(DevHam *H, int nroots, int mode, int block, double tol, double switch_tol, int maxiter,
                     int maxspace, int verbose, const double *guess, double *energies, double *vecs);

// ------------------------------------------------------------------------------------------
// host network: the flat SVD-group arrays of build_network_otf (src/include/ham/otf.hpp:223-349)
// ------------------------------------------------------------------------------------------
struct HostGroup {
    u32 ax, bx;
    int rank, nza, nzb;
    std::vector<u32> zas, zbs;
    std::vector<double> wa, wb;  // [r * nz + k]
};

struct HostNet {
    int norb = 0;
    std::vector<HostGroup> groups;
    std::unordered_map<u64, int> by_x;  // (ax << 32 | bx) -> group
    // f_g(a, b) = sum_r Pa_r(a) Pb_r(b), P_r(s) = sum_k w[r,k] (-1)^popc(s & z_k)
    // (src/include/core/math.hpp:37-124); the matrix element <(a,b)^x| H |(a,b)>.
    double element(u32 ax, u32 bx, u32 a, u32 b) const {
        auto it = by_x.find(((u64)ax << 32) | bx);
        if (it == by_x.end()) return 0.0;
        const HostGroup &g = groups[it->second];
        double v = 0.0;
        for (int r = 0; r < g.rank; ++r) {
            double pa = 0.0, pb = 0.0;
            for (int k = 0; k < g.nza; ++k)
                pa += (__builtin_popcount(a & g.zas[k]) & 1) ? -g.wa[r * g.nza + k] : g.wa[r * g.nza + k];
            for (int k = 0; k < g.nzb; ++k)
                pb += (__builtin_popcount(b & g.zbs[k]) & 1) ? -g.wb[r * g.nzb + k] : g.wb[r * g.nzb + k];
            v += pa * pb;
        }
        return v;
    }
};

// ------------------------------------------------------------------------------------------
// device operator handle: either a Hamiltonian (integrals) or a generator pool; the
// basis-dependent tables are built on first use with a basis and cached under its uid
// ------------------------------------------------------------------------------------------
struct DevNet {
    int kind = 0;  // 1 Hamiltonian (integrals), 2 pool, 3 generic Pauli-group network
    int n = 0;
    DevGenNet *gen = nullptr;  // kind 3
    ~DevNet() { destroy_dev_gennet(gen); }
    double ecore = 0.0;
    std::vector<double> h, eri;
    std::vector<int32_t> offsets, orbs;
    std::map<i64, std::unique_ptr<DevHam>> hams;
    std::map<i64, std::unique_ptr<DevPool>> pools;
    DevHam *ham_for(DevBasis *b) {
        HIPOTF_REQUIRE(kind != 3, "this network is evaluated group by group (it is not a spin-free two-body "
                                  "Hamiltonian): only hvec_cuda and get_diags_elements_cuda take it");
        HIPOTF_REQUIRE(kind == 1, "operator handle is not a Hamiltonian");
        auto it = hams.find(b->uid);
        if (it != hams.end()) return it->second.get();
        DevHam *H = make_dev_ham(b, n, h, eri, ecore);
        hams[b->uid].reset(H);
        return H;
    }
    // the Hamiltonian's tables for b if they exist (no allocation), else an error
    DevHam *ham_prepared(DevBasis *b) {
        HIPOTF_REQUIRE(kind != 3, "this network is evaluated group by group (it is not a spin-free two-body "
                                  "Hamiltonian): only hvec_cuda and get_diags_elements_cuda take it");
        HIPOTF_REQUIRE(kind == 1, "operator handle is not a Hamiltonian");
        auto it = hams.find(b->uid);
        HIPOTF_REQUIRE(it != hams.end(), "the Hamiltonian's device tables for this basis are not built: call "
                                         "hipotf_prepare (or hvec_cuda) once first");
        return it->second.get();
    }
    DevPool *pool_for(DevBasis *b) {
        HIPOTF_REQUIRE(kind == 2, "operator handle is not a generator pool");
        auto it = pools.find(b->uid);
        if (it != pools.end()) return it->second.get();
        HIPOTF_REQUIRE(n == b->host.norb, "orbital count of the pool does not match the basis");
        DevPool *P = make_dev_pool(b, (i64)offsets.size() - 1, offsets.data(), orbs.data());
        pools[b->uid].reset(P);
        return P;
    }
};

static inline hipStream_t S0() { return (hipStream_t)0; }

static DevBasis *as_basis(void *p) {
    HIPOTF_REQUIRE(p != nullptr, "device basis handle is NULL");
    return static_cast<DevBasis *>(p);
}
static DevNet *as_net(void *p) {
    HIPOTF_REQUIRE(p != nullptr, "device operator handle is NULL");
    return static_cast<DevNet *>(p);
}

// ---- decoding a reference-style network by probing matrix elements -----------------------
static double probe_diag(const HostNet &N, u32 a, u32 b) { return N.element(0, 0, a, b); }

static void decode_hamiltonian(const HostNet &N, DevNet &D) {
    const int n = N.norb;
    D.kind = 1;
    D.n = n;
    D.h.assign((size_t)n * n, 0.0);
    D.eri.assign((size_t)n * n * n * n, 0.0);
    auto E = [&](int p, int q, int r, int s) -> double & {
        return D.eri[(((size_t)p * n + q) * n + r) * n + s];
    };
    const double e0 = probe_diag(N, 0, 0);
    D.ecore = e0;
    std::vector<double> ea(n), eb(n);
    for (int p = 0; p < n; ++p) {
        ea[p] = probe_diag(N, 1u << p, 0);
        eb[p] = probe_diag(N, 0, 1u << p);
        D.h[p * n + p] = ea[p] - e0;
    }
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q)
            if (p != q) D.h[p * n + q] = N.element((1u << p) | (1u << q), 0, 1u << q, 0);
    // Coulomb (pp|rr) from two-electron diagonals
    for (int p = 0; p < n; ++p)
        for (int r = 0; r < n; ++r) E(p, p, r, r) = probe_diag(N, 1u << p, 1u << r) - ea[p] - eb[r] + e0;
    // (pq|rr), p != q: alpha move q -> p with a spectator beta electron in r; the interleaved
    // sign is (-1)^[min(p,q) <= r < max(p,q)]
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) {
            if (p == q) continue;
            for (int r = 0; r < n; ++r) {
                double f = N.element((1u << p) | (1u << q), 0, 1u << q, 1u << r);
                const int lo = std::min(p, q), hi = std::max(p, q);
                if (r >= lo && r < hi) f = -f;
                E(p, q, r, r) = f - D.h[p * n + q];
            }
        }
    // (pq|rs), p != q, r != s: alpha q -> p and beta s -> r on a two-electron state; gauge signs
    // (-1)^([r < p] + [s < q]) convert interleaved to alpha-major order
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) {
            if (p == q) continue;
            for (int r = 0; r < n; ++r)
                for (int s = 0; s < n; ++s) {
                    if (r == s) continue;
                    double f = N.element((1u << p) | (1u << q), (1u << r) | (1u << s), 1u << q, 1u << s);
                    if (((r < p) + (s < q)) & 1) f = -f;
                    E(p, q, r, s) = f;
                }
        }
    // (pp|rs), r != s by symmetry
    for (int p = 0; p < n; ++p)
        for (int r = 0; r < n; ++r)
            for (int s = 0; s < n; ++s)
                if (r != s) E(p, p, r, s) = E(r, s, p, p);
}

// ---- checking a decode against the network it came from ------------------------------------
// deterministic pseudo-random subset of the bits of m (splitmix64 of the key)
static u32 probe_subset(u32 m, u64 key) {
    u64 z = key + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return m & (u32)(z ^ (z >> 32));
}

// The decode reads a few matrix elements per excitation pattern and assumes the network is a
// spin-free two-body Hamiltonian with real 8-fold-symmetric integrals.  Any other network
// (spin models, S^2 penalties, three-body terms, missing Jordan-Wigner strings) would decode into
// a different operator, so every number- and spin-conserving pattern of up to two electrons
// per spin is re-evaluated from the decoded integrals on probe determinants -- every
// annihilated/created split of the pattern, with no, all, single and pseudo-random spectator
// electrons of each spin -- and compared with the network's own element
// (src/include/core/math.hpp:37-124).  Networks with groups outside those patterns are
// rejected outright.
static void verify_hamiltonian(const HostNet &N, const DevNet &D) {
    const int n = N.norb;
    const u32 full = n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1);
    for (size_t g = 0; g < N.groups.size(); ++g) {
        const int ka = __builtin_popcount(N.groups[g].ax), kb = __builtin_popcount(N.groups[g].bx);
        HIPOTF_REQUIRE((N.groups[g].ax & ~full) == 0 && (N.groups[g].bx & ~full) == 0 && ka % 2 == 0 &&
                           kb % 2 == 0 && ka + kb <= 4,
                       "network group " + std::to_string(g) +
                           " is not a number- and spin-conserving one- or two-body term; only two-body "
                           "Hamiltonian networks are decoded (generic Pauli networks are not supported)");
    }
    double scale = std::fabs(D.ecore);
    for (double v : D.h) scale = std::max(scale, std::fabs(v));
    for (double v : D.eri) scale = std::max(scale, std::fabs(v));
    const double tol = 1e-9 * std::max(1.0, scale);
    // the sigma kernels use (pq|rs) = (qp|rs) = (rs|pq) (real orbitals): a spin-free operator
    // without that symmetry (an S^2 penalty: (pq|qp) without (pq|pq)) is not representable
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) {
            HIPOTF_REQUIRE(std::fabs(D.h[p * n + q] - D.h[q * n + p]) <= tol,
                           "network's one-body part is not symmetric (generic Pauli networks are not supported)");
            for (int r = 0; r < n; ++r)
                for (int s = 0; s < n; ++s) {
                    const double v = D.eri[(((size_t)p * n + q) * n + r) * n + s];
                    const double v1 = D.eri[(((size_t)q * n + p) * n + r) * n + s];
                    const double v2 = D.eri[(((size_t)r * n + s) * n + p) * n + q];
                    HIPOTF_REQUIRE(std::fabs(v - v1) <= tol && std::fabs(v - v2) <= tol,
                                   "network is not a spin-free two-body Hamiltonian with real 8-fold symmetric "
                                   "integrals (generic Pauli networks are not supported)");
                }
        }
    double worst = 0.0;
    u64 nprobe = 0;
    auto check = [&](u32 ax, u32 bx, u32 a, u32 b) {
        const double fn = N.element(ax, bx, a, b);
        const double fs = slater_condon(D.n, D.h.data(), D.eri.data(), D.ecore, interleave_occ(a ^ ax, b ^ bx), interleave_occ(a, b));
        const double d = std::fabs(fn - fs);
        ++nprobe;
        if (!(d <= tol)) {
            char msg[320];
            std::snprintf(msg, sizeof msg,
                          "network is not a spin-free two-body Hamiltonian: element <a^%#x,b^%#x|H|a=%#x,b=%#x> "
                          "is %.12g in the network and %.12g from the decoded integrals (generic Pauli networks "
                          "are not supported)",
                          ax, bx, a, b, fn, fs);
            throw Error(msg);
        }
        worst = std::max(worst, d);
    };
    std::vector<int> bits;
    auto subsets = [](u32 m, int k, std::vector<u32> &out) {  // subsets of m with k bits
        out.clear();
        std::vector<int> b;
        for (u32 t = m; t; t &= t - 1) b.push_back(__builtin_ctz(t));
        for (u32 s = 0; s < (1u << b.size()); ++s) {
            if (__builtin_popcount(s) != k) continue;
            u32 v = 0;
            for (size_t i = 0; i < b.size(); ++i)
                if ((s >> i) & 1) v |= 1u << b[i];
            out.push_back(v);
        }
    };
    std::vector<u32> pa_list, pb_list, sa_list, sb_list;
    // every pattern: alpha / beta flip masks of 0, 2 or 4 orbitals, at most 4 in total
    std::vector<u32> masks[5];
    for (int k : {0, 2, 4}) subsets(full, k, masks[k]);
    for (int ka : {0, 2, 4})
        for (int kb : {0, 2, 4}) {
            if (ka + kb > 4) continue;
            for (u32 ax : masks[ka])
                for (u32 bx : masks[kb]) {
                    subsets(ax, ka / 2, pa_list);
                    subsets(bx, kb / 2, pb_list);
                    const u32 fa = full & ~ax, fb = full & ~bx;
                    const int nrand = ka + kb == 0 ? 256 : 4;
                    for (u32 anna : pa_list)
                        for (u32 annb : pb_list) {
                            check(ax, bx, anna, annb);
                            check(ax, bx, anna | fa, annb | fb);
                            if (ka + kb <= 2) {  // single spectators (the diagonal: also pairs)
                                for (u32 t = fa; t; t &= t - 1) check(ax, bx, anna | (t & -t), annb);
                                for (u32 t = fb; t; t &= t - 1) check(ax, bx, anna, annb | (t & -t));
                            }
                            if (ka + kb == 0)
                                for (u32 s = fa; s; s &= s - 1)
                                    for (u32 t = fb; t; t &= t - 1) check(ax, bx, (s & -s), (t & -t));
                            if (ka + kb == 0)
                                for (u32 s = fa; s; s &= s - 1)
                                    for (u32 t = s & (s - 1); t; t &= t - 1) {
                                        check(ax, bx, (s & -s) | (t & -t), 0);
                                        check(ax, bx, 0, (s & -s) | (t & -t));
                                    }
                            for (int r = 0; r < nrand; ++r) {
                                const u64 key = ((u64)ax << 32 | bx) * 1315423911ull + (u64)anna * 2654435761ull +
                                                annb * 40503ull + (u64)r * 0x632BE59BD9B4E019ull;
                                check(ax, bx, anna | probe_subset(fa, key), annb | probe_subset(fb, key ^ 0xA5A5A5A5ull));
                            }
                        }
                }
        }
    (void)worst;
    (void)nprobe;
}

// recognise each pool group as +-(T - T^dagger) for a product T of ladder operators
static void decode_pool(const HostNet &N, DevNet &D) {
    const int n = N.norb;
    D.kind = 2;
    D.n = n;
    D.offsets.assign(1, 0);
    D.orbs.clear();
    for (size_t g = 0; g < N.groups.size(); ++g) {
        const HostGroup &G = N.groups[g];
        const u32 ax = G.ax, bx = G.bx;
        const int ka = __builtin_popcount(ax), kb = __builtin_popcount(bx);
        HIPOTF_REQUIRE(ka % 2 == 0 && kb % 2 == 0 && ka + kb >= 2 && ka + kb <= 8,
                       "pool group " + std::to_string(g) + " is not a number-conserving excitation");
        std::vector<int> oa, ob;
        for (int p = 0; p < n; ++p) {
            if ((ax >> p) & 1) oa.push_back(p);
            if ((bx >> p) & 1) ob.push_back(p);
        }
        // try every split of the alpha (beta) orbitals into annihilated / created halves
        bool found = false;
        const int na = ka / 2, nb = kb / 2;
        for (u32 ma = 0; ma < (1u << ka) && !found; ++ma) {
            if (__builtin_popcount(ma) != na) continue;
            for (u32 mb = 0; mb < (1u << kb) && !found; ++mb) {
                if (__builtin_popcount(mb) != nb) continue;
                u32 anna = 0, annb = 0;
                for (int i = 0; i < ka; ++i)
                    if ((ma >> i) & 1) anna |= 1u << oa[i];
                for (int i = 0; i < kb; ++i)
                    if ((mb >> i) & 1) annb |= 1u << ob[i];
                const double f = N.element(ax, bx, anna, annb);
                if (std::fabs(std::fabs(f) - 1.0) > 1e-9) continue;
                // T = a+_{cre a..} a+_{cre b..} a_{ann a..} a_{ann b..} (jl/ansatz.jl order)
                std::vector<int> term;
                for (int p = 0; p < n; ++p)
                    if (((ax & ~anna) >> p) & 1) term.push_back(2 * p);
                for (int p = 0; p < n; ++p)
                    if (((bx & ~annb) >> p) & 1) term.push_back(2 * p + 1);
                for (int p = 0; p < n; ++p)
                    if ((anna >> p) & 1) term.push_back(2 * p);
                for (int p = 0; p < n; ++p)
                    if ((annb >> p) & 1) term.push_back(2 * p + 1);
                // sign of T on the probe state: apply ladder ops right to left
                u32 a = anna, b = annb;
                int par = 0;
                for (int i = (int)term.size() - 1; i >= 0; --i) {
                    const int pos = term[i], q = pos >> 1;
                    if (pos & 1) {
                        par += __builtin_popcount(a & ((2u << q) - 1)) + __builtin_popcount(b & ((1u << q) - 1));
                        b ^= 1u << q;
                    } else {
                        par += __builtin_popcount(a & ((1u << q) - 1)) + __builtin_popcount(b & ((1u << q) - 1));
                        a ^= 1u << q;
                    }
                }
                const double t = (par & 1) ? -1.0 : 1.0;
                if (f * t < 0) {
                    // group equals -(T - T^dagger) = (T' - T'^dagger) with T' = T^dagger
                    std::vector<int> rev;
                    const int h = (int)term.size() / 2;
                    for (int i = h; i < (int)term.size(); ++i) rev.push_back(term[i]);
                    for (int i = 0; i < h; ++i) rev.push_back(term[i]);
                    // T^dagger = a+_{ann..}(reversed) a_{cre..}(reversed): reverse each half
                    std::reverse(rev.begin(), rev.begin() + h);
                    std::reverse(rev.begin() + h, rev.end());
                    term = rev;
                }
                for (int v : term) D.orbs.push_back(v);
                D.offsets.push_back((int32_t)D.orbs.size());
                found = true;
            }
        }
        HIPOTF_REQUIRE(found, "pool group " + std::to_string(g) +
                                  " is not a fermionic excitation generator (generic networks unsupported)");
    }
}

// <T-decoded generator> == network group on probe determinants: every split of the group's
// flip pattern, with no, all and pseudo-random spectators.  G = T - T^dagger, T the ladder
// product of offsets[g]..offsets[g+1] (creators first), so <x^f|G|x> = sign(T x) when T maps x
// to x^f, -sign(T x^f) when T maps x^f to x, 0 otherwise.
static void verify_pool(const HostNet &N, const DevNet &D) {
    const int n = N.norb;
    const u32 full = n >= 32 ? 0xFFFFFFFFu : ((1u << n) - 1);
    auto apply_term = [&](const std::vector<std::pair<int, bool>> &ops, u64 x, u64 &out) -> int {
        int par = 0;
        for (int i = (int)ops.size() - 1; i >= 0; --i)
            if (!ladder_apply(x, par, ops[i].first, ops[i].second)) return 0;
        out = x;
        return (par & 1) ? -1 : 1;
    };
    for (size_t g = 0; g < N.groups.size(); ++g) {
        const HostGroup &G = N.groups[g];
        const int h = (D.offsets[g + 1] - D.offsets[g]) / 2;
        std::vector<std::pair<int, bool>> T, Td;
        for (int i = D.offsets[g]; i < D.offsets[g + 1]; ++i) T.emplace_back(D.orbs[i], i - D.offsets[g] < h);
        for (int i = (int)T.size() - 1; i >= 0; --i) Td.emplace_back(T[i].first, !T[i].second);
        const u32 fa = full & ~G.ax, fb = full & ~G.bx;
        for (u32 anna = G.ax;; anna = (anna - 1) & G.ax) {
            for (u32 annb = G.bx;; annb = (annb - 1) & G.bx) {
                for (int r = 0; r < 6; ++r) {
                    const u64 key = (u64)g * 0x9E3779B97F4A7C15ull + (u64)anna * 977 + annb * 131 + r;
                    const u32 sa = r == 0 ? 0u : r == 1 ? fa : probe_subset(fa, key);
                    const u32 sb = r == 0 ? 0u : r == 1 ? fb : probe_subset(fb, key ^ 0x5Aull);
                    const u32 a = anna | sa, b = annb | sb;
                    const u64 ket = interleave_occ(a, b), bra = interleave_occ(a ^ G.ax, b ^ G.bx);
                    u64 o = 0;
                    double e = 0.0;
                    int s = apply_term(T, ket, o);
                    if (s && o == bra) e += s;
                    s = apply_term(Td, ket, o);
                    if (s && o == bra) e -= s;
                    const double f = N.element(G.ax, G.bx, a, b);
                    if (!(std::fabs(f - e) <= 1e-9)) {
                        char msg[256];
                        std::snprintf(msg, sizeof msg,
                                      "pool group %zu is not a fermionic excitation generator T - T^dagger: element "
                                      "at a=%#x b=%#x is %.12g, expected %.12g (generic networks unsupported)",
                                      g, a, b, f, e);
                        throw Error(msg);
                    }
                }
                if (annb == 0) break;
            }
            if (anna == 0) break;
        }
    }
}

}  // namespace hipotf

using namespace hipotf;

#define API_TRY try {
#define API_CATCH(ret)                                        \
    }                                                         \
    catch (const std::exception &e) {                         \
        set_last_error(e.what());                             \
        return ret;                                           \
    }                                                         \
    catch (...) {                                             \
        set_last_error("unknown error");                      \
        return ret;                                           \
    }
#define API_CATCH_VOID                                        \
    }                                                         \
    catch (const std::exception &e) {                         \
        set_last_error(e.what());                             \
    }                                                         \
    catch (...) {                                             \
        set_last_error("unknown error");                      \
    }

static const double kNaN = std::numeric_limits<double>::quiet_NaN();

// index ia * Nb + ib of determinant |astr, bstr> in the basis, or -1 if it is not there
// (basis.hpp:298-330: strings outside the space are ignored)
static i64 det_index(const HostBasis *hb, uint32_t astr, uint32_t bstr) {
    if (hb->dim == 0) return -1;
    auto ia = std::lower_bound(hb->A.strs.begin(), hb->A.strs.end(), (u64)astr);
    auto ib = std::lower_bound(hb->B.strs.begin(), hb->B.strs.end(), (u64)bstr);
    if (ia == hb->A.strs.end() || *ia != astr || ib == hb->B.strs.end() || *ib != bstr) return -1;
    return (i64)(ia - hb->A.strs.begin()) * hb->B.N + (i64)(ib - hb->B.strs.begin());
}

// complex float64 as passed by value across the boundary (two doubles: real, imaginary)
struct HipotfC64 {
    double re, im;
};

extern "C" {

const char *hipotf_last_error(void) { return last_error_cstr(); }
void hipotf_clear_error(void) { clear_last_error(); }

// ---- host basis (libbasis) -----------------------------------------------------------------
void *create_basis_manager(int64_t norb, int64_t na, int64_t nb, int64_t total_sym,
                           const int64_t *orbsym, int64_t num_irreps) {
    API_TRY
    return make_host_basis(norb, na, nb, total_sym, orbsym, num_irreps);
    API_CATCH(nullptr)
}

int64_t get_subspace_dim(void *basis) {
    API_TRY
    HIPOTF_REQUIRE(basis, "basis handle is NULL");
    return static_cast<HostBasis *>(basis)->dim;
    API_CATCH(-1)
}

int64_t get_num_symmetry_blocks(void *basis) {
    API_TRY
    HIPOTF_REQUIRE(basis, "basis handle is NULL");
    return static_cast<HostBasis *>(basis)->num_blocks;
    API_CATCH(-1)
}

void destroy_basis_manager(void *basis) { delete static_cast<HostBasis *>(basis); }

void set_det_coeff_f64(void *basis, uint32_t astr, uint32_t bstr, double coeff, double *vec) {
    API_TRY
    HIPOTF_REQUIRE(basis && vec, "NULL argument");
    const i64 at = det_index(static_cast<HostBasis *>(basis), astr, bstr);
    if (at >= 0) vec[at] += coeff;
    API_CATCH_VOID
}

// complex vector: (re, im) pairs, the layout of the reference's complexf64
void set_det_coeff_c64(void *basis, uint32_t astr, uint32_t bstr, HipotfC64 coeff, HipotfC64 *vec) {
    API_TRY
    HIPOTF_REQUIRE(basis && vec, "NULL argument");
    const i64 at = det_index(static_cast<HostBasis *>(basis), astr, bstr);
    if (at >= 0) {
        vec[at].re += coeff.re;
        vec[at].im += coeff.im;
    }
    API_CATCH_VOID
}

// ---- host network (libotf_native build/destroy) ----------------------------------------------
void *build_network_otf_f64(const int64_t *orbsym, int64_t norb, int64_t ngs, const uint32_t *axs,
                            const uint32_t *bxs, const int64_t *ranks, const int64_t *num_zas,
                            const int64_t *num_zbs, const uint32_t *flat_zas, const uint32_t *flat_zbs,
                            const double *flat_wa, const double *flat_wb) {
    API_TRY
    (void)orbsym;
    HIPOTF_REQUIRE(norb >= 1 && norb <= 32 && ngs >= 0, "bad network sizes");
    auto N = std::make_unique<HostNet>();
    N->norb = (int)norb;
    size_t za = 0, zb = 0, wa = 0, wb = 0;
    for (int64_t g = 0; g < ngs; ++g) {
        HostGroup G;
        G.ax = axs[g];
        G.bx = bxs[g];
        G.rank = (int)ranks[g];
        G.nza = (int)num_zas[g];
        G.nzb = (int)num_zbs[g];
        HIPOTF_REQUIRE(G.rank >= 0 && G.nza >= 0 && G.nzb >= 0, "bad group sizes");
        G.zas.assign(flat_zas + za, flat_zas + za + G.nza);
        G.zbs.assign(flat_zbs + zb, flat_zbs + zb + G.nzb);
        G.wa.assign(flat_wa + wa, flat_wa + wa + (size_t)G.nza * G.rank);
        G.wb.assign(flat_wb + wb, flat_wb + wb + (size_t)G.nzb * G.rank);
        za += G.nza;
        zb += G.nzb;
        wa += (size_t)G.nza * G.rank;
        wb += (size_t)G.nzb * G.rank;
        N->by_x[((u64)G.ax << 32) | G.bx] = (int)N->groups.size();
        N->groups.push_back(std::move(G));
    }
    return N.release();
    API_CATCH(nullptr)
}

void destroy_network_otf_f64(void *net) { delete static_cast<HostNet *>(net); }

// ---- device handles (libcuotf) ---------------------------------------------------------------
void sync_device_cuda(void) {
    API_TRY
    HIPOTF_CHECK(hipDeviceSynchronize());
    API_CATCH_VOID
}

void *build_basisdev_f64(void *host_basis) {
    API_TRY
    HIPOTF_REQUIRE(host_basis, "host basis handle is NULL");
    return make_dev_basis(*static_cast<HostBasis *>(host_basis));
    API_CATCH(nullptr)
}

void destroy_basisdev_f64(void *basis) { delete static_cast<DevBasis *>(basis); }

void *build_networkdev_f64(void *host_net) {
    API_TRY
    HIPOTF_REQUIRE(host_net, "host network handle is NULL");
    const HostNet &N = *static_cast<HostNet *>(host_net);
    auto D = std::make_unique<DevNet>();
    bool has_diag = N.by_x.count(0) > 0;
    if (has_diag) {
        // a spin-free two-body Hamiltonian decodes into integrals (the fast path); any other
        // network with a diagonal group is kept as its Pauli groups and evaluated as such
        try {
            decode_hamiltonian(N, *D);
            verify_hamiltonian(N, *D);
        } catch (const Error &) {
            D = std::make_unique<DevNet>();
            D->kind = 3;
            D->n = N.norb;
            std::vector<u32> ax, bx;
            std::vector<int> rank;
            std::vector<std::vector<u32>> zas, zbs;
            std::vector<std::vector<double>> was, wbs;
            for (const HostGroup &G : N.groups) {
                ax.push_back(G.ax);
                bx.push_back(G.bx);
                rank.push_back(G.rank);
                zas.push_back(G.zas);
                zbs.push_back(G.zbs);
                was.push_back(G.wa);
                wbs.push_back(G.wb);
            }
            D->gen = make_dev_gennet(N.norb, ax, bx, rank, zas, zbs, was, wbs);
        }
    } else {
        decode_pool(N, *D);
        verify_pool(N, *D);
    }
    return D.release();
    API_CATCH(nullptr)
}

void destroy_networkdev_f64(void *net) { delete static_cast<DevNet *>(net); }

int hipotf_network_kind(void *net) {
    API_TRY
    HIPOTF_REQUIRE(net, "operator handle is NULL");
    return as_net(net)->kind;
    API_CATCH(0)
}

void *hipotf_build_ham_f64(int64_t norb, const double *one_body_mo, const double *two_body_mo,
                           double energy_nuc) {
    API_TRY
    auto D = std::make_unique<DevNet>();
    D->kind = 1;
    D->n = (int)norb;
    D->ecore = energy_nuc;
    integrals_from_reference_layout((int)norb, one_body_mo, two_body_mo, D->h, D->eri);
    return D.release();
    API_CATCH(nullptr)
}

void *hipotf_build_pool_f64(int64_t norb, int64_t nops, const int32_t *offsets, const int32_t *orbs) {
    API_TRY
    HIPOTF_REQUIRE(norb >= 1 && norb <= 32 && nops >= 0, "bad pool sizes");
    validate_pool((int)norb, nops, offsets, orbs);
    auto D = std::make_unique<DevNet>();
    D->kind = 2;
    D->n = (int)norb;
    D->offsets.assign(offsets, offsets + nops + 1);
    D->orbs.assign(orbs, orbs + (nops ? offsets[nops] : 0));
    return D.release();
    API_CATCH(nullptr)
}

int hipotf_prepare(void *basis, void *net) {
    API_TRY
    DevBasis *B = as_basis(basis);
    DevNet *N = as_net(net);
    if (N->kind == 1) N->ham_for(B);
    else if (N->kind == 2) N->pool_for(B);
    HIPOTF_CHECK(hipDeviceSynchronize());
    return 0;
    API_CATCH(1)
}

// launch choices of the sigma kernels (process-wide, per precision 0 FP64 / 1 FP32):
// v[0..n) = {ab_kernel, ab_split, ab_db, ab_nvs, rs_passes, deterministic, rs_block,
// sweep_mode}; values out of range are rejected
int hipotf_set_tuning(int precision, const int32_t *v, int n) {
    API_TRY
    HIPOTF_REQUIRE(precision == 0 || precision == 1, "tuning: precision must be 0 or 1");
    HIPOTF_REQUIRE(v != nullptr && n == TUNING_FIELDS, "tuning: expected 8 values");
    HIPOTF_REQUIRE(v[0] >= 0 && v[0] <= 5, "tuning: ab_kernel must be 0..5");
    HIPOTF_REQUIRE(v[1] == 1 || v[1] == 2, "tuning: ab_split must be 1 or 2");
    HIPOTF_REQUIRE(v[2] == 1 || v[2] == 2, "tuning: ab_db must be 1 or 2");
    HIPOTF_REQUIRE(v[3] == 1 || v[3] == 2, "tuning: ab_nvs must be 1 or 2");
    HIPOTF_REQUIRE(v[4] == 0 || v[4] == 1 || v[4] == 2 || v[4] == 4 || v[4] == 8,
                   "tuning: rs_passes must be 0, 1, 2, 4 or 8");
    HIPOTF_REQUIRE(v[5] == 0 || v[5] == 1, "tuning: deterministic must be 0 or 1");
    HIPOTF_REQUIRE(v[6] == 0 || v[6] == 4 || v[6] == 8 || v[6] == 16, "tuning: rs_block must be 0, 4, 8 or 16");
    HIPOTF_REQUIRE(v[7] >= 0 && v[7] <= 7, "tuning: sweep_mode must be 0..7");
    SigmaTuning &t = tuning(precision);
    t.ab_kernel = v[0];
    t.ab_split = v[1];
    t.ab_db = v[2];
    t.ab_nvs = v[3];
    t.rs_passes = v[4];
    t.deterministic = v[5];
    t.rs_block = v[6];
    t.sweep_mode = v[7];
    return 0;
    API_CATCH(1)
}

int hipotf_get_tuning(int precision, int32_t *v, int n) {
    API_TRY
    HIPOTF_REQUIRE(precision == 0 || precision == 1, "tuning: precision must be 0 or 1");
    HIPOTF_REQUIRE(v != nullptr && n == TUNING_FIELDS, "tuning: expected 8 values");
    const SigmaTuning &t = tuning(precision);
    v[0] = t.ab_kernel;
    v[1] = t.ab_split;
    v[2] = t.ab_db;
    v[3] = t.ab_nvs;
    v[4] = t.rs_passes;
    v[5] = t.deterministic;
    v[6] = t.rs_block;
    v[7] = t.sweep_mode;
    return 0;
    API_CATCH(1)
}

// alpha-beta section timing: enable 1 zeroes the totals and starts, 0 stops, -1 leaves the state;
// returns (through non-null pointers) the totals since the last start
int hipotf_ab_timing(int enable, double *ms, int64_t *sections, int64_t *mma_launches) {
    API_TRY
    HIPOTF_REQUIRE(enable >= -1 && enable <= 1, "ab timing: enable must be -1, 0 or 1");
    AbTiming &t = ab_timing();
    if (enable == 1) {
        if (!t.e0) HIPOTF_CHECK(hipEventCreate(&t.e0));
        if (!t.e1) HIPOTF_CHECK(hipEventCreate(&t.e1));
        t.ms = 0.0;
        t.sections = t.mma_launches = 0;
        t.on = true;
    } else if (enable == 0) {
        t.on = false;
    }
    if (ms) *ms = t.ms;
    if (sections) *sections = t.sections;
    if (mma_launches) *mma_launches = t.mma_launches;
    return 0;
    API_CATCH(1)
}

int64_t hipotf_num_generators(void *net) {
    API_TRY
    DevNet *N = as_net(net);
    return N->kind == 2 ? (int64_t)N->offsets.size() - 1 : 0;
    API_CATCH(-1)
}

void hvec_cuda(void *basis, void *net, const double *src, double *dst) {
    API_TRY
    DevBasis *B = as_basis(basis);
    if (as_net(net)->kind == 3) {
        gennet_hvec(as_net(net)->gen, B, src, dst, S0());
        return;
    }
    DevHam *H = as_net(net)->ham_for(B);
    sigma(H, 1, &src, &dst, 0, S0());
    API_CATCH_VOID
}

// allocation-free hvec_cuda (for callers that cannot allocate device memory); the limit and
// the tables it needs are checked before any work is issued
int hipotf_hvec_noalloc(void *basis, void *net, const double *src, double *dst) {
    API_TRY
    DevBasis *B = as_basis(basis);
    HIPOTF_REQUIRE(B->host.dim <= NOALLOC_MAX_DIM,
                   "hipotf_hvec_noalloc: " + std::to_string(B->host.dim) + " determinants exceed the limit of " +
                       std::to_string(NOALLOC_MAX_DIM));
    DevHam *H = as_net(net)->ham_prepared(B);
    sigma_noalloc(H, src, dst, S0());
    return 0;
    API_CATCH(1)
}

int hipotf_hvec_block(void *basis, void *net, int nvec, const double *const *src, double *const *dst,
                      int precision) {
    API_TRY
    HIPOTF_REQUIRE(precision == 0 || precision == 1, "precision must be 0 (fp64) or 1 (fp32)");
    DevBasis *B = as_basis(basis);
    DevHam *H = as_net(net)->ham_for(B);
    sigma(H, nvec, src, dst, precision, S0());
    return 0;
    API_CATCH(1)
}

void get_diags_elements_cuda(void *basis, void *net, double *diags) {
    API_TRY
    DevBasis *B = as_basis(basis);
    if (as_net(net)->kind == 3) {
        gennet_diag(as_net(net)->gen, B, diags, S0());
        return;
    }
    diag_elements(as_net(net)->ham_for(B), diags, S0());
    API_CATCH_VOID
}

void expm_cuda(void *basis, void *net, int64_t idx, double theta, double *vec) {
    API_TRY
    DevBasis *B = as_basis(basis);
    rot_expm(as_net(net)->pool_for(B), idx, theta, vec, S0());
    API_CATCH_VOID
}

double grad_cuda(void *basis, void *net, int64_t idx, double theta, const double *lp, const double *rp) {
    API_TRY
    DevBasis *B = as_basis(basis);
    DevPool *P = as_net(net)->pool_for(B);
    double *d = P->scal.as<double>();
    rot_grad(P, idx, theta, lp, rp, d, S0());
    double h = 0.0;
    HIPOTF_CHECK(hipMemcpy(&h, d, sizeof(double), hipMemcpyDeviceToHost));
    return h;
    API_CATCH(kNaN)
}

double backgrad_cuda(void *basis, void *net, int64_t idx, double theta, double *lp, double *rp) {
    API_TRY
    DevBasis *B = as_basis(basis);
    DevPool *P = as_net(net)->pool_for(B);
    double *d = P->scal.as<double>();
    rot_backgrad(P, idx, theta, lp, rp, d, S0());
    double h = 0.0;
    HIPOTF_CHECK(hipMemcpy(&h, d, sizeof(double), hipMemcpyDeviceToHost));
    return h;
    API_CATCH(kNaN)
}

void expm_cuda_2d(void *basis, void *net, int64_t idx, double theta, double *vec) {
    expm_cuda(basis, net, idx, theta, vec);
}
double grad_cuda_2d(void *basis, void *net, int64_t idx, double theta, const double *lp, const double *rp) {
    return grad_cuda(basis, net, idx, theta, lp, rp);
}
double backgrad_cuda_2d(void *basis, void *net, int64_t idx, double theta, double *lp, double *rp) {
    return backgrad_cuda(basis, net, idx, theta, lp, rp);
}

// ---- real-time / batched pool operators (libotf_native only in the reference:
// src/include/op/declare.hpp:76-130; the same arguments with device pointers) ------------------
void tvec_cuda(void *basis, void *net, int64_t idx, const double *src, double *dst) {
    API_TRY
    DevBasis *B = as_basis(basis);
    rot_tvec(as_net(net)->pool_for(B), idx, src, dst, S0());
    API_CATCH_VOID
}

void backtran_cuda(void *basis, void *net, int64_t idx, double theta, double *lp, double *rp, double *bp) {
    API_TRY
    DevBasis *B = as_basis(basis);
    rot_backtran(as_net(net)->pool_for(B), idx, theta, lp, rp, bp, S0());
    API_CATCH_VOID
}

void batchexpm_cuda(void *basis, void *net, int64_t idx, double theta, double *matrix, int64_t ld,
                    int64_t nvec) {
    API_TRY
    DevBasis *B = as_basis(basis);
    rot_batch_expm(as_net(net)->pool_for(B), idx, theta, matrix, ld, nvec, S0());
    API_CATCH_VOID
}

void batchtran_cuda(void *basis, void *net, const double *lp, const double *rp, double *trans) {
    API_TRY
    DevBasis *B = as_basis(basis);
    DevPool *P = as_net(net)->pool_for(B);
    HIPOTF_REQUIRE(lp && rp && trans, "NULL argument");
    batch_tran(P, lp, rp, trans, S0());
    API_CATCH_VOID
}

// grads[g] = grad(g, thetas[g], lp, rp) for every generator (src/include/cuda/batchgrad.cuh:106-378);
// thetas / grads are device arrays.  One gradient-only pass over the sweeps' fused segments:
// the generators sharing an alpha pattern (or a row kind) read their rows once per segment
void batchgrad_cuda(void *basis, void *net, const double *thetas, const double *lp, const double *rp,
                    double *grads) {
    API_TRY
    DevBasis *B = as_basis(basis);
    DevPool *P = as_net(net)->pool_for(B);
    HIPOTF_REQUIRE(thetas && grads && lp && rp, "NULL argument");
    fused_grads(P, thetas, lp, rp, grads, S0());
    API_CATCH_VOID
}

// ---- fused VQE objective (jl/vqe.jl:1-21) ----------------------------------------------------
// psi: device, in = reference state, out = reference state again (uncomputed by the reverse
// sweep); lam: device workspace (out = H psi rotated back); grads: host [nparams];
// out3: host {energy, dH2, |psi|^2}.
int hipotf_vqe_energy_grad(void *basis, void *ham, void *pool, int64_t nparams, const int64_t *idxs,
                           const double *thetas, double *psi, double *lam, double *grads, double *out3) {
    API_TRY
    DevBasis *B = as_basis(basis);
    DevHam *H = as_net(ham)->ham_for(B);
    DevPool *P = as_net(pool)->pool_for(B);
    HIPOTF_REQUIRE(psi && lam && grads && out3 && (nparams == 0 || (idxs && thetas)), "NULL argument");
    const i64 dim = B->host.dim;
    hipStream_t st = S0();
    fused_sweeps(P, nparams, idxs, thetas, psi, lam, true, false, nullptr, st);
    const double *src = psi;
    double *dst = lam;
    sigma(H, 1, &src, &dst, 0, st);
    P->vqe_res.ensure((size_t)(nparams + 4) * sizeof(double));  // kept across calls
    double *d = P->vqe_res.as<double>();
    dot_dev(dim, psi, psi, d + nparams, st);
    dot_dev(dim, lam, lam, d + nparams + 1, st);
    dot_dev(dim, psi, lam, d + nparams + 2, st);
    fused_sweeps(P, nparams, idxs, thetas, psi, lam, false, true, d, st);
    std::vector<double> h(nparams + 3);
    HIPOTF_CHECK(hipMemcpyAsync(h.data(), d, (nparams + 3) * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPOTF_CHECK(hipStreamSynchronize(st));
    const double lnorm = h[nparams], rnorm = h[nparams + 1], e = h[nparams + 2] / lnorm;
    for (i64 i = 0; i < nparams; ++i) grads[i] = h[i] * 2.0 / lnorm;
    out3[0] = e;
    out3[1] = std::max(0.0, rnorm / lnorm - e * e);
    out3[2] = lnorm;
    return 0;
    API_CATCH(1)
}

// energy only (the line-search / finite-difference evaluations of jl/vqe.jl): forward sweep,
// lam = H psi, energy and variance; psi is left as the ansatz state U(theta)|ref>
int hipotf_vqe_energy(void *basis, void *ham, void *pool, int64_t nparams, const int64_t *idxs,
                      const double *thetas, double *psi, double *lam, double *out3) {
    API_TRY
    DevBasis *B = as_basis(basis);
    DevPool *P = as_net(pool)->pool_for(B);
    HIPOTF_REQUIRE(psi && lam && (nparams == 0 || (idxs && thetas)), "NULL argument");
    const i64 dim = B->host.dim;
    hipStream_t st = S0();
    fused_sweeps(P, nparams, idxs, thetas, psi, lam, true, false, nullptr, st);
    if (!ham) {  // forward sweep only
        HIPOTF_CHECK(hipStreamSynchronize(st));
        return 0;
    }
    HIPOTF_REQUIRE(out3, "NULL argument");
    DevHam *H = as_net(ham)->ham_for(B);
    const double *src = psi;
    double *dst = lam;
    sigma(H, 1, &src, &dst, 0, st);
    P->vqe_res.ensure(4 * sizeof(double));
    double *d = P->vqe_res.as<double>();
    dot_dev(dim, psi, psi, d, st);
    dot_dev(dim, lam, lam, d + 1, st);
    dot_dev(dim, psi, lam, d + 2, st);
    double h[3];
    HIPOTF_CHECK(hipMemcpyAsync(h, d, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPOTF_CHECK(hipStreamSynchronize(st));
    const double e = h[2] / h[0];
    out3[0] = e;
    out3[1] = std::max(0.0, h[1] / h[0] - e * e);
    out3[2] = h[0];
    return 0;
    API_CATCH(1)
}

// ---- Davidson driver (jl/davidson.jl outer loop on the device; see davidson.cpp) -----------
// info (host, 7 doubles): iterations, fp64 sigma vectors, fp32 sigma vectors, max residual,
// wall seconds, iteration of the fp32 -> fp64 switch (-1 if none), converged flag
int hipotf_davidson_f64(void *basis, void *ham, int nroots, int mode, int block, double tol,
                        double switch_tol, int maxiter, int maxspace, int verbose, const double *guess,
                        double *energies, double *vecs, double *info) {
    API_TRY
    HIPOTF_REQUIRE(energies && info, "NULL output");
    HIPOTF_REQUIRE(mode == 0 || mode == 1, "mode must be 0 (fp64) or 1 (mixed block)");
    DevBasis *B = as_basis(basis);
    DevHam *H = as_net(ham)->ham_for(B);
    DavidsonOut o = davidson(H, nroots, mode, block, tol, switch_tol, maxiter, maxspace, verbose, guess,
                             energies, vecs);
    info[0] = o.iterations;
    info[1] = (double)o.sigma64;
    info[2] = (double)o.sigma32;
    info[3] = o.max_residual;
    info[4] = o.seconds;
    info[5] = o.refine_iter;
    info[6] = o.converged;
    return o.converged == 1 ? 0 : o.converged == 2 ? 3 : 2;
    API_CATCH(1)
}

// ---- libdiag on the device (src/include/diag/declare.hpp:4-75) -------------------------------
double build_ritz_and_residual_f64(int64_t N, int dim, const double *coeffs, double min_e,
                                   const double **V, const double **AV, double *axt) {
    API_TRY
    return ritz_and_residual(N, dim, coeffs, min_e, V, AV, axt, S0());
    API_CATCH(kNaN)
}

void build_ritz_vector_f64(int64_t N, int dim, const double *coeffs, const double **V, double *xt) {
    API_TRY
    ritz_vector(N, dim, coeffs, V, xt, S0());
    API_CATCH_VOID
}

void apply_preconditioner_inplace_f64(int64_t N, double *axt, const double *diags, double min_e,
                                      double shift) {
    API_TRY
    precondition(N, axt, diags, min_e, shift, S0());
    API_CATCH_VOID
}

double mgs_orthogonalize_f64(int64_t N, int dim, const double **V, double *vt) {
    API_TRY
    return mgs(N, dim, V, vt, S0());
    API_CATCH(kNaN)
}

double inplace_normalize_f64(int64_t N, double *vec) {
    API_TRY
    return normalize(N, vec, S0());
    API_CATCH(kNaN)
}

double fast_dot_f64(int64_t N, const double *v1, const double *v2) {
    API_TRY
    return dot(N, v1, v2, S0());
    API_CATCH(kNaN)
}

void compute_heff_col_f64(int64_t N, int dim, const double **V, const double *AV_dim, double *out_col) {
    API_TRY
    multidot(N, dim, V, AV_dim, out_col, S0());
    API_CATCH_VOID
}

void rebuild_heff_f64(int64_t N, int dim, const double **V, const double **AV, double *heff, int ldh) {
    API_TRY
    std::vector<double> col(dim);
    for (int c = 0; c < dim; ++c) {
        multidot(N, dim, V, AV[c], col.data(), S0());
        for (int r = 0; r < dim; ++r) heff[r + (size_t)c * ldh] = col[r];
    }
    API_CATCH_VOID
}

}  // extern "C"
