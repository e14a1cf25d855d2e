// Givens-rotation operators of a UCC generator pool: expm / grad / backgrad.
//
// Replaces expm_svd_network_otf_gpu, grad_svd_network_otf_gpu and
// backgrad_svd_network_otf_gpu (src/include/cuda/expm.cuh:141-197, grad.cuh:144-203,
// backgrad.cuh:152-212) and their CPU counterparts (src/include/op/expm.hpp:374-438,
// grad.hpp:397-465, backgrad.hpp:406-474).
//
// A generator G = T - T^dagger with T = a+_{c1} .. a+_{cm} a_{a1} .. a_{am} (interleaved
// spin-orbitals, jl/ansatz.jl:230-250) couples each determinant u that T maps to a nonzero
// state with w = T u / t(u), t(u) = +-1 its Jordan-Wigner sign.  Instead of scanning every
// determinant and evaluating the SVD phase network, the kernels walk exactly those pairs:
// per spin a pair list (u index, w index) of the strings matching the generator's pattern
// (identity list for a spin the generator does not touch), and t(u) is computed on the fly
// by applying the ladder operators to (a, b) with popcounts.  Per pair, with c = cos(theta),
// s = sin(theta) (update formulas of src/include/core/math.hpp:132-215, |vt| = 1):
//   expm     : x_u <- c x_u - t s x_w ,  x_w <- t s x_u + c x_w
//   grad     : g += l_u(-s r_u + t c r_w) + l_w(-s r_w - t c r_u)
//   backgrad : l <- R(-theta) l ; g += [same form with new l, old r] ; r <- R(-theta) r
#include <numeric>
#include "hipotf.hpp"

#include <algorithm>
#include <cmath>

namespace hipotf {

i64 host_find_str(const SpinSpace &S, u64 s);

struct GenDev {
    int nops;
    int pos[8];
    int cre[8];
};

__device__ __forceinline__ int jw_sign(const GenDev &g, u64 a, u64 b) {
    int par = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (i >= g.nops) break;
        const int p = g.pos[i];
        const int q = p >> 1;
        const u64 lowq = (1ull << q) - 1;
        if ((p & 1) == 0) {
            par += __popcll(a & lowq) + __popcll(b & lowq);
            a ^= 1ull << q;
        } else {
            par += __popcll(a & ((lowq << 1) | 1ull)) + __popcll(b & lowq);
            b ^= 1ull << q;
        }
    }
    return par & 1;
}

struct PairView {
    const int32_t *la, *lb;   // (u, w) index pairs
    i64 na, nb;               // list lengths
    i64 NB;                   // beta strings (row length)
    const u64 *strsA, *strsB;
};

__device__ __forceinline__ void pair_at(const PairView &v, i64 i, i64 &u, i64 &w, u64 &ua, u64 &ub) {
    const i64 ia = i / v.nb, ib = i - ia * v.nb;
    const int32_t uA = v.la[2 * ia], wA = v.la[2 * ia + 1];
    const int32_t uB = v.lb[2 * ib], wB = v.lb[2 * ib + 1];
    u = (i64)uA * v.NB + uB;
    w = (i64)wA * v.NB + wB;
    ua = v.strsA[uA];
    ub = v.strsB[uB];
}

__global__ __launch_bounds__(256) void k_expm(PairView v, GenDev g, double c, double s,
                                              double *__restrict__ x) {
    const i64 tot = v.na * v.nb;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (i64)gridDim.x * blockDim.x) {
        i64 u, w;
        u64 ua, ub;
        pair_at(v, i, u, w, ua, ub);
        const double ts = jw_sign(g, ua, ub) ? -s : s;
        const double xu = x[u], xw = x[w];
        x[u] = c * xu - ts * xw;
        x[w] = ts * xu + c * xw;
    }
}

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double red[16];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) t += red[i];
    }
    return t;  // valid in thread 0
}

__global__ __launch_bounds__(256) void k_grad(PairView v, GenDev g, double c, double s,
                                              const double *__restrict__ l, const double *__restrict__ r,
                                              double *__restrict__ partial) {
    const i64 tot = v.na * v.nb;
    double acc = 0.0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (i64)gridDim.x * blockDim.x) {
        i64 u, w;
        u64 ua, ub;
        pair_at(v, i, u, w, ua, ub);
        const double tc = jw_sign(g, ua, ub) ? -c : c;
        const double ru = r[u], rw = r[w];
        acc += l[u] * (-s * ru + tc * rw) + l[w] * (-s * rw - tc * ru);
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_backgrad(PairView v, GenDev g, double c, double s,
                                                  double *__restrict__ l, double *__restrict__ r,
                                                  double *__restrict__ partial) {
    const i64 tot = v.na * v.nb;
    double acc = 0.0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (i64)gridDim.x * blockDim.x) {
        i64 u, w;
        u64 ua, ub;
        pair_at(v, i, u, w, ua, ub);
        const bool neg = jw_sign(g, ua, ub);
        const double ts = neg ? -s : s, tc = neg ? -c : c;
        // R(-theta): x_u <- c x_u + t s x_w ; x_w <- -t s x_u + c x_w
        const double lu = l[u], lw = l[w], ru = r[u], rw = r[w];
        const double nlu = c * lu + ts * lw, nlw = -ts * lu + c * lw;
        acc += nlu * (-s * ru + tc * rw) + nlw * (-s * rw - tc * ru);
        l[u] = nlu;
        l[w] = nlw;
        r[u] = c * ru + ts * rw;
        r[w] = -ts * ru + c * rw;
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// tvec (src/include/op/tvec.hpp:373-445, tvec_update of core/math.hpp:164-170): dst = G src,
// per pair dst_u = -t src_w, dst_w = t src_u (dst zeroed first by k_fill)
__global__ __launch_bounds__(256) void k_tvec(PairView v, GenDev g, const double *__restrict__ src,
                                              double *__restrict__ dst) {
    const i64 tot = v.na * v.nb;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (i64)gridDim.x * blockDim.x) {
        i64 u, w;
        u64 ua, ub;
        pair_at(v, i, u, w, ua, ub);
        const double t = jw_sign(g, ua, ub) ? -1.0 : 1.0;
        dst[u] = -t * src[w];
        dst[w] = t * src[u];
    }
}

// backtran (op/backtran.hpp:383-452, backtran_update of math.hpp:217-225): l and r rotated as
// expm rotates, then b = G l of the rotated l (b zeroed first)
__global__ __launch_bounds__(256) void k_backtran(PairView v, GenDev g, double c, double s,
                                                  double *__restrict__ l, double *__restrict__ r,
                                                  double *__restrict__ b) {
    const i64 tot = v.na * v.nb;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (i64)gridDim.x * blockDim.x) {
        i64 u, w;
        u64 ua, ub;
        pair_at(v, i, u, w, ua, ub);
        const bool neg = jw_sign(g, ua, ub);
        const double t = neg ? -1.0 : 1.0, ts = neg ? -s : s;
        const double lu = l[u], lw = l[w], ru = r[u], rw = r[w];
        const double nlu = c * lu - ts * lw, nlw = ts * lu + c * lw;
        l[u] = nlu;
        l[w] = nlw;
        r[u] = c * ru - ts * rw;
        r[w] = ts * ru + c * rw;
        b[u] = -t * nlw;
        b[w] = t * nlu;
    }
}

// batch expm (op/batchexpm.hpp:376-444, expm_batch_update_matrix of math.hpp:227-252): the
// rotation on columns 0..nvec-1 of a det-major matrix m[i * ld + k]; a pair's two rows are
// ld-strided, the columns of a row contiguous
__global__ __launch_bounds__(256) void k_batch_expm(PairView v, GenDev g, double c, double s,
                                                    double *__restrict__ m, i64 ld, int nvec) {
    const i64 tot = v.na * v.nb;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (i64)gridDim.x * blockDim.x) {
        i64 u, w;
        u64 ua, ub;
        pair_at(v, i, u, w, ua, ub);
        const double ts = jw_sign(g, ua, ub) ? -s : s;
        double *mu = m + u * ld, *mw = m + w * ld;
        for (int k = 0; k < nvec; ++k) {
            const double xu = mu[k], xw = mw[k];
            mu[k] = c * xu - ts * xw;
            mw[k] = ts * xu + c * xw;
        }
    }
}

__global__ __launch_bounds__(256) void k_fill(double *__restrict__ x, i64 n, double val) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        x[i] = val;
}

// sum of partial[0..n) into *out (+ optional scale) by one block, fixed order -> deterministic
__global__ __launch_bounds__(256) void k_sum_partials(const double *__restrict__ partial, int n,
                                                      double scale, double *__restrict__ out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
    const double t = block_sum(acc);
    if (threadIdx.x == 0) *out = t * scale;
}

// ---------------------------------------------------------------------------------------
// pool construction
// ---------------------------------------------------------------------------------------
static int pattern_id(DevPool &P, const SpinSpace &S, int spin, u64 cre, u64 ann) {
    if (cre == 0 && ann == 0) return spin;  // identity list of that spin
    for (size_t i = 2; i < P.lists.size(); ++i)
        if (P.lists[i].spin == spin && P.lists[i].cre == cre && P.lists[i].ann == ann) return (int)i;
    PairList L;
    L.cre = cre;
    L.ann = ann;
    L.spin = spin;
    std::vector<int32_t> uw;
    for (i64 i = 0; i < S.N; ++i) {
        const u64 s = S.strs[i];
        if ((s & ann) == ann && (s & cre) == 0) {
            const i64 j = host_find_str(S, s ^ ann ^ cre);
            HIPOTF_REQUIRE(j >= 0, "internal: excited string missing from the basis");
            uw.push_back((int32_t)i);
            uw.push_back((int32_t)j);
        }
    }
    L.len = (i64)(uw.size() / 2);
    L.uw.upload(uw);
    P.lists.push_back(std::move(L));
    return (int)P.lists.size() - 1;
}

// structural checks of a generator list (no basis needed): creators first, <= 8 ladder
// operators, no repeated spin-orbital, N_alpha and N_beta conserved, no number operators
void validate_pool(int n, i64 nops, const int32_t *offsets, const int32_t *orbs) {
    HIPOTF_REQUIRE(nops >= 0 && (nops == 0 || (offsets && orbs)), "bad pool arrays");
    HIPOTF_REQUIRE(nops == 0 || offsets[0] == 0, "offsets[0] must be 0");
    for (i64 g = 0; g < nops; ++g) {
        const int32_t o0 = offsets[g], o1 = offsets[g + 1];
        const int len = o1 - o0;
        const std::string tag = "generator " + std::to_string(g) + ": ";
        HIPOTF_REQUIRE(len >= 2 && len <= 8 && len % 2 == 0,
                       tag + "need 2..8 spin-orbitals, creators first");
        u64 m[2][2] = {{0, 0}, {0, 0}};  // [spin][cre]
        for (int i = 0; i < len; ++i) {
            const int pos = orbs[o0 + i];
            HIPOTF_REQUIRE(pos >= 0 && pos < 2 * n, tag + "spin-orbital out of range");
            const u64 bit = 1ull << (pos >> 1);
            u64 &mm = m[pos & 1][i < len / 2 ? 1 : 0];
            HIPOTF_REQUIRE((mm & bit) == 0, tag + "repeats a spin-orbital (T = 0)");
            mm |= bit;
        }
        for (int sp = 0; sp < 2; ++sp) {
            HIPOTF_REQUIRE((m[sp][0] & m[sp][1]) == 0,
                           tag + "number-operator (diagonal) generators are not supported");
            HIPOTF_REQUIRE(__builtin_popcountll(m[sp][0]) == __builtin_popcountll(m[sp][1]),
                           tag + "does not conserve N_alpha and N_beta");
        }
    }
}

DevPool *make_dev_pool(DevBasis *b, i64 nops, const int32_t *offsets, const int32_t *orbs) {
    HIPOTF_REQUIRE(b != nullptr, "basis handle is NULL");
    HIPOTF_REQUIRE(nops >= 0 && (nops == 0 || (offsets && orbs)), "bad pool arrays");
    auto P = std::make_unique<DevPool>();
    P->basis = b;
    const HostBasis &hb = b->host;
    const int n = (int)hb.norb;
    // identity lists
    for (int sp = 0; sp < 2; ++sp) {
        const SpinSpace &S = sp == 0 ? hb.A : hb.B;
        PairList L;
        L.identity = true;
        L.spin = sp;
        std::vector<int32_t> uw(2 * S.N);
        for (i64 i = 0; i < S.N; ++i) uw[2 * i] = uw[2 * i + 1] = (int32_t)i;
        L.len = S.N;
        L.uw.upload(uw);
        P->lists.push_back(std::move(L));
    }
    for (i64 g = 0; g < nops; ++g) {
        const int32_t o0 = offsets[g], o1 = offsets[g + 1];
        const int len = o1 - o0;
        HIPOTF_REQUIRE(len >= 2 && len <= 8 && len % 2 == 0,
                       "generator " + std::to_string(g) + ": need 2..8 spin-orbitals, creators first");
        Generator G;
        G.nops = len;
        u64 cre_a = 0, ann_a = 0, cre_b = 0, ann_b = 0;
        // T = a+_{l0} .. a+_{l(mid-1)} a_{l(mid)} .. a_{l(len-1)}; applied right to left
        for (int i = 0; i < len; ++i) {
            const int pos = orbs[o0 + i];
            HIPOTF_REQUIRE(pos >= 0 && pos < 2 * n, "generator spin-orbital out of range");
            const bool is_cre = i < len / 2;
            const u64 bit = 1ull << (pos >> 1);
            u64 &m = (pos & 1) ? (is_cre ? cre_b : ann_b) : (is_cre ? cre_a : ann_a);
            HIPOTF_REQUIRE((m & bit) == 0, "generator repeats a spin-orbital (T = 0)");
            m |= bit;
            G.pos[len - 1 - i] = pos;
            G.cre[len - 1 - i] = is_cre ? 1 : 0;
        }
        HIPOTF_REQUIRE((cre_a & ann_a) == 0 && (cre_b & ann_b) == 0,
                       "number-operator (diagonal) generators are not supported");
        HIPOTF_REQUIRE(__builtin_popcountll(cre_a) == __builtin_popcountll(ann_a) &&
                           __builtin_popcountll(cre_b) == __builtin_popcountll(ann_b),
                       "generator does not conserve N_alpha and N_beta");
        const bool ta = cre_a != 0, tb = cre_b != 0;
        G.kind = ta && tb ? 3 : (ta ? 1 : 2);
        G.pa = pattern_id(*P, hb.A, 0, cre_a, ann_a);
        G.pb = pattern_id(*P, hb.B, 1, cre_b, ann_b);
        P->gens.push_back(G);
    }
    P->partial.alloc(4096 * sizeof(double));
    P->scal.alloc(64 * sizeof(double));
    return P.release();
}

static void views(DevPool *P, i64 idx, PairView &v, GenDev &gd) {
    HIPOTF_REQUIRE(P != nullptr, "pool handle is NULL");
    HIPOTF_REQUIRE(idx >= 0 && idx < (i64)P->gens.size(),
                   "generator index " + std::to_string(idx) + " out of range");
    const Generator &G = P->gens[idx];
    DevBasis *B = P->basis;
    const PairList &LA = P->lists[G.pa], &LB = P->lists[G.pb];
    v.la = LA.uw.as<int32_t>();
    v.lb = LB.uw.as<int32_t>();
    v.na = LA.len;
    v.nb = LB.len;
    v.NB = B->pB->N;
    v.strsA = B->pA->strs.as<u64>();
    v.strsB = B->pB->strs.as<u64>();
    gd.nops = G.nops;
    for (int i = 0; i < 8; ++i) {
        gd.pos[i] = i < G.nops ? G.pos[i] : 0;
        gd.cre[i] = i < G.nops ? G.cre[i] : 0;
    }
}

static int grid_for(i64 tot) {
    const i64 want = (tot + 255) / 256;
    return (int)std::max<i64>(1, std::min<i64>(want, 2048));
}

void rot_expm(DevPool *P, i64 idx, double theta, double *vec, hipStream_t st) {
    PairView v;
    GenDev g;
    views(P, idx, v, g);
    const i64 tot = v.na * v.nb;
    if (tot == 0) return;
    HIPOTF_REQUIRE(vec != nullptr, "NULL vector");
    hipLaunchKernelGGL(k_expm, dim3(grid_for(tot)), dim3(256), 0, st, v, g, std::cos(theta),
                       std::sin(theta), vec);
    HIPOTF_CHECK(hipGetLastError());
}

static void fill(double *x, i64 n, double val, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, st, x, n, val);
    HIPOTF_CHECK(hipGetLastError());
}

void rot_tvec(DevPool *P, i64 idx, const double *src, double *dst, hipStream_t st) {
    PairView v;
    GenDev g;
    views(P, idx, v, g);
    HIPOTF_REQUIRE(src != nullptr && dst != nullptr, "NULL vector");
    HIPOTF_REQUIRE(src != dst, "tvec: src and dst must be different vectors");
    fill(dst, P->basis->host.dim, 0.0, st);
    const i64 tot = v.na * v.nb;
    if (tot == 0) return;
    hipLaunchKernelGGL(k_tvec, dim3(grid_for(tot)), dim3(256), 0, st, v, g, src, dst);
    HIPOTF_CHECK(hipGetLastError());
}

void rot_backtran(DevPool *P, i64 idx, double theta, double *lp, double *rp, double *bp, hipStream_t st) {
    PairView v;
    GenDev g;
    views(P, idx, v, g);
    HIPOTF_REQUIRE(lp && rp && bp, "NULL vector");
    HIPOTF_REQUIRE(bp != lp && bp != rp && lp != rp, "backtran: lp, rp and bp must be different vectors");
    fill(bp, P->basis->host.dim, 0.0, st);
    const i64 tot = v.na * v.nb;
    if (tot == 0) return;
    hipLaunchKernelGGL(k_backtran, dim3(grid_for(tot)), dim3(256), 0, st, v, g, std::cos(theta),
                       std::sin(theta), lp, rp, bp);
    HIPOTF_CHECK(hipGetLastError());
}

void rot_batch_expm(DevPool *P, i64 idx, double theta, double *m, i64 ld, i64 nvec, hipStream_t st) {
    PairView v;
    GenDev g;
    views(P, idx, v, g);
    HIPOTF_REQUIRE(m != nullptr, "NULL matrix");
    HIPOTF_REQUIRE(ld >= 1 && nvec >= 0 && nvec <= ld && nvec < (1ll << 31),
                   "batch expm: need 0 <= nvec <= ld");
    const i64 tot = v.na * v.nb;
    if (tot == 0 || nvec == 0) return;
    hipLaunchKernelGGL(k_batch_expm, dim3(grid_for(tot)), dim3(256), 0, st, v, g, std::cos(theta),
                       std::sin(theta), m, ld, (int)nvec);
    HIPOTF_CHECK(hipGetLastError());
}

void rot_grad(DevPool *P, i64 idx, double theta, const double *lp, const double *rp, double *dres,
              hipStream_t st) {
    PairView v;
    GenDev g;
    views(P, idx, v, g);
    const i64 tot = v.na * v.nb;
    HIPOTF_REQUIRE(lp && rp && dres, "NULL vector");
    const int nb = tot ? grid_for(tot) : 1;
    if (tot) {
        hipLaunchKernelGGL(k_grad, dim3(nb), dim3(256), 0, st, v, g, std::cos(theta), std::sin(theta),
                           lp, rp, P->partial.as<double>());
        HIPOTF_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, P->partial.as<double>(),
                       tot ? nb : 0, 1.0, dres);
    HIPOTF_CHECK(hipGetLastError());
}

// rot_grad with the angle on the device
static void rot_grad_dev(DevPool *P, i64 idx, const double *theta, const double *lp, const double *rp,
                         double *dres, hipStream_t st);

void rot_backgrad(DevPool *P, i64 idx, double theta, double *lp, double *rp, double *dres,
                  hipStream_t st) {
    PairView v;
    GenDev g;
    views(P, idx, v, g);
    const i64 tot = v.na * v.nb;
    HIPOTF_REQUIRE(lp && rp && dres, "NULL vector");
    const int nb = tot ? grid_for(tot) : 1;
    if (tot) {
        hipLaunchKernelGGL(k_backgrad, dim3(nb), dim3(256), 0, st, v, g, std::cos(theta),
                           std::sin(theta), lp, rp, P->partial.as<double>());
        HIPOTF_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, P->partial.as<double>(),
                       tot ? nb : 0, 1.0, dres);
    HIPOTF_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------------------
// Fused sweeps.  Consecutive generators that only move amplitudes inside the same rows are
// applied by one launch with the rows staged in LDS:
//   ROW_A : pure-beta generators act inside each alpha row (alpha-major layout);
//   ROW_B : pure-alpha generators act inside each beta row of the beta-major copy;
//   PAIR  : generators sharing one alpha excitation pattern (an alpha single and the
//           alpha-beta doubles built on it) act inside each row pair (u_a, w_a) of that pattern.
// The rotations of a segment are applied in sweep order (reverse order in the adjoint sweep),
// so the result equals the per-generator launches; gradients are reduced per generator from
// per-block partials in fixed order.
// ---------------------------------------------------------------------------------------
struct SegGen {
    GenDev g;
    const int32_t *lst;  // (u, w) pairs over the elements of a staged row
    i64 len;
    double c, s;
};

enum { SEG_ROW_A = 0, SEG_ROW_B = 1, SEG_PAIR = 2, SEG_SINGLE = 3 };

// A thread's first pair of a generator, (u, w) and the element string of u, loaded while the
// previous generator runs (list(): the pair, early in that generator; str(): the string, once the
// pair has arrived) so that the two dependent global loads overlap its work and barrier.  Row
// segments use it (a 16-orbital row generator is one pair per thread, so these loads are most of
// its latency); pair segments do not: there the early loads delay the first pair's reads of the
// rows in global memory more than they save
struct SegPrefetch {
    int32_t u = 0, w = 0;
    u64 e = 0;
    __device__ __forceinline__ void list(const SegGen &g, int tid) {
        if (tid < g.len) {
            u = g.lst[2 * tid];
            w = g.lst[2 * tid + 1];
        }
    }
    __device__ __forceinline__ void str(const SegGen &g, int tid, const u64 *__restrict__ strs) {
        if (tid < g.len) e = strs[u];
    }
};

// A generator's descriptor held one dword per lane (lanes 0..SEG_GEN_DW-1 of every wave), loaded
// with a vector load some generators ahead and read back with readlane when its turn comes: a
// scalar load issued that early would be waited for at the first LDS access after it (the two
// share one wait counter), so it would not overlap anything
constexpr int SEG_GEN_DW = (int)(sizeof(SegGen) / 4);
static_assert(sizeof(SegGen) % 4 == 0 && SEG_GEN_DW <= 64, "descriptor must fit one wave, whole dwords");
__device__ __forceinline__ uint32_t seg_gen_ld(const SegGen *__restrict__ gens, int k) {
    const int lane = threadIdx.x & 63;
    return lane < SEG_GEN_DW ? reinterpret_cast<const uint32_t *>(gens + k)[lane] : 0u;
}
__device__ __forceinline__ SegGen seg_gen_get(uint32_t d) {
    uint32_t w[SEG_GEN_DW];
#pragma unroll
    for (int i = 0; i < SEG_GEN_DW; ++i) w[i] = (uint32_t)__builtin_amdgcn_readlane((int)d, i);
    SegGen g;
    __builtin_memcpy(&g, w, sizeof(SegGen));
    return g;
}

// end of a generator in a segment block: the block barrier, and with SUM the block's partial of
// acc (wave sums in red, then thread 0 adds them in wave order, after the barrier, while the
// other waves go on; red alternates by generator, so the next generator's wave sums do not
// overwrite it before the following barrier)
template <bool SUM>
__device__ __forceinline__ void seg_partial(double acc, double *red, double *__restrict__ partial, i64 at) {
    if (SUM) {
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    }
    __syncthreads();  // (also orders this generator's global updates before the next one's reads)
    if (SUM && threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
        partial[at] = t;
    }
}

// A row of a segment block: in LDS (M = 0), in global memory (M = 2), or split (M = 1): its
// first sp elements staged in LDS, the rest in global memory
template <int M>
struct SegRow {
    double *s, *g;
    i64 sp;
    __device__ __forceinline__ double ld(i64 i) const {
        if constexpr (M == 0) return s[i];
        else if constexpr (M == 2) return g[i];
        else return i < sp ? s[i] : g[i];
    }
    __device__ __forceinline__ void st(i64 i, double v) const {
        if constexpr (M == 0) s[i] = v;
        else if constexpr (M == 2) g[i] = v;
        else if (i < sp) s[i] = v;
        else g[i] = v;
    }
};
// row k of a block with NST fully staged rows: staged, split (the next one), or global
constexpr int seg_mode(int k, int NST) { return k < NST ? 0 : (k == NST ? 1 : 2); }

// one block = one row of nel elements; BACK: adjoint step on (l, r) with gradient partials.
// NST of the block's rows (l, then r) are staged in LDS, the next one's first sp elements too
// (rows are staged in that order, row k at LDS offset k * nel); rows that do not fit stay in global
// memory and are updated in place, one generator after the other, with the launch sized to
// one block per CU so that the rows in flight stay in the L2 / Infinity Cache across the
// segment's generators.  All rows staged (NST = 1 forward, 2 adjoint) is the LDS form.
template <bool BACK, int NST, bool GRAD = false>
__global__ __launch_bounds__(1024) void k_seg_row(double *__restrict__ l, double *__restrict__ r,
                                                             i64 nrows, i64 nel, const u64 *__restrict__ row_strs,
                                                             const u64 *__restrict__ el_strs, int row_is_alpha,
                                                             const SegGen *__restrict__ gens, int ng,
                                                             double *__restrict__ partial, i64 sp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const i64 row = blockIdx.x;
    if (row >= nrows) return;
    const int tid = threadIdx.x, bd = blockDim.x;
    const u64 rs = row_strs[row];
    double *gl = l + row * nel;
    double *gr = BACK ? r + row * nel : nullptr;
    double *const sm = reinterpret_cast<double *>(smem_raw);
    const SegRow<seg_mode(0, NST)> xl{sm, gl, NST == 0 ? sp : nel};
    const SegRow<seg_mode(1, NST)> xr{sm + nel, gr, sp};
    // stage: rows below NST whole, row NST up to sp
    auto stage = [&](bool in) {
        for (int k = 0; k < (BACK ? 2 : 1); ++k) {
            const i64 m = k < NST ? nel : (k == NST ? sp : 0);
            double *g = k == 0 ? gl : gr, *sk = sm + (i64)k * nel;
            for (i64 t = tid; t < m; t += bd) {
                if (in) sk[t] = g[t];
                else g[t] = sk[t];
            }
        }
    };
    stage(true);
    __syncthreads();
    auto gen_at = [&](int kk) { return BACK ? ng - 1 - kk : kk; };
    uint32_t dc = ng > 0 ? seg_gen_ld(gens, gen_at(0)) : 0u, dn = ng > 1 ? seg_gen_ld(gens, gen_at(1)) : 0u;
    SegPrefetch pf;
    if (ng > 0) {
        const SegGen g0 = seg_gen_get(dc);
        pf.list(g0, tid);
        pf.str(g0, tid, el_strs);
    }
    __shared__ double red[2][16];
    for (int kk = 0; kk < ng; ++kk) {
        const int k = gen_at(kk);
        const SegGen sg = seg_gen_get(dc), sn = seg_gen_get(dn);  // sn: the next one (kk + 1 < ng)
        dc = dn;
        if (kk + 2 < ng) dn = seg_gen_ld(gens, gen_at(kk + 2));
        const SegPrefetch cur = pf;
        if (kk + 1 < ng) pf.list(sn, tid);
        double acc = 0.0;
        auto body = [&](int32_t u, int32_t w, u64 es) {
            const bool neg = row_is_alpha ? jw_sign(sg.g, rs, es) : jw_sign(sg.g, es, rs);
            const double ts = neg ? -sg.s : sg.s;
            if (GRAD) {
                // <l| G e^{theta G} |r> of this generator alone (k_grad), nothing stored
                const double tc = neg ? -sg.c : sg.c;
                const double lu = xl.ld(u), lw = xl.ld(w), ru = xr.ld(u), rw = xr.ld(w);
                acc += lu * (-sg.s * ru + tc * rw) + lw * (-sg.s * rw - tc * ru);
            } else if (!BACK) {
                const double xu = xl.ld(u), xw = xl.ld(w);
                xl.st(u, sg.c * xu - ts * xw);
                xl.st(w, ts * xu + sg.c * xw);
            } else {
                const double tc = neg ? -sg.c : sg.c;
                const double lu = xl.ld(u), lw = xl.ld(w), ru = xr.ld(u), rw = xr.ld(w);
                const double nlu = sg.c * lu + ts * lw, nlw = -ts * lu + sg.c * lw;
                acc += nlu * (-sg.s * ru + tc * rw) + nlw * (-sg.s * rw - tc * ru);
                xl.st(u, nlu);
                xl.st(w, nlw);
                xr.st(u, sg.c * ru + ts * rw);
                xr.st(w, -ts * ru + sg.c * rw);
            }
        };
        if (tid < sg.len) body(cur.u, cur.w, cur.e);
        for (i64 j = tid + bd; j < sg.len; j += bd) {
            const int32_t u = sg.lst[2 * j], w = sg.lst[2 * j + 1];
            body(u, w, el_strs[u]);
        }
        if (kk + 1 < ng) pf.str(sn, tid, el_strs);
        seg_partial<BACK>(acc, red[kk & 1], partial, (i64)k * nrows + row);
    }
    if (!GRAD) stage(false);
}

// one block = one alpha pair (u_a, w_a) of the segment's alpha pattern; rows of NB elements
// (the u and w strings of one pattern are disjoint, so blocks own disjoint row pairs).  NST of
// the rows (l_u, l_w, then r_u, r_w) are staged in LDS, the others updated in place in global
// memory, as in k_seg_row; NST = 2 forward / 4 adjoint is the LDS form
template <bool BACK, int NST, bool GRAD = false>
__global__ __launch_bounds__(1024) void k_seg_pair(double *__restrict__ l, double *__restrict__ r,
                                                              const int32_t *__restrict__ alist, i64 npairs, i64 NB,
                                                              const u64 *__restrict__ strsA,
                                                              const u64 *__restrict__ strsB,
                                                              const SegGen *__restrict__ gens, int ng,
                                                              double *__restrict__ partial, i64 sp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const i64 p = blockIdx.x;
    if (p >= npairs) return;
    const int tid = threadIdx.x, bd = blockDim.x;
    const int32_t ua = alist[2 * p], wa = alist[2 * p + 1];
    const u64 as = strsA[ua];
    double *const g0 = l + (i64)ua * NB, *const g1 = l + (i64)wa * NB;
    double *const h0 = BACK ? r + (i64)ua * NB : nullptr, *const h1 = BACK ? r + (i64)wa * NB : nullptr;
    double *const sm = reinterpret_cast<double *>(smem_raw);
    const SegRow<seg_mode(0, NST)> l0{sm, g0, NST == 0 ? sp : NB};
    const SegRow<seg_mode(1, NST)> l1{sm + NB, g1, NST == 1 ? sp : NB};
    const SegRow<seg_mode(2, NST)> r0{sm + 2 * NB, h0, NST == 2 ? sp : NB};
    const SegRow<seg_mode(3, NST)> r1{sm + 3 * NB, h1, sp};
    auto stage = [&](bool in) {
        for (int k = 0; k < (BACK ? 4 : 2); ++k) {
            const i64 m = k < NST ? NB : (k == NST ? sp : 0);
            double *g = k == 0 ? g0 : (k == 1 ? g1 : (k == 2 ? h0 : h1)), *sk = sm + (i64)k * NB;
            for (i64 t = tid; t < m; t += bd) {
                if (in) sk[t] = g[t];
                else g[t] = sk[t];
            }
        }
    };
    stage(true);
    __syncthreads();
    __shared__ double red[2][16];
    // descriptors loaded ahead in the adjoint and gradient forms; the forward form's first global
    // row reads would queue behind that load, so it takes each descriptor when it starts
    auto gen_at = [&](int kk) { return BACK ? ng - 1 - kk : kk; };
    uint32_t dc = BACK && ng > 0 ? seg_gen_ld(gens, gen_at(0)) : 0u;
    uint32_t dn = BACK && ng > 1 ? seg_gen_ld(gens, gen_at(1)) : 0u;
    for (int kk = 0; kk < ng; ++kk) {
        const int k = gen_at(kk);
        const SegGen sg = BACK ? seg_gen_get(dc) : gens[k];
        if (BACK) {
            dc = dn;
            if (kk + 2 < ng) dn = seg_gen_ld(gens, gen_at(kk + 2));
        }
        double acc = 0.0;
        auto body = [&](int32_t ub, int32_t wb, u64 bs) {
            const bool neg = jw_sign(sg.g, as, bs);
            const double ts = neg ? -sg.s : sg.s;
            if (GRAD) {
                const double tc = neg ? -sg.c : sg.c;
                const double lu = l0.ld(ub), lw = l1.ld(wb), ru = r0.ld(ub), rw = r1.ld(wb);
                acc += lu * (-sg.s * ru + tc * rw) + lw * (-sg.s * rw - tc * ru);
            } else if (!BACK) {
                const double xu = l0.ld(ub), xw = l1.ld(wb);
                l0.st(ub, sg.c * xu - ts * xw);
                l1.st(wb, ts * xu + sg.c * xw);
            } else {
                const double tc = neg ? -sg.c : sg.c;
                const double lu = l0.ld(ub), lw = l1.ld(wb), ru = r0.ld(ub), rw = r1.ld(wb);
                const double nlu = sg.c * lu + ts * lw, nlw = -ts * lu + sg.c * lw;
                acc += nlu * (-sg.s * ru + tc * rw) + nlw * (-sg.s * rw - tc * ru);
                l0.st(ub, nlu);
                l1.st(wb, nlw);
                r0.st(ub, sg.c * ru + ts * rw);
                r1.st(wb, -ts * ru + sg.c * rw);
            }
        };
        for (i64 j = tid; j < sg.len; j += bd) {
            const int32_t ub = sg.lst[2 * j], wb = sg.lst[2 * j + 1];
            body(ub, wb, strsB[ub]);
        }
        seg_partial<BACK>(acc, red[kk & 1], partial, (i64)k * npairs + p);
    }
    if (!GRAD) stage(false);
}

// dres[k] = sum_b partial[k * nb + b], one block per generator of the segment
__global__ __launch_bounds__(256) void k_seg_reduce(const double *__restrict__ partial, i64 nb,
                                                    double *__restrict__ dres) {
    const int k = blockIdx.x;
    double acc = 0.0;
    for (i64 i = threadIdx.x; i < nb; i += blockDim.x) acc += partial[(i64)k * nb + i];
    const double t = block_sum(acc);
    if (threadIdx.x == 0) dres[k] = t;
}

// ---- adjoint pair segments on interleaved rows: element e of the state holds (l[e], r[e]) as
// one 16-byte value, so each side of a rotation pair is one access instead of two (the accesses
// that miss the LDS are scattered, and each moves a whole sector whatever its width).  Row
// segments keep separate l and r rows: a whole l row in LDS serves them better.
template <int M>
struct SegRow2 {
    double2 *s, *g;
    i64 sp;
    __device__ __forceinline__ double2 ld(i64 i) const {
        if constexpr (M == 0) return s[i];
        else if constexpr (M == 2) return g[i];
        else return i < sp ? s[i] : g[i];
    }
    __device__ __forceinline__ void st(i64 i, double2 v) const {
        if constexpr (M == 0) s[i] = v;
        else if constexpr (M == 2) g[i] = v;
        else if (i < sp) s[i] = v;
        else g[i] = v;
    }
};

// one adjoint step of a pair: a = (l_u, r_u), b = (l_w, r_w); the arithmetic of k_seg_pair
__device__ __forceinline__ void adj_rot(double c, double s, bool neg, double2 &a, double2 &b, double &acc) {
    const double ts = neg ? -s : s, tc = neg ? -c : c;
    const double nlu = c * a.x + ts * b.x, nlw = -ts * a.x + c * b.x;
    acc += nlu * (-s * a.y + tc * b.y) + nlw * (-s * b.y - tc * a.y);
    const double nru = c * a.y + ts * b.y, nrw = -ts * a.y + c * b.y;
    a = make_double2(nlu, nru);
    b = make_double2(nlw, nrw);
}

// k_seg_pair<true, *> on interleaved rows: rows u, w of the alpha pair; NST (0..2) of them
// staged whole, the next one up to sp elements
template <int NST>
__global__ __launch_bounds__(1024) void k_seg_pair2(double2 *__restrict__ z, const int32_t *__restrict__ alist,
                                                    i64 npairs, i64 NB, const u64 *__restrict__ strsA,
                                                    const u64 *__restrict__ strsB,
                                                    const SegGen *__restrict__ gens, int ng,
                                                    double *__restrict__ partial, i64 sp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const i64 p = blockIdx.x;
    if (p >= npairs) return;
    const int tid = threadIdx.x, bd = blockDim.x;
    const int32_t ua = alist[2 * p], wa = alist[2 * p + 1];
    const u64 as = strsA[ua];
    double2 *const g0 = z + (i64)ua * NB, *const g1 = z + (i64)wa * NB;
    double2 *const sm = reinterpret_cast<double2 *>(smem_raw);
    const SegRow2<seg_mode(0, NST)> z0{sm, g0, NST == 0 ? sp : NB};
    const SegRow2<seg_mode(1, NST)> z1{sm + NB, g1, sp};
    auto stage = [&](bool in) {
        for (int k = 0; k < 2; ++k) {
            const i64 m = k < NST ? NB : (k == NST ? sp : 0);
            double2 *g = k == 0 ? g0 : g1, *sk = sm + (i64)k * NB;
            for (i64 t = tid; t < m; t += bd) {
                if (in) sk[t] = g[t];
                else g[t] = sk[t];
            }
        }
    };
    stage(true);
    __syncthreads();
    __shared__ double red[2][16];
    uint32_t dc = ng > 0 ? seg_gen_ld(gens, ng - 1) : 0u, dn = ng > 1 ? seg_gen_ld(gens, ng - 2) : 0u;
    for (int kk = 0; kk < ng; ++kk) {
        const int k = ng - 1 - kk;
        const SegGen sg = seg_gen_get(dc);
        dc = dn;
        if (kk + 2 < ng) dn = seg_gen_ld(gens, k - 2);
        double acc = 0.0;
        auto body = [&](int32_t ub, int32_t wb, u64 bs) {
            double2 a = z0.ld(ub), b = z1.ld(wb);
            adj_rot(sg.c, sg.s, jw_sign(sg.g, as, bs), a, b, acc);
            z0.st(ub, a);
            z1.st(wb, b);
        };
        for (i64 j = tid; j < sg.len; j += bd) {
            const int32_t ub = sg.lst[2 * j], wb = sg.lst[2 * j + 1];
            body(ub, wb, strsB[ub]);
        }
        seg_partial<true>(acc, red[kk & 1], partial, (i64)k * npairs + p);
    }
    stage(false);
}

// z[i] = (l[i], r[i]) and back
__global__ __launch_bounds__(256) void k_zip(const double *__restrict__ l, const double *__restrict__ r,
                                             double2 *__restrict__ z, i64 n) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        z[i] = make_double2(l[i], r[i]);
}
__global__ __launch_bounds__(256) void k_unzip(const double2 *__restrict__ z, double *__restrict__ l,
                                               double *__restrict__ r, i64 n) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const double2 v = z[i];
        l[i] = v.x;
        r[i] = v.y;
    }
}

// out[c * R + r] = in[r * Cc + c]  (R x Cc -> Cc x R), elements of type E
template <typename E>
__global__ __launch_bounds__(256) void k_transpose(const E *__restrict__ in, E *__restrict__ out, i64 R, i64 Cc) {
    __shared__ E tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const i64 c0 = (i64)blockIdx.x * 32, r0 = (i64)blockIdx.y * 32;
    for (int i = 0; i < 4; ++i) {
        const i64 rr = r0 + ty + 8 * i, cc = c0 + tx;
        if (rr < R && cc < Cc) tile[ty + 8 * i][tx] = in[rr * Cc + cc];
    }
    __syncthreads();
    for (int i = 0; i < 4; ++i) {
        const i64 cc = c0 + ty + 8 * i, rr = r0 + tx;
        if (rr < R && cc < Cc) out[cc * R + rr] = tile[tx][ty + 8 * i];
    }
}

struct Segment {
    int kind, first, count, pa;
    int nst;    // rows per block staged in LDS whole (the rest updated in place in global memory)
    i64 sp;     // elements of the next row staged too
    int nrow;   // rows per block: 1 (ROW) or 2 (PAIR), times 1 forward / 2 adjoint
};

static size_t seg_lds_budget() {
    int dev = 0, v = 0;
    HIPOTF_CHECK(hipGetDevice(&dev));
    HIPOTF_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    return std::min<size_t>((size_t)v, 160 * 1024) - 2048;
}

// group the sweep (generator indices in order) into fusable segments; nvec rows per element
// (1 for the forward expm sweep, 2 for the adjoint sweep) decide what fits LDS.  il: the
// adjoint sweep's pair segments on interleaved (l, r) rows (two rows of 16-byte elements)
static std::vector<Segment> plan_sweep(DevPool *P, const i64 *idxs, i64 n, int nvec, bool il = false) {
    const DevBasis *B = P->basis;
    const i64 NA = B->pA->N, NB = B->pB->N;
    const size_t budget = seg_lds_budget();
    // as many of a block's rows as fit the LDS are staged there, the others updated in place
    // in global memory; sweep_mode 1 stages none, 2 launches per generator, 3 stages one row
    // and half of the next (+4: the adjoint sweep on interleaved rows)
    const int mode = tuning(0).sweep_mode & 3;
    int esz = 8;
    // whole rows, then the part of the next row that fits (at least 1024 elements); mode 3 at
    // most one whole row and half of the next
    auto stage = [&](Segment &s, i64 len) {
        const i64 cap = mode == 1 ? 0 : (i64)(budget / esz);
        s.nst = (int)std::min<i64>(s.nrow, cap / len);
        s.sp = 0;
        if (mode == 3) {
            s.nst = std::min(s.nst, 1);
            if (s.nst < s.nrow && cap - s.nst * len >= len / 2) s.sp = len / 2;
        } else if (s.nst < s.nrow && cap - s.nst * len >= 1024) {
            s.sp = cap - s.nst * len;
        }
    };
    std::vector<Segment> segs;
    for (i64 i = 0; i < n;) {
        const Generator &G = P->gens[idxs[i]];
        Segment s{SEG_SINGLE, (int)i, 1, G.pa, 0, 0, 0};
        auto gen = [&](i64 j) -> const Generator & { return P->gens[idxs[j]]; };
        if (mode == 2) {
            // one launch per generator
        } else if (G.kind == 2) {
            s.kind = SEG_ROW_A;
            s.nrow = nvec;
            stage(s, NB);
            while (i + s.count < n && gen(i + s.count).kind == 2) ++s.count;
        } else if (G.kind == 3 || (G.kind == 1 && i + 1 < n && gen(i + 1).kind == 3 && gen(i + 1).pa == G.pa)) {
            s.kind = SEG_PAIR;
            s.nrow = il ? 2 : 2 * nvec;
            esz = il ? 16 : 8;
            stage(s, NB);
            esz = 8;
            while (i + s.count < n && gen(i + s.count).pa == G.pa && gen(i + s.count).kind != 2) ++s.count;
        } else if (G.kind == 1) {
            s.kind = SEG_ROW_B;
            s.nrow = nvec;
            stage(s, NA);
            while (i + s.count < n && gen(i + s.count).kind == 1 &&
                   !(i + s.count + 1 < n && gen(i + s.count + 1).kind == 3 &&
                     gen(i + s.count + 1).pa == gen(i + s.count).pa))
                ++s.count;
        }
        if (s.kind != SEG_SINGLE && s.count > 4096) s.count = 4096;
        segs.push_back(s);
        i += s.count;
    }
    return segs;
}

// c, s of sweep position i from the device angles (the gradient-only pass: no angles on the host)
__global__ __launch_bounds__(256) void k_seg_cs(SegGen *__restrict__ d, const double *__restrict__ th, i64 n) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        d[i].c = cos(th[i]);
        d[i].s = sin(th[i]);
    }
}

// k_grad with the angle read on the device
__global__ __launch_bounds__(256) void k_grad_dev(PairView v, GenDev g, const double *__restrict__ th,
                                                  const double *__restrict__ l, const double *__restrict__ r,
                                                  double *__restrict__ partial) {
    const double c = cos(*th), s = sin(*th);
    const i64 tot = v.na * v.nb;
    double acc = 0.0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (i64)gridDim.x * blockDim.x) {
        i64 u, w;
        u64 ua, ub;
        pair_at(v, i, u, w, ua, ub);
        const double tc = jw_sign(g, ua, ub) ? -c : c;
        const double ru = r[u], rw = r[w];
        acc += l[u] * (-s * ru + tc * rw) + l[w] * (-s * rw - tc * ru);
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

static void rot_grad_dev(DevPool *P, i64 idx, const double *theta, const double *lp, const double *rp,
                         double *dres, hipStream_t st) {
    PairView v;
    GenDev g;
    views(P, idx, v, g);
    const i64 tot = v.na * v.nb;
    const int nb = tot ? grid_for(tot) : 1;
    if (tot) {
        hipLaunchKernelGGL(k_grad_dev, dim3(nb), dim3(256), 0, st, v, g, theta, lp, rp, P->partial.as<double>());
        HIPOTF_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, P->partial.as<double>(), tot ? nb : 0, 1.0, dres);
    HIPOTF_CHECK(hipGetLastError());
}

// Forward expm sweep then adjoint sweep, both fused where possible, or (grad) the gradient-only
// pass: dres[i] = <psi| G_i e^{theta_i G_i} |lam> for every position, each generator on the same
// two vectors, with the angles on the device (thetas_dev).  psi / lam device vectors; dres[i]
// receives the raw backgrad (grad) value of sweep position i.
static void sweeps(DevPool *P, i64 n, const i64 *idxs, const double *thetas, const double *thetas_dev,
                   double *psi, double *lam, bool forward, bool backward, bool grad, double *dres,
                   hipStream_t st) {
    DevBasis *B = P->basis;
    const i64 NA = B->pA->N, NB = B->pB->N, dim = NA * NB;
    for (i64 i = 0; i < n; ++i)
        HIPOTF_REQUIRE(idxs[i] >= 0 && idxs[i] < (i64)P->gens.size(), "generator index out of range");
    // descriptors of every sweep position (for ROW_B the list is the alpha pattern)
    std::vector<SegGen> host(n);
    auto fill = [&](const std::vector<Segment> &segs) {
        for (const Segment &s : segs)
            for (int k = 0; k < s.count; ++k) {
                const i64 i = s.first + k;
                const Generator &G = P->gens[idxs[i]];
                SegGen &d = host[i];
                d.g.nops = G.nops;
                for (int q = 0; q < 8; ++q) {
                    d.g.pos[q] = q < G.nops ? G.pos[q] : 0;
                    d.g.cre[q] = q < G.nops ? G.cre[q] : 0;
                }
                const PairList &L = (s.kind == SEG_ROW_B) ? P->lists[G.pa] : P->lists[G.pb];
                d.lst = L.uw.as<int32_t>();
                d.len = L.len;
                d.c = thetas ? std::cos(thetas[i]) : 0.0;
                d.s = thetas ? std::sin(thetas[i]) : 0.0;
            }
    };
    bool transposed = false;  // grad: the beta-major copies are made once (the vectors do not change)
    auto run = [&](const std::vector<Segment> &segs, bool back, bool il) {
        fill(segs);
        P->segdesc.ensure((size_t)std::max<i64>(n, 1) * sizeof(SegGen));
        if (n) HIPOTF_CHECK(hipMemcpyAsync(P->segdesc.p, host.data(), n * sizeof(SegGen), hipMemcpyHostToDevice, st));
        if (grad && n) {
            hipLaunchKernelGGL(k_seg_cs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                               P->segdesc.as<SegGen>(), thetas_dev, n);
            HIPOTF_CHECK(hipGetLastError());
        }
        const SegGen *dg = P->segdesc.as<SegGen>();
        const int nvec = back ? 2 : 1;
        // interleaved pair segments: z = (psi, lam) element by element, made before the first
        // pair segment of a run of them and split back before a row segment or the end
        double2 *z = nullptr;
        bool zipped = false;
        if (il) {
            P->tz.ensure((size_t)dim * sizeof(double2));
            z = P->tz.as<double2>();
        }
        auto zip = [&](bool on) {
            if (on == zipped) return;
            if (on) hipLaunchKernelGGL(k_zip, dim3(grid_for(dim)), dim3(256), 0, st, psi, lam, z, dim);
            else hipLaunchKernelGGL(k_unzip, dim3(grid_for(dim)), dim3(256), 0, st, z, psi, lam, dim);
            HIPOTF_CHECK(hipGetLastError());
            zipped = on;
        };
        // segments with rows in global memory: one block per CU (LDS reserved for that where no
        // row is staged), so the rows the blocks in flight work on (256 CUs x up to 4 rows)
        // stay within the L2 / Infinity Cache
        const size_t glob_smem = seg_lds_budget() / 2 + 4096;
        for (size_t si = 0; si < segs.size(); ++si) {
            const Segment &s = back ? segs[segs.size() - 1 - si] : segs[si];
            zip(il && s.kind == SEG_PAIR);
            if (s.kind == SEG_SINGLE) {
                for (int kk = 0; kk < s.count; ++kk) {
                    const int k = back ? s.count - 1 - kk : kk;
                    const i64 i = s.first + k;
                    if (grad) rot_grad_dev(P, idxs[i], thetas_dev + i, psi, lam, dres + i, st);
                    else if (back) rot_backgrad(P, idxs[i], thetas[i], psi, lam, dres + i, st);
                    else rot_expm(P, idxs[i], thetas[i], psi, st);
                }
                continue;
            }
            if (s.kind == SEG_ROW_A || s.kind == SEG_ROW_B) {
                const bool A = s.kind == SEG_ROW_A;
                const i64 nrows = A ? NA : NB, nel = A ? NB : NA;
                double *l = psi, *r = lam;
                if (!A) {
                    P->tl.ensure(dim * sizeof(double));
                    P->tr.ensure(dim * sizeof(double));
                    dim3 g((unsigned)((NB + 31) / 32), (unsigned)((NA + 31) / 32));
                    if (!transposed) {
                        hipLaunchKernelGGL(k_transpose<double>, g, dim3(256), 0, st, psi, P->tl.as<double>(), NA, NB);
                        if (back)
                            hipLaunchKernelGGL(k_transpose<double>, g, dim3(256), 0, st, lam, P->tr.as<double>(), NA, NB);
                        transposed = grad;
                    }
                    l = P->tl.as<double>();
                    r = P->tr.as<double>();
                }
                const size_t staged = (size_t)(nel * s.nst + s.sp) * sizeof(double);
                const size_t smem = staged ? staged : glob_smem;
                // 1024 threads where the LDS holds one block per CU (latency of the per-generator
                // pair-list and string loads), 256 where several blocks share a CU
                const int bd = smem > glob_smem / 2 ? 1024 : 256;
                double *part = nullptr;
                if (back) {
                    P->segpart.ensure((size_t)s.count * nrows * sizeof(double));
                    part = P->segpart.as<double>();
                }
                const u64 *rstr = A ? B->pA->strs.as<u64>() : B->pB->strs.as<u64>();
                const u64 *estr = A ? B->pB->strs.as<u64>() : B->pA->strs.as<u64>();
                auto go_row = [&](auto kern, double *rr, double *pp) {
                    HIPOTF_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)smem));
                    hipLaunchKernelGGL(kern, dim3((unsigned)nrows), dim3(bd), smem, st, l, rr, nrows, nel, rstr,
                                       estr, A ? 1 : 0, dg + s.first, s.count, pp, s.sp);
                };
                if (grad) {
                    if (s.nst == 0) go_row(k_seg_row<true, 0, true>, r, part);
                    else if (s.nst == 1) go_row(k_seg_row<true, 1, true>, r, part);
                    else go_row(k_seg_row<true, 2, true>, r, part);
                } else if (back) {
                    if (s.nst == 0) go_row(k_seg_row<true, 0>, r, part);
                    else if (s.nst == 1) go_row(k_seg_row<true, 1>, r, part);
                    else go_row(k_seg_row<true, 2>, r, part);
                } else {
                    if (s.nst == 0) go_row(k_seg_row<false, 0>, (double *)nullptr, (double *)nullptr);
                    else go_row(k_seg_row<false, 1>, (double *)nullptr, (double *)nullptr);
                }
                HIPOTF_CHECK(hipGetLastError());
                if (!A && !grad) {
                    dim3 g((unsigned)((NA + 31) / 32), (unsigned)((NB + 31) / 32));
                    hipLaunchKernelGGL(k_transpose<double>, g, dim3(256), 0, st, P->tl.as<double>(), psi, NB, NA);
                    if (back)
                        hipLaunchKernelGGL(k_transpose<double>, g, dim3(256), 0, st, P->tr.as<double>(), lam, NB, NA);
                }
                if (back)
                    hipLaunchKernelGGL(k_seg_reduce, dim3(s.count), dim3(256), 0, st, part, nrows, dres + s.first);
                HIPOTF_CHECK(hipGetLastError());
            } else {  // SEG_PAIR
                const PairList &LA = P->lists[s.pa];
                const i64 np = LA.len;
                if (np == 0) {
                    if (back) HIPOTF_CHECK(hipMemsetAsync(dres + s.first, 0, s.count * sizeof(double), st));
                    continue;
                }
                const size_t staged = (size_t)(NB * s.nst + s.sp) * (il ? sizeof(double2) : sizeof(double));
                const size_t smem = staged ? staged : glob_smem;
                const int bd = smem > glob_smem / 2 ? 1024 : 256;
                double *part = nullptr;
                if (back) {
                    P->segpart.ensure((size_t)s.count * np * sizeof(double));
                    part = P->segpart.as<double>();
                }
                if (il) {
                    auto go = [&](auto kern) {
                        HIPOTF_CHECK(hipFuncSetAttribute((const void *)kern,
                                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
                        hipLaunchKernelGGL(kern, dim3((unsigned)np), dim3(bd), smem, st, z, LA.uw.as<int32_t>(), np,
                                           NB, B->pA->strs.as<u64>(), B->pB->strs.as<u64>(), dg + s.first,
                                           s.count, part, s.sp);
                    };
                    if (s.nst == 0) go(k_seg_pair2<0>);
                    else if (s.nst == 1) go(k_seg_pair2<1>);
                    else go(k_seg_pair2<2>);
                    HIPOTF_CHECK(hipGetLastError());
                    hipLaunchKernelGGL(k_seg_reduce, dim3(s.count), dim3(256), 0, st, part, np, dres + s.first);
                    HIPOTF_CHECK(hipGetLastError());
                    continue;
                }
                auto go_pair = [&](auto kern, double *rr, double *pp) {
                    HIPOTF_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)smem));
                    hipLaunchKernelGGL(kern, dim3((unsigned)np), dim3(bd), smem, st, psi, rr, LA.uw.as<int32_t>(),
                                       np, NB, B->pA->strs.as<u64>(), B->pB->strs.as<u64>(), dg + s.first,
                                       s.count, pp, s.sp);
                };
                if (grad) {
                    if (s.nst == 0) go_pair(k_seg_pair<true, 0, true>, lam, part);
                    else if (s.nst == 1) go_pair(k_seg_pair<true, 1, true>, lam, part);
                    else if (s.nst == 2) go_pair(k_seg_pair<true, 2, true>, lam, part);
                    else if (s.nst == 3) go_pair(k_seg_pair<true, 3, true>, lam, part);
                    else go_pair(k_seg_pair<true, 4, true>, lam, part);
                } else if (back) {
                    if (s.nst == 0) go_pair(k_seg_pair<true, 0>, lam, part);
                    else if (s.nst == 1) go_pair(k_seg_pair<true, 1>, lam, part);
                    else if (s.nst == 2) go_pair(k_seg_pair<true, 2>, lam, part);
                    else if (s.nst == 3) go_pair(k_seg_pair<true, 3>, lam, part);
                    else go_pair(k_seg_pair<true, 4>, lam, part);
                } else {
                    if (s.nst == 0) go_pair(k_seg_pair<false, 0>, (double *)nullptr, (double *)nullptr);
                    else if (s.nst == 1) go_pair(k_seg_pair<false, 1>, (double *)nullptr, (double *)nullptr);
                    else go_pair(k_seg_pair<false, 2>, (double *)nullptr, (double *)nullptr);
                }
                HIPOTF_CHECK(hipGetLastError());
                if (back)
                    hipLaunchKernelGGL(k_seg_reduce, dim3(s.count), dim3(256), 0, st, part, np, dres + s.first);
                HIPOTF_CHECK(hipGetLastError());
            }
        }
        zip(false);
    };
    if (grad) {
        run(plan_sweep(P, idxs, n, 2), true, false);
        return;
    }
    if (forward) run(plan_sweep(P, idxs, n, 1), false, false);
    if (backward) {
        // sweep_mode + 4: the pair segments on interleaved rows
        const bool il = (tuning(0).sweep_mode & 4) != 0;
        run(plan_sweep(P, idxs, n, 2, il), true, il);
    }
}

void fused_sweeps(DevPool *P, i64 n, const i64 *idxs, const double *thetas, double *psi, double *lam,
                  bool forward, bool backward, double *dres, hipStream_t st) {
    sweeps(P, n, idxs, thetas, nullptr, psi, lam, forward, backward, false, dres, st);
}

// batch tran (op/batchtran.hpp:517-528): tran_update summed per generator, i.e. the gradient
// of every generator at theta = 0 (the gradient-only pass on a zero angle vector)
void batch_tran(DevPool *P, const double *lp, const double *rp, double *trans, hipStream_t st) {
    const i64 n = (i64)P->gens.size();
    P->zeros.ensure((size_t)std::max<i64>(n, 1) * sizeof(double));
    fill(P->zeros.as<double>(), n, 0.0, st);
    fused_grads(P, P->zeros.as<double>(), lp, rp, trans, st);
}

void fused_grads(DevPool *P, const double *thetas, const double *lp, const double *rp, double *grads,
                 hipStream_t st) {
    const i64 n = (i64)P->gens.size();
    std::vector<i64> idxs(n);
    std::iota(idxs.begin(), idxs.end(), 0);
    // the vectors are only read in the gradient-only pass
    sweeps(P, n, idxs.data(), nullptr, thetas, const_cast<double *>(lp), const_cast<double *>(rp), false,
           false, true, grads, st);
}

// ---------------------------------------------------------------------------------------
// dot product (two-stage, deterministic)
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dot_partial(const double *__restrict__ a,
                                                     const double *__restrict__ b, i64 n,
                                                     double *__restrict__ partial) {
    double acc = 0.0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        acc += a[i] * b[i];
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

void dev_dot(const double *a, const double *b, i64 n, double *dres, DevBuf &scratch, hipStream_t st) {
    const int nb = (int)std::max<i64>(1, std::min<i64>((n + 255) / 256, 2048));
    scratch.ensure(2048 * sizeof(double));
    if (n > 0) {
        hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(256), 0, st, a, b, n, scratch.as<double>());
        HIPOTF_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, scratch.as<double>(), n > 0 ? nb : 0,
                       1.0, dres);
    HIPOTF_CHECK(hipGetLastError());
}

}  // namespace hipotf
