// Same-spin operator tables built on the device (the host build in ham.cpp stays for rows longer
// than the in-LDS sort takes).
//
// For each string I of one spin: the row G[I, .] of the same-spin operator (the diagonal, the
// k (n-k) single and C(k,2) C(n-k,2) double excitations, coefficients as ham.cpp's
// build_same_spin forms them from h_pq and (pq|rs)), its source strings as colex ranks, the row
// sorted by source in LDS; then the row-block tables of the same-spin SpMM -- per block of rb
// consecutive strings the ascending union of their sources and an rb-wide coefficient column per
// union entry -- with the union sorted in LDS.  Replaces the host construction of
// src/include/ham/ham.hpp:635-996 / :998-1069 (JW terms and SVD groups) with a direct build.
#include "hipotf.hpp"

#include <algorithm>
#include <climits>

namespace hipotf {

constexpr int TB_LP = 2048;  // padded row length of the in-LDS row sort
constexpr int TB_BW = 33;    // binomial table C(a, b), a, b <= 32

__device__ __forceinline__ u32 tb_binom(int a, int b) {
    if (b < 0 || b > a) return 0u;
    u64 c = 1;
    for (int i = 0; i < b; ++i) c = c * (u64)(a - i) / (u64)(i + 1);
    return (u32)c;
}

// ascending bitonic sort of key[0..LP) (with val when V) by nt threads
template <bool V, typename VT>
__device__ void tb_bitonic(int *key, VT *val, int LP, int tid, int nt) {
    for (int size = 2; size <= LP; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < LP; i += nt) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const int ki = key[i], kj = key[j];
                    if ((ki > kj) == up) {
                        key[i] = kj;
                        key[j] = ki;
                        if constexpr (V) {
                            const VT t = val[i];
                            val[i] = val[j];
                            val[j] = t;
                        }
                    }
                }
            }
            __syncthreads();
        }
}

// first index in sorted a[0..m) whose value is >= x
__device__ __forceinline__ int tb_lower(const int *a, int m, i64 x) {
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((i64)a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// one workgroup per string I: row I of the same-spin table, sorted by source
__global__ __launch_bounds__(256) void k_ss_rows(const u64 *__restrict__ strs, i64 N, int n, int k, int L,
                                                 const double *__restrict__ h, const double *__restrict__ eri,
                                                 int32_t *__restrict__ J, double *__restrict__ c64,
                                                 float *__restrict__ c32, int32_t *__restrict__ P,
                                                 double *__restrict__ gdiag) {
    __shared__ int key[TB_LP];
    __shared__ double val[TB_LP];
    __shared__ u32 bn[TB_BW * TB_BW];
    __shared__ int occ[32], vir[32];
    const int tid = threadIdx.x, nt = blockDim.x;
    const i64 I = blockIdx.x;
    const u64 s = strs[I];
    for (int t = tid; t < TB_BW * TB_BW; t += nt) bn[t] = tb_binom(t / TB_BW, t % TB_BW);
    if (tid == 0) {
        int no = 0, nv = 0;
        for (int p = 0; p < n; ++p) {
            if ((s >> p) & 1) occ[no++] = p;
            else vir[nv++] = p;
        }
    }
    __syncthreads();
    const int nv = n - k;
    const i64 n2 = (i64)n * n;
    auto E = [&](int p, int q, int r, int t) { return eri[((i64)p * n + q) * n2 + (i64)r * n + t]; };
    auto rank = [&](u64 x) {
        u32 r = 0;
        int j = 0;
        while (x) {
            r += bn[__builtin_ctzll(x) * TB_BW + j + 1];
            x &= x - 1;
            ++j;
        }
        return (int)r;
    };
    auto between = [](int a, int b) {
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        return ((1ull << hi) - 1) & ~((1ull << (lo + 1)) - 1);
    };
    // (the index of pair e of m items in lexicographic (i < j) order)
    auto pair_of = [](int e, int m, int &i, int &j) {
        i = 0;
        while (e >= m - 1 - i) {
            e -= m - 1 - i;
            ++i;
        }
        j = i + 1 + e;
    };
    const int ns = k * nv, np = nv * (nv - 1) / 2;
    for (int e = tid; e < TB_LP; e += nt) {
        int kk = INT_MAX;
        double v = 0.0;
        if (e == 0) {
            for (int a = 0; a < k; ++a) {
                const int p = occ[a];
                v += h[p * n + p];
                for (int b = 0; b < k; ++b) {
                    const int q = occ[b];
                    v += 0.5 * (E(p, p, q, q) - E(p, q, q, p));
                }
            }
            kk = (int)I;
            gdiag[I] = v;
        } else if (e <= ns) {
            const int io = (e - 1) / nv, iv = (e - 1) % nv;
            const int r = occ[io], t = vir[iv];
            v = h[r * n + t];
            for (int a = 0; a < k; ++a) {
                const int j = occ[a];
                if (j == r) continue;
                v += E(r, t, j, j) - E(r, j, j, t);
            }
            if (__popcll(s & between(r, t)) & 1) v = -v;
            kk = rank(s ^ (1ull << r) ^ (1ull << t));
        } else if (e < L) {
            const int q = e - 1 - ns;
            int a1, a2, b1, b2;
            pair_of(q / np, k, a1, a2);
            pair_of(q % np, nv, b1, b2);
            const int r1 = occ[a1], r2 = occ[a2], s1 = vir[b1], s2 = vir[b2];
            const u64 js = s ^ (1ull << r1) ^ (1ull << r2) ^ (1ull << s1) ^ (1ull << s2);
            const int g1 = __popcll(js & between(r1, s1)) & 1;
            const u64 j1 = js ^ (1ull << s1) ^ (1ull << r1);
            const int g2 = __popcll(j1 & between(r2, s2)) & 1;
            v = E(r1, s1, r2, s2) - E(r1, s2, r2, s1);
            if (g1 ^ g2) v = -v;
            kk = rank(js);
        }
        key[e] = kk;
        val[e] = v;
    }
    __syncthreads();
    tb_bitonic<true>(key, val, TB_LP, tid, nt);
    for (int e = tid; e < L; e += nt) {
        J[I * L + e] = key[e];
        c64[I * L + e] = val[e];
        c32[I * L + e] = (float)val[e];
    }
    if (tid <= SS_RANGES)
        P[I * (SS_RANGES + 1) + tid] = tid == SS_RANGES ? L : tb_lower(key, L, (i64)tid * N / SS_RANGES);
}

// per block of rb rows: the ascending union of their sources.  FILL = false: its size into
// cnt[g]; FILL = true: the union into Jr[off[g]..], the rb-wide coefficient columns and the
// range offsets.  LP (power of two) >= rb * L keys in LDS.
template <bool FILL>
__global__ __launch_bounds__(1024) void k_rb_union(const int32_t *__restrict__ J, const double *__restrict__ c64,
                                                   i64 N, int L, int rb, int LP, int32_t *__restrict__ cnt,
                                                   const i64 *__restrict__ off, int32_t *__restrict__ Jr,
                                                   double *__restrict__ cr64, float *__restrict__ cr32,
                                                   int32_t *__restrict__ Pr) {
    extern __shared__ int lds[];
    int *key = lds;              // LP keys
    int *tsum = lds + LP;        // per-thread unique counts, then their exclusive scan
    const int tid = threadIdx.x, nt = blockDim.x;
    const i64 g = blockIdx.x, r0 = g * rb, r1 = r0 + rb < N ? r0 + rb : N;
    const int m = (int)((r1 - r0) * L);
    for (int i = tid; i < LP; i += nt) key[i] = i < m ? J[r0 * L + i] : INT_MAX;
    __syncthreads();
    tb_bitonic<false, int>(key, nullptr, LP, tid, nt);
    // each thread owns a contiguous chunk of the sorted keys
    const int per = LP / nt, c0 = tid * per;
    int mine[32];
    int u = 0;
    for (int i = 0; i < per; ++i) {
        const int x = key[c0 + i];
        const bool first = x != INT_MAX && (c0 + i == 0 || key[c0 + i - 1] != x);
        mine[i] = first ? x : INT_MAX;
        u += first ? 1 : 0;
    }
    tsum[tid] = u;
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int t = 0; t < nt; ++t) {
            const int c = tsum[t];
            tsum[t] = acc;
            acc += c;
        }
        tsum[nt] = acc;
    }
    __syncthreads();
    const int U = tsum[nt];
    if (!FILL) {
        if (tid == 0) cnt[g] = U;
        return;
    }
    // compact the union in place (every thread has read its chunk before the barrier above)
    int pos = tsum[tid];
    for (int i = 0; i < per; ++i)
        if (mine[i] != INT_MAX) key[pos++] = mine[i];
    __syncthreads();
    const i64 o = off[g];
    for (int i = tid; i < U; i += nt) Jr[o + i] = key[i];
    for (int i = tid; i < m; i += nt) {
        const int row = i / L;
        const i64 src = r0 * L + i;
        const int p = tb_lower(key, U, J[src]);
        cr64[(o + p) * rb + row] = c64[src];
        cr32[(o + p) * rb + row] = (float)c64[src];
    }
    if (tid <= SS_RANGES)
        Pr[g * (SS_RANGES + 1) + tid] =
            (int32_t)(o + (tid == SS_RANGES ? U : tb_lower(key, U, (i64)tid * N / SS_RANGES)));
}

// single-excitation table of one spin: sx[slot][I], slot = io (n - k) + iv, r = io-th occupied and
// t = iv-th empty orbital of string I, packed with the colex rank of I - r + t and the parity of
// the occupied orbitals strictly between r and t
__global__ __launch_bounds__(256) void k_sx(const u64 *__restrict__ strs, i64 N, int n, int k,
                                            u32 *__restrict__ sx) {
    __shared__ u32 bn[TB_BW * TB_BW];
    for (int t = threadIdx.x; t < TB_BW * TB_BW; t += blockDim.x) bn[t] = tb_binom(t / TB_BW, t % TB_BW);
    __syncthreads();
    const int nv = n - k;
    const i64 tot = N * (i64)(k * nv);
    for (i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x; x < tot; x += (i64)gridDim.x * blockDim.x) {
        const i64 slot = x / N, I = x - slot * N;
        const int io = (int)(slot / nv), iv = (int)(slot - (i64)io * nv);
        const u64 s = strs[I];
        u64 m = s, e = ~s & ((n < 64) ? ((1ull << n) - 1) : ~0ull);
        for (int i = 0; i < io; ++i) m &= m - 1;
        for (int i = 0; i < iv; ++i) e &= e - 1;
        const int r = __builtin_ctzll(m), t = __builtin_ctzll(e);
        const u64 j = s ^ (1ull << r) ^ (1ull << t);
        u32 rank = 0;
        int q = 0;
        for (u64 y = j; y; y &= y - 1, ++q) rank += bn[__builtin_ctzll(y) * TB_BW + q + 1];
        const int lo = r < t ? r : t, hi = r < t ? t : r;
        const bool neg = __popcll(s & ((1ull << hi) - 1) & ~((1ull << (lo + 1)) - 1)) & 1;
        sx[x] = sx_pack(rank, r, t, neg);
    }
}

void build_sx_dev(DevSpin &D) {
    const i64 tot = D.N * (i64)D.nsx;
    HIPOTF_REQUIRE(D.N < (1 << 20), "too many strings per spin for the packed tables (2^20)");
    HIPOTF_REQUIRE(D.n <= 32, "internal: excitation table build for n > 32");
    D.sx.alloc((size_t)tot * sizeof(u32));
    if (tot == 0) return;
    const i64 nb = std::min<i64>((tot + 255) / 256, 65536);
    hipLaunchKernelGGL(k_sx, dim3((unsigned)nb), dim3(256), 0, 0, D.strs.as<u64>(), D.N, D.n, D.k, D.sx.as<u32>());
    HIPOTF_CHECK(hipGetLastError());
    HIPOTF_CHECK(hipDeviceSynchronize());
}

__global__ __launch_bounds__(256) void k_jab(const u64 *__restrict__ strs, i64 N, int n,
                                             const double *__restrict__ eri, double *__restrict__ jab) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * n) return;
    const i64 a = i / n;
    const int r = (int)(i - a * n);
    const u64 s = strs[a];
    const i64 n2 = (i64)n * n;
    double v = 0.0;
    for (int p = 0; p < n; ++p)
        if ((s >> p) & 1) v += eri[((i64)p * n + p) * n2 + (i64)r * n + r];
    jab[i] = v;
}

bool build_same_spin_dev(SameSpinTable &T, const DevSpin &S, const DevBuf &hdev, const DevBuf &eri64, int n) {
    const int L = 1 + S.nsx + S.ndx;
    if (L > TB_LP || n > 32 || S.N < 1 || S.N >= (1ll << 31) || (i64)S.N * L >= (1ll << 31)) return false;
    T.L = L;
    T.N = S.N;
    T.J.alloc((size_t)S.N * L * sizeof(int32_t));
    T.c64.alloc((size_t)S.N * L * sizeof(double));
    T.c32.alloc((size_t)S.N * L * sizeof(float));
    T.P.alloc((size_t)S.N * (SS_RANGES + 1) * sizeof(int32_t));
    T.gdiag.alloc((size_t)S.N * sizeof(double));
    hipLaunchKernelGGL(k_ss_rows, dim3((unsigned)S.N), dim3(256), 0, 0, S.strs.as<u64>(), S.N, n, S.k, L,
                       hdev.as<double>(), eri64.as<double>(), T.J.as<int32_t>(), T.c64.as<double>(),
                       T.c32.as<float>(), T.P.as<int32_t>(), T.gdiag.as<double>());
    HIPOTF_CHECK(hipGetLastError());
    HIPOTF_CHECK(hipDeviceSynchronize());
    T.dev_built = true;
    return true;
}

void row_blocks_dev(SameSpinTable &T, RowBlockTable &R, int rb) {
    const i64 N = T.N, nblk = (N + rb - 1) / rb;
    const int L = T.L;
    int LP = 1024;
    while (LP < rb * L) LP <<= 1;
    HIPOTF_REQUIRE(LP <= 32768, "internal: row block too long for the in-LDS union");
    const size_t smem = ((size_t)LP + 1024 + 1) * sizeof(int);
    DevBuf cnt, off;
    cnt.alloc((size_t)nblk * sizeof(int32_t));
    HIPOTF_CHECK(hipFuncSetAttribute((const void *)k_rb_union<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)smem));
    HIPOTF_CHECK(hipFuncSetAttribute((const void *)k_rb_union<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)smem));
    hipLaunchKernelGGL(k_rb_union<false>, dim3((unsigned)nblk), dim3(1024), smem, 0, T.J.as<int32_t>(),
                       T.c64.as<double>(), N, L, rb, LP, cnt.as<int32_t>(), nullptr, nullptr, nullptr, nullptr,
                       nullptr);
    HIPOTF_CHECK(hipGetLastError());
    std::vector<int32_t> hc(nblk);
    HIPOTF_CHECK(hipMemcpy(hc.data(), cnt.p, (size_t)nblk * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<i64> ho(nblk + 1, 0);
    for (i64 g = 0; g < nblk; ++g) ho[g + 1] = ho[g] + hc[g];
    const i64 E = ho[nblk];
    HIPOTF_REQUIRE(E < (1ll << 31), "row-block table too large for 32-bit offsets");
    off.upload(ho);
    R.rb = rb;
    R.nblk = nblk;
    R.entries = E;
    R.J.alloc((size_t)E * sizeof(int32_t));
    R.c64.alloc((size_t)E * rb * sizeof(double));
    R.c32.alloc((size_t)E * rb * sizeof(float));
    R.P.alloc((size_t)nblk * (SS_RANGES + 1) * sizeof(int32_t));
    HIPOTF_CHECK(hipMemset(R.c64.p, 0, (size_t)E * rb * sizeof(double)));
    HIPOTF_CHECK(hipMemset(R.c32.p, 0, (size_t)E * rb * sizeof(float)));
    hipLaunchKernelGGL(k_rb_union<true>, dim3((unsigned)nblk), dim3(1024), smem, 0, T.J.as<int32_t>(),
                       T.c64.as<double>(), N, L, rb, LP, nullptr, off.as<i64>(), R.J.as<int32_t>(),
                       R.c64.as<double>(), R.c32.as<float>(), R.P.as<int32_t>());
    HIPOTF_CHECK(hipGetLastError());
    HIPOTF_CHECK(hipDeviceSynchronize());
}

void build_jab_dev(DevBuf &jab, const DevSpin &A, int n, const DevBuf &eri64) {
    const i64 tot = A.N * n;
    jab.alloc((size_t)tot * sizeof(double));
    if (tot == 0) return;
    hipLaunchKernelGGL(k_jab, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, A.strs.as<u64>(), A.N, n,
                       eri64.as<double>(), jab.as<double>());
    HIPOTF_CHECK(hipGetLastError());
    HIPOTF_CHECK(hipDeviceSynchronize());
}

}  // namespace hipotf
