// Alpha-beta part of the sigma build on the matrix cores (included by sigma.hip), n <= 16.
//
// For a destination alpha row Ia, with c running over its alpha connections -- the nsxA single
// excitations Ja_c = Ia - p_c + q_c with sign s_c, then Ia itself with the diagonal integral
// row sum_{p in Ia} (pp|..) -- the alpha-beta part of the sigma row is
//
//   sigma~[Ia, Ib] += sum_{r <= t} sum_{Jb : Ib = E_(r,t) Jb} sgn(r, t, Jb) Y[(r,t), Jb]
//   Y[(r,t), Jb]    = sum_c W[(r,t), c] x~[Ja_c, Jb],      W[(r,t), c] = s_c (p_c q_c | r t)
//
// E_(r,t) (r < t) moves the one beta electron of Jb that sits on r or t to the other orbital,
// with sgn the parity of the occupied orbitals of Jb strictly between r and t; it does not
// apply when both or neither are occupied.  The rows r = t apply when r is occupied and leave
// Jb as it is (sgn = +1).  Y is a dense product of the n(n+1)/2 x nca matrix W with the
// nca x NB matrix of source rows, so it runs as 16x16x4 MFMA tiles (FP64 or FP32 units) with
// no gathers: the rows that do not apply to a Jb are computed and dropped (47% of them at 16
// orbitals / 8 electrons).  Each applying element of Y is then added into an FP64 LDS copy of
// the destination row (ds_add_f64), with the target rank read from two colex tables (low and
// high byte of the 16-bit string).  Same sums as the source-lane kernel (ab_src.hpp), in
// another order.
//
// Workgroup = destination alpha row (XCD-interleaved like the other alpha-beta kernels),
// NWV waves; beta strings in tiles of 16 (the MFMA's N), tiles dealt to the waves round robin.
// W sits in LDS in MFMA-operand order for KSP k-steps (4 connections each) at a time: the
// A operand of tile (k-step s, row tile m) for lane l is entry (s * MT + m) * 64 + l, a
// conflict-free 8- or 4-byte read.  When all k-steps do not fit beside the FP64 row (FP64 at
// 16 orbitals: 78 KB of W + 103 KB of row), the connections are done in passes, each
// scattering its partial Y.
//
// Operand layout of v_mfma_{f64,f32}_16x16x4 on gfx950 (measured): lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; result element v of lane l is D[4 (l / 16) + v][l % 16] for f32 and
// D[4 v + l / 16][l % 16] for f64.
#pragma once

constexpr int AB_MMA_RANKS = 256 + 9 * 256;
// waves per workgroup (one workgroup per CU: the LDS holds the FP64 destination row): FP64 8
// (224 VGPRs), FP32 12 (136 VGPRs; 8 waves: 4% slower)
constexpr int AB_MMA_NWV64 = 8, AB_MMA_NWV32 = 12;  // colex rank tables: low byte, (low count, high byte)

__device__ __forceinline__ u32 ab_mma_binom(int a, int b) {
    if (b < 0 || b > a) return 0u;
    u32 c = 1;
    for (int i = 0; i < b; ++i) c = c * (u32)(a - i) / (u32)(i + 1);
    return c;
}

template <typename T>
struct AbMmaOps;
template <>
struct AbMmaOps<double> {
    typedef double V4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V4 mma(double a, double b, V4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static constexpr __device__ int row(int g, int v) { return 4 * v + g; }  // of result element v
};
template <>
struct AbMmaOps<float> {
    typedef float V4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V4 mma(float a, float b, V4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static constexpr __device__ int row(int g, int v) { return 4 * g + v; }
};

// LDS bytes of k_ab_mma for NB beta strings
template <typename T, int MT, int KSP>
constexpr size_t ab_mma_lds(i64 NB) {
    return (size_t)NB * sizeof(double) + (size_t)KSP * MT * 64 * sizeof(T) + MT * 16 * 4 +
           AB_MMA_RANKS * 2 + AB_SRC_CONN * 4;
}

// nks = k-steps of all connections, ceil((nsxA + 1) / 4); dw rows are nwd entries wide
template <typename T, int MT, int KSP, int NWV>
__global__ __launch_bounds__(NWV * 64) __attribute__((amdgpu_waves_per_eu(NWV / 4, NWV / 4)))
void k_ab_mma(const T *__restrict__ xt, T *__restrict__ s1, int nvst, int v0, const u32 *__restrict__ sxA,
              const u64 *__restrict__ strsB, const T *__restrict__ eri, const T *__restrict__ dw, int nwd, int n,
              i64 NA, int NB, int nsxA, int nks) {
    using Ops = AbMmaOps<T>;
    using V4 = typename Ops::V4;
    constexpr int BD = NWV * 64;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *S = reinterpret_cast<double *>(smem_raw);                       // destination row
    T *Wl = reinterpret_cast<T *>(smem_raw + (size_t)NB * sizeof(double));  // KSP x MT x 64
    u32 *rinfo = reinterpret_cast<u32 *>(Wl + KSP * MT * 64);               // per W row
    unsigned short *rk = reinterpret_cast<unsigned short *>(rinfo + MT * 16);
    u32 *conn = reinterpret_cast<u32 *>(rk + AB_MMA_RANKS);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, col = lane & 15;
    const i64 per = (NA + 7) / 8;
    const i64 Ia = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (Ia >= NA) return;
    const int n2 = n * n;
    static_assert(MT == 9, "row classes: 1 diagonal, 2 + 2 one-byte and 4 cross-byte tiles");

    for (int t = tid; t < nsxA; t += BD) conn[t] = sxA[(i64)t * NA + Ia];
    // W rows by class of the orbital pair, so the target rank is formed with the fewest table
    // reads: tile 0 the rows r = t (r < 16), tiles 1-2 the pairs inside the low byte (r < t < 8),
    // tiles 3-4 inside the high byte (8 <= r < t), tiles 5-8 the pairs across (r < 8 <= t,
    // row 80 + 8 r + t - 8).  A row's entry holds its orbital mask (low half) and the mask of
    // the orbitals strictly between r and t (high half); rows past n, and padding, hold 0 and
    // never apply.
    for (int i = tid; i < MT * 16; i += BD) {
        int r = -1, t = -1;
        if (i < 16) {
            r = t = i;
        } else if (i < 80) {
            const int lo = i < 48 ? 0 : 8, e0 = i - (i < 48 ? 16 : 48);
            int e = e0, a = 0;
            while (a < 7 && e >= 7 - a) {
                e -= 7 - a;
                ++a;
            }
            if (e0 < 28) {
                r = lo + a;
                t = r + 1 + e;
            }
        } else {
            r = (i - 80) >> 3;
            t = 8 + ((i - 80) & 7);
        }
        u32 v = 0;
        if (r >= 0 && t < n)
            v = r == t ? 1u << r : (1u << r) | (1u << t) | ((((1u << t) - (2u << r)) & 0xffffu) << 16);
        rinfo[i] = v;
    }
    // colex ranks rank(s) = rk[s & 255] + rk[256 + 256 * popc(s & 255) + (s >> 8)], from the
    // binomials C(a, b), a < 16, b <= 16, staged at the start of the (not yet used) destination
    // row and W area
    u32 *bn = reinterpret_cast<u32 *>(S);
    for (int e = tid; e < 16 * 17; e += BD) bn[e] = ab_mma_binom(e / 17, e % 17);
    __syncthreads();
    for (int e = tid; e < AB_MMA_RANKS; e += BD) {
        const int hi = e >= 256;
        const int c0 = hi ? (e - 256) >> 8 : 0, bits = hi ? (e - 256) & 255 : e, base = hi ? 8 : 0;
        u32 v = 0;
        int j = c0;
        for (int q = 0; q < 8; ++q)
            if ((bits >> q) & 1) {
                v += j + 1 <= 16 ? bn[(base + q) * 17 + j + 1] : 0u;
                ++j;
            }
        rk[e] = (unsigned short)v;
    }
    __syncthreads();
    for (int t = tid; t < NB; t += BD) S[t] = 0.0;
    __syncthreads();

    // FP64: this lane's W rows (result element v of row tile m) stay in registers; the FP32
    // kernel (12 waves, 168 VGPRs) reads them from LDS
    constexpr bool HOIST = sizeof(T) == 8;
    u32 rl[HOIST ? MT : 1][4];
    if constexpr (HOIST) {
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v) rl[m][v] = rinfo[m * 16 + Ops::row(g, v)];
    }
    auto row_info = [&](int m, int v) -> u32 {
        if constexpr (HOIST) return rl[m % MT][v];
        return rinfo[m * 16 + Ops::row(g, v)];
    };
    const int ntiles = (NB + 15) >> 4;
    const int npass = (nks + KSP - 1) / KSP;
    for (int pass = 0; pass < npass; ++pass) {
        const int s0 = pass * KSP;
        // W of k-steps s0 .. s0 + KSP - 1 in operand order (zero past the last connection)
        for (int e = tid; e < KSP * MT * 64; e += BD) {
            const int l = e & 63, m = (e >> 6) % MT, s = (e >> 6) / MT;
            const int i = m * 16 + (l & 15), c = 4 * (s0 + s) + (l >> 4);
            T v = T(0);
            const u32 om = rinfo[i] & 0xffffu;
            if (om != 0u && c <= nsxA) {
                const int r = __builtin_ctz(om), t = 31 - __builtin_clz(om);
                if (c < nsxA) {
                    const u32 d = conn[c];
                    v = eri[(i64)(sx_r(d) * n + sx_s(d)) * n2 + r * n + t];
                    if (sx_neg(d)) v = -v;
                } else {
                    v = dw[Ia * nwd + r * n + t];
                }
            }
            Wl[e] = v;
        }
        // this lane's source rows: connection 4 (s0 + s) + g (rows past the last connection
        // read row Ia against zero W)
        u32 jo[KSP];
#pragma unroll
        for (int s = 0; s < KSP; ++s) {
            const int c = 4 * (s0 + s) + g;
            const u32 Ja = c < nsxA ? sx_J(conn[c]) : (u32)Ia;
            jo[s] = Ja * (u32)NB;
        }
        __syncthreads();

        // B operands and string of a tile, loaded unconditionally with clamped indices (a
        // load on one path only would make the compiler drain every outstanding load)
        auto load_tile = [&](T (&b)[KSP], u32 &jb, int nt) {
            const int ntc = nt < ntiles ? nt : ntiles - 1;
            int Jb = ntc * 16 + col;
            Jb = Jb < NB ? Jb : NB - 1;
#pragma unroll
            for (int s = 0; s < KSP; ++s) b[s] = xt[(size_t)((jo[s] + (u32)Jb) * (u32)nvst + (u32)v0)];
            jb = (u32)strsB[Jb];
        };
        T bq[KSP];
        u32 jbq;
        load_tile(bq, jbq, wave);
        for (int nt = wave; nt < ntiles; nt += NWV) {
            T b[KSP];
#pragma unroll
            for (int s = 0; s < KSP; ++s) b[s] = bq[s];
            const u32 jb = jbq;
            load_tile(bq, jbq, nt + NWV);
            V4 acc[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m) acc[m] = V4{T(0), T(0), T(0), T(0)};
            // A operands of the next k-step are read while this one's MFMAs run
            T a0[MT], a1[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m) a0[m] = Wl[m * 64 + lane];
#pragma unroll
            for (int s = 0; s < KSP; ++s) {
                T(&ac)[MT] = (s & 1) ? a1 : a0;
                T(&an)[MT] = (s & 1) ? a0 : a1;
                if (s + 1 < KSP) {
#pragma unroll
                    for (int m = 0; m < MT; ++m) an[m] = Wl[((s + 1) * MT + m) * 64 + lane];
                }
#pragma unroll
                for (int m = 0; m < MT; ++m) acc[m] = Ops::mma(ac[m], b[s], acc[m]);
            }
            // scatter.  Colex rank of a target Ib = Jb ^ om: rank = rk[lo(Ib)] + rk[256 +
            // 256 popc(lo(Ib)) + hi(Ib)]; a pair inside one byte leaves the other byte's term
            // as it is for Jb.  Two row tiles at a time, every target rank is read first
            // (branch-free, so the reads are issued together), then the adds; the sign goes
            // into the value.
            const int Jb = nt * 16 + col;
            const bool live = Jb < NB;
            const u32 jlo = jb & 255u, jbase = 256u + ((u32)__popc(jlo) << 8);
            const u32 RL = rk[jlo], RH = rk[jbase + (jb >> 8)];
            double dsum = 0.0;
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (jb & row_info(0, v)) dsum += (double)acc[0][v];
#pragma unroll
            for (int m0 = 1; m0 < MT; m0 += 2) {
                u32 tg[2][4];
#pragma unroll
                for (int mm = 0; mm < 2; ++mm)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int m = m0 + mm;
                        const u32 ri = row_info(m, v), om = ri & 0xffffu, x = jb & om;
                        const u32 ib = jb ^ om;
                        u32 rank;
                        if (m <= 2) {
                            rank = (u32)rk[ib & 255u] + RH;
                        } else if (m <= 4) {
                            rank = RL + (u32)rk[jbase + (ib >> 8)];
                        } else {
                            const u32 l2 = ib & 255u;
                            rank = (u32)rk[l2] + (u32)rk[256u + ((u32)__popc(l2) << 8) + (ib >> 8)];
                        }
                        acc[m][v] = flip_sign(acc[m][v], (u32)(__popc(jb & (ri >> 16)) & 1) << 31);
                        u32 tv = (live && x != 0u && x != om) ? rank : 0xffffffffu;
                        asm volatile("" : "+v"(tv));  // test the value itself at the add
                        tg[mm][v] = tv;
                    }
#pragma unroll
                for (int mm = 0; mm < 2; ++mm)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        if (tg[mm][v] != 0xffffffffu) atomicAdd(&S[tg[mm][v]], (double)acc[m0 + mm][v]);
            }
            if (live) atomicAdd(&S[Jb], dsum);
        }
        __syncthreads();  // W is refilled by the next pass
    }
    for (int t = tid; t < NB; t += BD) {
        T &o = s1[(size_t)((Ia * NB + t) * nvst + v0)];
        o = (T)((double)o + S[t]);
    }
}
