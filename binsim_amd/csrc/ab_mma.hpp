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
// W sits in LDS in MFMA-operand order for KSP k-steps (4 connections each) at a time: the A
// operand of tile (k-step s, row tile m) for lane l is entry (s * MT + m) * 64 + l, a
// conflict-free 8- or 4-byte read.
//
// Two forms.  Whole row (FP32): the LDS holds all NB destination elements and W (39 KB at 16
// orbitals), one launch.  Class split (FP64, CLS): the 78 KB of FP64 W do not fit beside a
// whole 103 KB FP64 row, so the destinations are split by the top orbital tb = n - 1 into
// class N (tb empty: ranks [0, NN) in colex order) and class T (the rest), one launch per
// class with only that class in LDS.  A target stays in the class of its source unless the
// pair moves an electron onto or off tb, so the launch for class K computes the 8 row tiles of
// the pairs without tb (and r = t = tb) on the class-K sources, and the one tile of pairs
// (r, tb) on the sources of the other class: every element of Y is computed once over both
// launches.
//
// Operand layout of v_mfma_{f64,f32}_16x16x4 on gfx950 (measured): lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; result element v of lane l is D[4 (l / 16) + v][l % 16] for f32 and
// D[4 v + l / 16][l % 16] for f64.
#pragma once

// colex rank tables: low byte (256), (low-byte count, high byte) (9 x 256)
constexpr int AB_MMA_RANKS = 256 + 9 * 256;

// waves per workgroup (one workgroup per CU: the LDS holds the FP64 destination elements): 16
// for both forms (128 / 116 VGPRs).  Measured at 16 orbitals: FP64 with 8 waves (row masks in
// registers, B one tile ahead: 256 VGPRs) 1.4% slower, 12 waves 0.5% slower; FP32 with 12
// waves 2% slower, 8 waves 6% slower.
constexpr int AB_MMA_NWV64 = 16, AB_MMA_NWV32 = 16;

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

// LDS bytes of k_ab_mma with NS destination elements in LDS
template <typename T, int MT, int KSP, int NWV>
constexpr size_t ab_mma_lds(i64 NS) {
    return (size_t)NS * sizeof(double) + (size_t)KSP * MT * 64 * sizeof(T) + MT * 16 * 4 +
           AB_MMA_RANKS * 2 + AB_SRC_CONN * 4 + NWV * 64 * sizeof(double);
}

// Orbital pair (r <= t) of W row i, or r = t = -1 for a row that never applies.
//   whole row: tile 0 the rows r = t, tiles 1-2 the pairs inside the low byte (r < t < 8),
//     tiles 3-4 inside the high byte (8 <= r < t), tiles 5-8 across (r < 8 <= t, row
//     80 + 8 r + t - 8);
//   class split (tb = n - 1): tile 0 the rows r = t, tiles 1-2 the low-byte pairs with t < tb,
//     tiles 3-7 the pairs inside orbitals 8..14, then those across (r < 8 <= t <= 14), each
//     kept when t < tb, tile 8 the pairs (r, tb), r < tb.
// The tile of a row fixes how its target rank is formed: with one table read when the pair
// lies inside the low byte (tiles 1-2) or (whole row) inside the high byte (tiles 3-4).
template <bool CLS>
__device__ inline void ab_mma_pair(int i, int n, int &r, int &t) {
    // e-th pair a < b of the w orbitals from `base`, e < w (w - 1) / 2
    auto in_run = [](int e, int base, int w, int &a, int &b) {
        int p = 0;
        while (e >= w - 1 - p) {
            e -= w - 1 - p;
            ++p;
        }
        a = base + p;
        b = a + 1 + e;
    };
    const int tb = n - 1;
    r = t = -1;
    if (i < 16) {
        r = t = i;
    } else if (i < 48) {
        if (i - 16 < 28) in_run(i - 16, 0, 8, r, t);
    } else if (!CLS) {
        if (i < 80) {
            if (i - 48 < 28) in_run(i - 48, 8, 8, r, t);
        } else {
            r = (i - 80) >> 3;
            t = 8 + ((i - 80) & 7);
        }
    } else if (i < 128) {
        const int e = i - 48;
        if (e < 21) {
            in_run(e, 8, 7, r, t);
        } else if (e < 77) {
            r = (e - 21) / 7;
            t = 8 + (e - 21) % 7;
        }
    } else {
        r = i - 128;
        t = tb;
    }
    bool ok = r >= 0 && t < n;
    if (CLS && ok && i >= 16) ok = i < 128 ? t < tb : r < tb;
    if (!ok) r = t = -1;
}

// nks = k-steps of all connections, ceil((nsxA + 1) / 4); dw rows are nwd entries wide.
// CLS: destination class cls (0: ranks [0, NN), 1: [NN, NB)); otherwise the whole row.
template <typename T, int MT, int KSP, int NWV, bool CLS>
__global__ __launch_bounds__(NWV * 64) __attribute__((amdgpu_waves_per_eu(NWV / 4, NWV / 4)))
void k_ab_mma(const T *__restrict__ xt, T *__restrict__ s1, int nvst, int v0, const u32 *__restrict__ sxA,
              const u64 *__restrict__ strsB, const T *__restrict__ eri, const T *__restrict__ dw, int nwd, int n,
              i64 NA, int NB, int nsxA, int nks, int cls, int NN) {
    using Ops = AbMmaOps<T>;
    using V4 = typename Ops::V4;
    constexpr int BD = NWV * 64;
    static_assert(MT == 9, "row layout: 9 tiles of 16");
    // own columns and destinations [o0, o0 + NK); other columns [c0, c0 + NC) (CLS only)
    const int o0 = CLS && cls ? NN : 0, NK = CLS ? (cls ? NB - NN : NN) : NB;
    const int c0 = CLS && !cls ? NN : 0, NC = CLS ? NB - NK : 0;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *S = reinterpret_cast<double *>(smem_raw);                       // destinations
    T *Wl = reinterpret_cast<T *>(smem_raw + (size_t)NK * sizeof(double));  // KSP x MT x 64
    u32 *rinfo = reinterpret_cast<u32 *>(Wl + KSP * MT * 64);               // per W row
    unsigned short *rk = reinterpret_cast<unsigned short *>(rinfo + MT * 16);
    u32 *conn = reinterpret_cast<u32 *>(rk + AB_MMA_RANKS);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // per-thread slot that takes the adds of the elements that do not apply (every element is
    // added, unconditionally: no branches in the scatter; 1% faster than masked adds)
    const u32 dmy = (u32)(reinterpret_cast<double *>(conn + AB_SRC_CONN) - S) + (u32)tid;
    const int g = lane >> 4, col = lane & 15;
    const i64 per = (NA + 7) / 8;
    const i64 Ia = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (Ia >= NA) return;
    const int n2 = n * n;

    for (int t = tid; t < nsxA; t += BD) conn[t] = sxA[(i64)t * NA + Ia];
    // a row's entry: its orbital mask (low half) and the mask of the orbitals strictly between
    // r and t (high half); 0 for a row that never applies
    for (int i = tid; i < MT * 16; i += BD) {
        int r, t;
        ab_mma_pair<CLS>(i, n, r, t);
        u32 v = 0;
        if (r >= 0) v = r == t ? 1u << r : (1u << r) | (1u << t) | ((((1u << t) - (2u << r)) & 0xffffu) << 16);
        rinfo[i] = v;
    }
    // colex ranks rank(s) = rk[s & 255] + rk[256 + 256 * popc(s & 255) + (s >> 8)], from the
    // binomials C(a, b), a < 16, b <= 16, staged at the start of the (not yet used)
    // destination and W area
    u32 *bn = reinterpret_cast<u32 *>(S);
    for (int e = tid; e < 16 * 17; e += BD) bn[e] = ab_mma_binom(e / 17, e % 17);
    __syncthreads();
    for (int e = tid; e < AB_MMA_RANKS; e += BD) {
        const int hi = e >= 256;
        const int k0 = hi ? (e - 256) >> 8 : 0, bits = hi ? (e - 256) & 255 : e, base = hi ? 8 : 0;
        u32 v = 0;
        int j = k0;
        for (int q = 0; q < 8; ++q)
            if ((bits >> q) & 1) {
                v += j + 1 <= 16 ? bn[(base + q) * 17 + j + 1] : 0u;
                ++j;
            }
        rk[e] = (unsigned short)v;
    }
    __syncthreads();
    for (int t = tid; t < NK; t += BD) S[t] = 0.0;
    __syncthreads();

    // this lane's W row of result element v in row tile m
    auto row_info = [&](int m, int v) -> u32 { return rinfo[m * 16 + Ops::row(g, v)]; };
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

        // the column tiles of [b0, b0 + nb) on the row tiles [M0, M0 + NM)
        auto run = [&](auto m0_c, auto nm_c, int b0, int nb) {
            constexpr int M0 = decltype(m0_c)::value, NM = decltype(nm_c)::value;
            const int ntiles = (nb + 15) >> 4;
            if (ntiles == 0) return;
            // B operands and string of a tile, loaded unconditionally with clamped indices (a
            // load on one path only would make the compiler drain every outstanding load)
            auto load_tile = [&](T (&b)[KSP], u32 &jb, int nt) {
                const int ntc = nt < ntiles ? nt : ntiles - 1;
                const int Jl = ntc * 16 + col;
                const int Jb = b0 + (Jl < nb ? Jl : nb - 1);
#pragma unroll
                for (int s = 0; s < KSP; ++s) b[s] = xt[(size_t)((jo[s] + (u32)Jb) * (u32)nvst + (u32)v0)];
                jb = (u32)strsB[Jb];
            };
            // FP32: B operands loaded one tile ahead; FP64 (more registers per operand) loads
            // them at the tile, the other waves covering the latency
            constexpr bool PF = sizeof(T) == 4;
            T bq[PF ? KSP : 1];
            u32 jbq = 0;
            if constexpr (PF) load_tile(bq, jbq, wave);
            for (int nt = wave; nt < ntiles; nt += NWV) {
                T b[KSP];
                u32 jb;
                if constexpr (PF) {
#pragma unroll
                    for (int s = 0; s < KSP; ++s) b[s] = bq[s % (PF ? KSP : 1)];
                    jb = jbq;
                    load_tile(bq, jbq, nt + NWV);
                } else {
                    load_tile(b, jb, nt);
                }
                V4 acc[NM];
#pragma unroll
                for (int m = 0; m < NM; ++m) acc[m] = V4{T(0), T(0), T(0), T(0)};
                // A operands of the next k-step are read while this one's MFMAs run
                T a0[NM], a1[NM];
#pragma unroll
                for (int m = 0; m < NM; ++m) a0[m] = Wl[(M0 + m) * 64 + lane];
#pragma unroll
                for (int s = 0; s < KSP; ++s) {
                    T(&ac)[NM] = (s & 1) ? a1 : a0;
                    T(&an)[NM] = (s & 1) ? a0 : a1;
                    if (s + 1 < KSP) {
#pragma unroll
                        for (int m = 0; m < NM; ++m) an[m] = Wl[((s + 1) * MT + M0 + m) * 64 + lane];
                    }
#pragma unroll
                    for (int m = 0; m < NM; ++m) acc[m] = Ops::mma(ac[m], b[s], acc[m]);
                }
                // scatter.  Colex rank of a target Ib = Jb ^ om: rank = rk[lo(Ib)] + rk[256 +
                // 256 popc(lo(Ib)) + hi(Ib)]; a pair inside one byte leaves the other byte's
                // term as it is for Jb.  Two row tiles at a time, every target rank is read
                // first, then the adds -- an element that does not apply goes to the thread's
                // spare slot; the sign goes into the value.
                const int Jl = nt * 16 + col;
                const bool live = Jl < nb;
                const u32 jlo = jb & 255u, jbase = 256u + ((u32)__popc(jlo) << 8);
                const u32 RL = rk[jlo], RH = rk[jbase + (jb >> 8)];
                if constexpr (M0 == 0) {
                    double dsum = 0.0;
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        if (jb & row_info(0, v)) dsum += (double)acc[0][v];
                    if (live) atomicAdd(&S[b0 + Jl - o0], dsum);
                }
                constexpr int F = M0 == 0 ? 1 : 0;  // first off-diagonal tile (local index)
#pragma unroll
                for (int q0 = F; q0 < NM; q0 += 2) {
                    u32 tg[2][4];
#pragma unroll
                    for (int qq = 0; qq < 2; ++qq)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            if (q0 + qq >= NM) continue;
                            const int m = M0 + q0 + qq;  // row tile
                            const u32 ri = row_info(m, v), om = ri & 0xffffu, x = jb & om;
                            const u32 ib = jb ^ om;
                            u32 rank;
                            if (m <= 2) {
                                rank = (u32)rk[ib & 255u] + RH;
                            } else if (!CLS && m <= 4) {
                                rank = RL + (u32)rk[jbase + (ib >> 8)];
                            } else {
                                const u32 l2 = ib & 255u;
                                rank = (u32)rk[l2] + (u32)rk[256u + ((u32)__popc(l2) << 8) + (ib >> 8)];
                            }
                            acc[q0 + qq][v] =
                                flip_sign(acc[q0 + qq][v], (u32)(__popc(jb & (ri >> 16)) & 1) << 31);
                            tg[qq][v] = (live && __popc(x) == 1) ? rank - (u32)o0 : dmy;
                        }
#pragma unroll
                    for (int qq = 0; qq < 2; ++qq)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            if (q0 + qq < NM) atomicAdd(&S[tg[qq][v]], (double)acc[q0 + qq][v]);
                }
            }
        };
        using I0 = std::integral_constant<int, 0>;
        using I1 = std::integral_constant<int, 1>;
        using I8 = std::integral_constant<int, 8>;
        using I9 = std::integral_constant<int, 9>;
        if constexpr (CLS) {
            run(I0{}, I8{}, o0, NK);
            run(I8{}, I1{}, c0, NC);
        } else {
            run(I0{}, I9{}, 0, NB);
        }
        __syncthreads();  // W is refilled by the next pass
    }
    for (int t = tid; t < NK; t += BD) {
        T &o = s1[(size_t)((Ia * NB + o0 + t) * nvst + v0)];
        o = (T)((double)o + S[t]);
    }
}
