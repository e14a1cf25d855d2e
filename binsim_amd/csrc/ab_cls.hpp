// Alpha-beta part of the sigma build for pairs of FP32 vectors, class-split source-lane form
// (included by sigma.hip after ab_src.hpp).
//
// The beta strings of k electrons in n orbitals fall into two classes by the occupation of the
// top orbital t0 = n - 1; in colex order both are contiguous rank ranges:
//   class N (t0 empty):    [0, C(n-1, k)),          C(n-1, k) strings
//   class T (t0 occupied): [C(n-1, k), C(n, k)),    C(n-1, k-1) strings
// A single excitation r -> t of a beta string stays in its class unless it moves t0.  A
// workgroup owns destination alpha row Ia and ONE class C of it.  Its LDS holds the class's
// destination elements in FP64 for NVS vectors (at 16 orbitals / 8 electrons: 6,435 elements x
// 2 vectors = 103 KB, as much as one whole FP64 row), so the scatter uses the FP64 LDS atomic
// add, about 7x the throughput of the FP32 one on this GPU.  Per chunk of 512 and alpha
// connection, thread jl handles
//   its own source   (class C, rank base_C + chunk*512 + jl): the slots that keep t0 as it is
//                    (class T: r != t0, KO = k-1 occupied x VO = n-k virtual; class N: t != t0,
//                    KO = k x VO = n-k-1), and the diagonal term;
//   a foreign source (the other class, same position): the KF slots that move t0 into or out
//                    of it, whose targets land in C (class T: r -> t0 for the k occupied r;
//                    class N: t0 -> t for the n-k virtual t).
// At 16 orbitals / 8 electrons that is 56 + 8 = 64 slots per thread and round, the same as the
// whole-row kernel, over half the rounds per vector pair (two workgroups per alpha row).
// Everything else follows k_ab_src: the integral row W[r*n + t] = s_a (pq|rt) of the round's
// alpha connection replicated R times in LDS (thread l reads copy l % R), source elements
// loaded XPF connections ahead, accumulators in registers across the connections of a chunk,
// computed scatter targets (colex ranks from the binomial table; a target outside the class,
// which the rank arithmetic rules out, would be dropped).  (pq|rt) = (pq|tr), so the foreign
// slots of class N read W[t0*n + t].
#pragma once

// class-split kernel for class CT (true: T, false: N): NVS vectors (interleaved in xt / s1 with
// stride nvst), a KO x VO grid of own slots and KF foreign slots, of which the first ko x vo and
// kf are used (EX: all of them), R copies of the 256-entry integral row, 512 threads.  Unused
// slots read valid LDS words into accumulators that are never scattered.
template <int NVS, int KO, int VO, int KF, int R, bool CT, bool EX>
__global__ __launch_bounds__(AB_SRC_NJ) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_ab_cls(const float *__restrict__ xt, float *__restrict__ s1, int nvst, int v0,
              const u32 *__restrict__ sxA, const u64 *__restrict__ strsB, const float *__restrict__ eri,
              const float *__restrict__ dw, int n, i64 NA, int NB, int nsxA, int kB, int ko_arg,
              int vo_arg, int kf_arg) {
    constexpr int NJ = AB_SRC_NJ, NW = 256, BD = NJ;
    constexpr int XPF = 2;  // (3 as in k_ab_src spills here)
    constexpr int NWV = BD / 64, EPW = NW / NWV, LPE = 64 / EPW, CB = 4, NBLK = R / CB, BPL = NBLK / LPE;
    static_assert(NWV * EPW == NW && LPE * EPW == 64 && NBLK % LPE == 0 && BPL >= 1, "fill layout");
    typedef float TV __attribute__((ext_vector_type(CB)));
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n2 = n * n;
    float *Wr = reinterpret_cast<float *>(smem_raw);  // NW * R replicated integral row
    float *Z = Wr + NW * R;                            // 16 zeros
    double *S = reinterpret_cast<double *>(Z + 16);    // class elements x NVS
    constexpr int BW = 8 + 2;
    // class sizes: C(n-1, k) strings without t0, C(n-1, k-1) with it
    const int t0 = n - 1;
    const int tid = threadIdx.x;
    const i64 per = (NA + 7) / 8;
    constexpr bool clsT = CT;
    const int ko = EX ? KO : ko_arg, vo = EX ? VO : vo_arg, kf = EX ? KF : kf_arg;
    const i64 Ia = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (Ia >= NA) return;
    // C(n-1, k) from the string count: NB = C(n, k), C(n-1, k) = NB (n - k) / n
    const int nN = NB * (n - kB) / n, nT = NB - nN;
    const int own_base = clsT ? nN : 0, own_cnt = clsT ? nT : nN;
    const int for_base = clsT ? 0 : nN, for_cnt = clsT ? nN : nT;
    int *binom = reinterpret_cast<int *>(S + own_cnt * NVS);
    u32 *conn = reinterpret_cast<u32 *>(binom + AB_SRC_BINOM);
    for (int t = tid; t < own_cnt * NVS; t += BD) S[t] = 0.0;
    if (tid < 16) Z[tid] = 0.0f;
    for (int t = tid; t < n * BW; t += BD) {
        const int p = t / BW, j = t - p * BW;
        u64 c = j <= p ? 1ull : 0ull;
        for (int i = 0; i < j && c; ++i) c = c * (u64)(p - i) / (u64)(i + 1);
        binom[t] = (int)c;
    }
    for (int t = tid; t < nsxA; t += BD) conn[t] = sxA[(i64)t * NA + Ia];
    __syncthreads();

    const int nca = nsxA + 1;
    const int nchunks = ((own_cnt > for_cnt ? own_cnt : for_cnt) + NJ - 1) / NJ;
    const int nit = nchunks * nca;
    const u32 copy_off = (u32)((tid & (R - 1)) * sizeof(float));
    const int fe = (tid >> 6) * EPW + (tid & 63) / LPE;
    u32 foff[BPL];
#pragma unroll
    for (int q = 0; q < BPL; ++q) {
        const int b = ((tid & 63) % LPE * BPL + q + (tid & 63) / LPE) % NBLK;
        foff[q] = (u32)((fe * R + b * CB) * sizeof(float));
    }
    auto store_w = [&](float v) {
        const TV vv = v;
#pragma unroll
        for (int q = 0; q < BPL; ++q) *reinterpret_cast<TV *>(smem_raw + foff[q]) = vv;
    };
    // entries past n2 run into the next row or the zero padding after the table and land in
    // Wr slots no lane reads
    auto load_w = [&](float &pre, bool &neg, int ca) {
        const bool single = ca < nsxA;
        const u32 e = conn[single ? ca : 0];
        const float *er = single ? eri + (i64)(sx_r(e) * n + sx_s(e)) * n2 : dw + Ia * NW;
        neg = single && sx_neg(e);
        pre = er[fe];
    };
    // own and foreign source elements of round (chunk, ca); clamped to valid elements
    auto load_x = [&](float (&xo)[NVS], float (&xf)[NVS], int chunk, int ca) {
        const int c = chunk < nchunks ? chunk : nchunks - 1;
        int jo = c * NJ + tid, jf = jo;
        jo = own_base + (jo < own_cnt ? jo : own_cnt - 1);
        jf = for_base + (jf < for_cnt ? jf : for_cnt - 1);
        const i64 Ja = ca < nsxA ? (i64)sx_J(conn[ca < nsxA ? ca : 0]) : Ia;
#pragma unroll
        for (int v = 0; v < NVS; ++v) {
            xo[v] = xt[(Ja * NB + jo) * nvst + v0 + v];
            xf[v] = xt[(Ja * NB + jf) * nvst + v0 + v];
        }
    };
    float pre;
    bool pneg = false;
    load_w(pre, pneg, 0);
    float xqo[XPF][NVS], xqf[XPF][NVS];
    int pchunk = 0;
    int pca = 0;
    auto advance = [&](int &c, int &a) {
        if (++a == nca) {
            a = 0;
            ++c;
        }
    };
#pragma unroll
    for (int d = 0; d < XPF; ++d) {
        load_x(xqo[d], xqf[d], pchunk, pca);
        advance(pchunk, pca);
    }
    float acc[KO][VO][NVS], accf[KF][NVS], accd[NVS];
    // the orbitals of the slots, packed 4 bits each (n <= 16): own occupied po and virtual pv,
    // the diagonal's pd, the foreign slots' pf; their LDS offsets are formed per round (held
    // across rounds they would push the accumulators out of the register file)
    u32 po = 0, pv = 0, pd = 0, pf = 0;
    bool live_o = false, live_f = false;
    const u32 rowb = (u32)(n * R * sizeof(float)), colb = (u32)(R * sizeof(float));
    const u32 zero_off = (u32)(NW * R * sizeof(float));
    const u64 all = (n < 64) ? ((1ull << n) - 1) : ~0ull;
    int chunk = 0;
    int ca = 0;
    for (int it = 0; it < nit; ++it, advance(chunk, ca)) {
        if (ca == 0) {
            const int lo = chunk * NJ + tid;
            live_o = lo < own_cnt;
            live_f = lo < for_cnt;
            const u64 so = strsB[own_base + (live_o ? lo : own_cnt - 1)];
            const u64 sf = strsB[for_base + (live_f ? lo : for_cnt - 1)];
            // own slots: class T skips the last occupied orbital (t0), class N the last virtual
            u64 mo = so, mv = ~so & all;
            po = 0;
            pv = 0;
#pragma unroll
            for (int i = 0; i < KO; ++i) {
                const u32 p = mo ? (u32)__builtin_ctzll(mo) : 0u;
                mo &= mo - 1;
                po |= p << (4 * i);
            }
#pragma unroll
            for (int iv = 0; iv < VO; ++iv) {
                const u32 q = mv ? (u32)__builtin_ctzll(mv) : 0u;
                mv &= mv - 1;
                pv |= q << (4 * iv);
            }
            // diagonal: all k occupied orbitals of the own source
            u64 md = so;
            pd = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const u32 p = md ? (u32)__builtin_ctzll(md) : 0u;
                md &= md - 1;
                pd |= p << (4 * i);
            }
            // foreign slots: class T reads W[r*n + t0] for the occupied r of an N string, class
            // N reads W[t0*n + t] for the virtual t of a T string
            u64 mf = clsT ? sf : (~sf & all);
            pf = 0;
#pragma unroll
            for (int j = 0; j < KF; ++j) {
                const u32 p = mf ? (u32)__builtin_ctzll(mf) : 0u;
                mf &= mf - 1;
                pf |= p << (4 * j);
            }
#pragma unroll
            for (int i = 0; i < KO; ++i)
#pragma unroll
                for (int iv = 0; iv < VO; ++iv)
#pragma unroll
                    for (int v = 0; v < NVS; ++v) acc[i][iv][v] = 0.0f;
#pragma unroll
            for (int j = 0; j < KF; ++j)
#pragma unroll
                for (int v = 0; v < NVS; ++v) accf[j][v] = 0.0f;
#pragma unroll
            for (int v = 0; v < NVS; ++v) accd[v] = 0.0f;
        }
        __syncthreads();  // everyone is done reading the previous W
        store_w(pneg ? -pre : pre);
        __syncthreads();
        load_w(pre, pneg, ca + 1 == nca ? 0 : ca + 1);
        float xno[NVS], xnf[NVS];
        load_x(xno, xnf, pchunk, pca);
        advance(pchunk, pca);
        {
            // a zero read from LDS after this round's barrier makes the slot offsets opaque per
            // round, so the compiler forms the addresses in the loop instead of hoisting them
            const u32 lz = *reinterpret_cast<const u32 *>(smem_raw + zero_off);
            u32 oo[KO], vv[VO];
#pragma unroll
            for (int i = 0; i < KO; ++i) oo[i] = ((po >> (4 * i)) & 15u) * rowb + lz;
#pragma unroll
            for (int iv = 0; iv < VO; ++iv) vv[iv] = ((pv >> (4 * iv)) & 15u) * colb + copy_off;
            constexpr int PD = AB_SRC_PD < KO ? AB_SRC_PD : KO;
            float wv[PD + 1][VO];
#pragma unroll
            for (int d = 0; d < PD; ++d)
#pragma unroll
                for (int iv = 0; iv < VO; ++iv)
                    wv[d][iv] = *reinterpret_cast<const float *>(smem_raw + oo[d] + vv[iv]);
#pragma unroll
            for (int i = 0; i < KO; ++i) {
                if (i + PD < KO) {
#pragma unroll
                    for (int iv = 0; iv < VO; ++iv)
                        wv[(i + PD) % (PD + 1)][iv] = *reinterpret_cast<const float *>(smem_raw + oo[i + PD] + vv[iv]);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int iv = 0; iv < VO; ++iv)
#pragma unroll
                    for (int v = 0; v < NVS; ++v) acc[i][iv][v] += wv[i % (PD + 1)][iv] * xqo[0][v];
            }
            // foreign: class T entry r*n + t0, class N entry t0*n + t (r, t from pf)
            const u32 fmul = clsT ? rowb : colb, fadd = (clsT ? (u32)t0 * colb : (u32)t0 * rowb) + copy_off + lz;
            float wf[KF];
#pragma unroll
            for (int j = 0; j < KF; ++j)
                wf[j] = *reinterpret_cast<const float *>(smem_raw + ((pf >> (4 * j)) & 15u) * fmul + fadd);
            // diagonal entries p*n + p of the k occupied orbitals (slots past k read the zeros)
            float wd = 0.0f;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const u32 off = i < kB ? ((pd >> (4 * i)) & 15u) * (rowb + colb) + copy_off : zero_off;
                wd += *reinterpret_cast<const float *>(smem_raw + off + lz);
            }
#pragma unroll
            for (int j = 0; j < KF; ++j)
#pragma unroll
                for (int v = 0; v < NVS; ++v) accf[j][v] += wf[j] * xqf[0][v];
#pragma unroll
            for (int v = 0; v < NVS; ++v) accd[v] += wd * xqo[0][v];
        }
#pragma unroll
        for (int d = 0; d + 1 < XPF; ++d)
#pragma unroll
            for (int v = 0; v < NVS; ++v) {
                xqo[d][v] = xqo[d + 1][v];
                xqf[d][v] = xqf[d + 1][v];
            }
#pragma unroll
        for (int v = 0; v < NVS; ++v) {
            xqo[XPF - 1][v] = xno[v];
            xqf[XPF - 1][v] = xnf[v];
        }
        if (ca == nca - 1) {
            // scatter this chunk's sums into the class's LDS elements.  For a string with
            // occupied p_0 < .. < p_{k-1} (colex rank sum_j C(p_j, j+1)) and the slot moving
            // r = p_io to the iv-th virtual t, with m = t - iv occupied orbitals below t:
            //   t > r: rank' = rank - C(r, io+1) + C(t, m)   + A_m - A_{io+1}
            //   t < r: rank' = rank - C(r, io+1) + C(t, m+1) + B_io - B_m
            // A_j = sum_{i<j} C(p_i, i) - C(p_i, i+1), B_j = sum_{i<j} C(p_i, i+2) - C(p_i, i+1);
            // the sign is the parity of the occupied orbitals strictly between r and t.
            auto prefix = [&](u64 s, int (&pa)[8], int (&A)[9], int (&Bp)[9]) {
                u64 mo = s;
                A[0] = 0;
                Bp[0] = 0;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int p = mo ? __builtin_ctzll(mo) : 0;
                    mo &= mo - 1;
                    pa[j] = p;
                    const int c0 = binom[p * BW + j], c1 = binom[p * BW + j + 1], c2 = binom[p * BW + j + 2];
                    A[j + 1] = A[j] + (j < kB ? c0 - c1 : 0);
                    Bp[j + 1] = Bp[j] + (j < kB ? c2 - c1 : 0);
                }
            };
            const int Jo = own_base + (live_o ? chunk * NJ + tid : 0);
            const int Jf = for_base + (live_f ? chunk * NJ + tid : 0);
            if (live_o) {
                const u64 s = strsB[Jo];
                int pa[8], A[9], Bp[9];
                prefix(s, pa, A, Bp);
                u64 mv = ~s & all;
#pragma unroll
                for (int iv = 0; iv < VO; ++iv) {
                    const int t = mv ? __builtin_ctzll(mv) : 0;
                    mv &= mv - 1;
                    // m is in [0, k] for the used slots; unused ones (t = 0) are clamped
                    const int m = t - iv < 0 ? 0 : (t - iv > 8 ? 8 : t - iv);
                    const int Ct0 = binom[t * BW + m], Ct1 = binom[t * BW + m + 1];
#pragma unroll
                    for (int io = 0; io < KO; ++io) {
                        if (iv >= vo || io >= ko) continue;
                        const bool above = m > io;
                        const int r = pa[io];
                        const int cr = binom[r * BW + io + 1];
                        const int Ip = Jo - own_base - cr +
                                       (above ? Ct0 + A[m] - A[io + 1] : Ct1 + Bp[io] - Bp[m]);
                        const bool neg = (m + io + (above ? 1 : 0)) & 1;
                        if ((u32)Ip < (u32)own_cnt) {
#pragma unroll
                            for (int v = 0; v < NVS; ++v)
                                atomicAdd(&S[Ip * NVS + v], (double)(neg ? -acc[io][iv][v] : acc[io][iv][v]));
                        }
                    }
                }
#pragma unroll
                for (int v = 0; v < NVS; ++v) atomicAdd(&S[(Jo - own_base) * NVS + v], (double)accd[v]);
            }
            if (live_f) {
                const u64 s = strsB[Jf];
                int pa[8], A[9], Bp[9];
                prefix(s, pa, A, Bp);
                if (clsT) {
                    // N string, r = p_j -> t0 (above every occupied orbital: m = k)
                    const int cf = binom[t0 * BW + kB] + A[kB];
#pragma unroll
                    for (int j = 0; j < KF; ++j) {
                        if (j >= kf) continue;
                        const int cr = binom[pa[j] * BW + j + 1];
                        const int Ip = Jf - cr - A[j + 1] + cf - own_base;
                        const bool neg = (kB + j + 1) & 1;
                        if ((u32)Ip < (u32)own_cnt) {
#pragma unroll
                            for (int v = 0; v < NVS; ++v)
                                atomicAdd(&S[Ip * NVS + v], (double)(neg ? -accf[j][v] : accf[j][v]));
                        }
                    }
                } else {
                    // T string, r = t0 = p_{k-1} -> the j-th virtual t (below t0: m = t - j)
                    const int io = kB - 1;
                    const int cr = binom[t0 * BW + io + 1];
                    u64 mv = ~s & all;
#pragma unroll
                    for (int j = 0; j < KF; ++j) {
                        const int t = mv ? __builtin_ctzll(mv) : 0;
                        mv &= mv - 1;
                        if (j >= kf) continue;
                        const int m = t - j;
                        const int Ip = Jf - cr + binom[t * BW + m + 1] + Bp[io] - Bp[m] - own_base;
                        const bool neg = (m + io) & 1;
                        if ((u32)Ip < (u32)own_cnt) {
#pragma unroll
                            for (int v = 0; v < NVS; ++v)
                                atomicAdd(&S[Ip * NVS + v], (double)(neg ? -accf[j][v] : accf[j][v]));
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    __syncthreads();
    for (int t = tid; t < own_cnt * NVS; t += BD) {
        const int Ip = t / NVS;
        float &o = s1[(Ia * NB + own_base + Ip) * nvst + v0 + (t - Ip * NVS)];
        o = (float)((double)o + S[t]);
    }
}
