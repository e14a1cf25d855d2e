// Alpha-beta part of the sigma build, source-lane form (included by sigma.hip).
//
// Workgroup = destination alpha row Ia; every thread owns one SOURCE beta string Jb of the
// current chunk of 512 and a share (1/SPLIT) of its k*(n-k) single-excitation slots:
//
//   sigma~[Ia, Ib] += sum_{alpha conn. (Ja, pq, s_a)} sum_{slots (r in Jb, t not in Jb)}
//                     s_a (pq|rt) <Ib|E_tr|Jb> x~[Ja, Jb]          Ib = Jb - r + t
//                   + sum_{alpha conn.} (sum_{r in Jb} s_a (pq|rr)) x~[Ja, Jb]   (Ib = Jb)
//
// (the alpha connections are the nsxA single excitations Ja = Ia - p + q of Ia and Ia itself
// with W = sum_{p in Ia} (pp|..)).  The source element x~[Ja, Jb] is loaded once per alpha
// connection -- coalesced along the row, XPF connections ahead -- and reused from a register
// for all of the thread's slots; the accumulators acc[slot] stay in registers across the
// alpha connections and are scatter-added into an LDS copy of the destination row once per
// chunk.  Slot (io, iv) is orbital occ[io] annihilated / vir[iv] created; its target rank
// and sign are computed from the colex ranks (see the scatter), with <Ib|E_tr|Jb> the
// parity of the occupied orbitals of Jb between r and t, and (pq|rt) = (pq|tr).
//
// The integral row W[r*n + t] = s_a (pq|rt) is stored R times interleaved (entry e, copy c at
// e*R + c) and thread l reads copy l % R, so the lane-varying gathers W[occ*n + vir] of a
// group of R lanes never share a bank.  W of the next connection is prefetched into
// registers while the current one is consumed.  With `ordered` the contributions are added
// into the LDS row one wave at a time, in wave order, so the sum order -- and the FP result
// -- is reproducible; otherwise all waves add at once.  Fixed block size, so fill positions
// are compile-time offsets.
//
// The LDS row holds the destination elements [lo, lo + nbp) of row Ia; contributions to other
// elements are dropped.  A row too long for the LDS is done in parts, one launch per part
// (every part repeats the accumulation and keeps its own targets).
// Included inside namespace hipotf after sx_* and u32/u64/i64 are defined.
#pragma once

constexpr int AB_SRC_NJ = 512;   // source strings per chunk
// integral entries one W fill covers (template NW): 256 for n*n <= 256, 512 for n*n <= 512
constexpr int AB_SRC_XPF = 3;    // alpha connections of source elements in flight
constexpr int AB_SRC_BINOM = 192; // binomial table entries: n <= 18 orbitals x (KM + 2) <= 10
constexpr int AB_SRC_NTRI = 136;  // r <= t entries of a 16-orbital integral row (TRI layout)
constexpr int AB_SRC_CONN = 88;   // alpha connections per row: k (n - k) <= 81 for n <= 18
// slot rows of LDS gathers issued ahead of the FMAs (measured at 16 orbitals: 2 rows is no
// faster than 1, 3 is 2-4% slower)
constexpr int AB_SRC_PD = 1;
// keep slot addresses in VGPRs across rounds (NVS = 1, DB = 1 only): 1.5-3% faster there
constexpr bool AB_SRC_HOIST = true;

// word type of the LDS destination row: FP64 also for the single-vector FP32 kernel of up to 16
// orbitals (its 103 KB row fits beside the integral copies), whose scatter then uses the FP64
// LDS atomic add
template <typename T, int NVS, int NW>
using AbSrcRow = typename std::conditional<NW == 256 && NVS == 1, double, T>::type;

// TRI: the integral row is stored as its r <= t triangle, n(n+1)/2 entries ((pq|rt) = (pq|tr)),
// which leaves room for R = 32 copies of 8-byte words beside the FP64 destination row: lanes l
// and l + 16 then read different copies and the FP64 gathers are free of bank conflicts (with
// 16 copies of the full row, lanes l and l + 16 collide whenever their entries share a parity).
// The slot addresses are formed once per chunk and kept in VGPRs (single-vector FP64 form).
template <typename T, int NVS, int KM, int VM, int R, int SPLIT, int DB, int NW, bool TRI = false>
__global__ __launch_bounds__(AB_SRC_NJ *SPLIT) __attribute__((amdgpu_waves_per_eu(2 * SPLIT, 2 * SPLIT)))
void k_ab_src(const T *__restrict__ xt, T *__restrict__ s1, int nvst, int v0, const u32 *__restrict__ sxA,
              const u32 *__restrict__ sxB, const u64 *__restrict__ strsA, const u64 *__restrict__ strsB,
              const T *__restrict__ eri, const T *__restrict__ dw, int n, i64 NA, i64 NB, int nsxA, int kB,
              int vB, int ordered, i64 lo_arg, i64 nbp_arg) {
    constexpr int BD = AB_SRC_NJ * SPLIT;
    constexpr int KH = KM / SPLIT;                 // occupied-orbital slots per thread
    constexpr int XPF = AB_SRC_XPF;
    // only the FP64 512-entry form is launched in parts; every other form gets lo = 0, nbp = NB
    // and compiles the range checks of the scatter out
    constexpr bool PARTS = NW == 512 && sizeof(T) == 8;
    const i64 lo = PARTS ? lo_arg : 0, nbp = PARTS ? nbp_arg : NB;
    // W fill: wave w loads entries [w*EPW, (w+1)*EPW) of the row, LPE lanes per entry (one
    // coalesced load per thread), and each lane stores BPL 16-byte blocks of CB copies of its
    // entry.  The block order is rotated with the entry so that the 64 lanes of one store
    // instruction spread evenly over the LDS banks.
    constexpr int NWV = BD / 64;                   // waves
    constexpr int EPW = NW / NWV;                  // entries per wave
    constexpr int LPE = 64 / EPW;                  // lanes per entry
    constexpr int CB = 16 / (int)sizeof(T);        // copies per 16-byte block
    constexpr int NBLK = R / CB;                   // blocks per entry
    constexpr int NE = TRI ? AB_SRC_NTRI : NW;     // entries stored per replicated row
    static_assert(!TRI || (DB == 1 && NVS == 1 && SPLIT == 1 && NW == 256), "triangular row: single-vector form");
    constexpr int BPL = NBLK / LPE;                // blocks stored per lane
    static_assert(KM % SPLIT == 0 && NWV * EPW == NW && LPE * EPW == 64 && NBLK % LPE == 0 &&
                  BPL >= 1, "fill layout");
    typedef T TV __attribute__((ext_vector_type(CB)));
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n2 = n * n;
    // DB = 2: two replicated rows, the next connection's filled while the current one is read
    // (one barrier per connection instead of two)
    T *Wr = reinterpret_cast<T *>(smem_raw);       // DB x NE * R replicated integral rows
    T *Z = Wr + DB * NE * R;                       // 16 zeros (reads of unused slots)
    using ST = AbSrcRow<T, NVS, NW>;
    ST *S = reinterpret_cast<ST *>(Z + 16);        // nbp * NVS destination elements lo..
    constexpr int BW = KM + 2;
    int *binom = reinterpret_cast<int *>(S + nbp * NVS);  // C(p, j), p < n, j < BW
    // the row's alpha connections (single-excitation table entries of Ia), staged so the round
    // loop issues no scalar memory loads: those share the LDS wait counter and would force
    // every wait on the slot gathers to drain all of them
    u32 *conn = reinterpret_cast<u32 *>(binom + AB_SRC_BINOM);
    const int tid = threadIdx.x;
    const int jl = tid & (AB_SRC_NJ - 1);          // source string within the chunk
    const int half = SPLIT == 1 ? 0 : __builtin_amdgcn_readfirstlane(tid / AB_SRC_NJ);
    const i64 per = (NA + 7) / 8;
    const i64 Ia = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (Ia >= NA) return;
    (void)strsA;
    for (i64 t = tid; t < nbp * NVS; t += BD) S[t] = ST(0);
    if (tid < 16) Z[tid] = T(0);
    for (int t = tid; t < n * BW; t += BD) {
        const int p = t / BW, j = t - p * BW;
        u64 c = j <= p ? 1ull : 0ull;
        for (int i = 0; i < j && c; ++i) c = c * (u64)(p - i) / (u64)(i + 1);
        binom[t] = (int)c;
    }
    for (int t = tid; t < nsxA; t += BD) conn[t] = sxA[(i64)t * NA + Ia];
    __syncthreads();

    const int nca = nsxA + 1;
    const i64 nchunks = (NB + AB_SRC_NJ - 1) / AB_SRC_NJ;
    const i64 nit = nchunks * nca;
    const u32 copy_off = (u32)((tid & (R - 1)) * sizeof(T));
    const int fe = (tid >> 6) * EPW + (tid & 63) / LPE;  // entry this thread loads and stores
    u32 foff[BPL];                                        // byte offsets of its blocks in a row
#pragma unroll
    for (int q = 0; q < BPL; ++q) {
        const int b = ((tid & 63) % LPE * BPL + q + (tid & 63) / LPE) % NBLK;
        foff[q] = (u32)((fe * R + b * CB) * sizeof(T));
    }
    // TRI: fill entry fe < n(n+1)/2 is (a, b), a <= b, in row-major triangle order; it loads
    // row element a*n + b, and threads with fe past the triangle store nothing
    const int ntri = n * (n + 1) / 2;
    int fsrc = fe;
    if (TRI) {
        int a = 0, start = 0;
        while (a < n - 1 && start + (n - a) <= fe) {
            start += n - a;
            ++a;
        }
        fsrc = fe < ntri ? a * n + a + (fe - start) : 0;
    }
    auto store_w = [&](T *W, T v) {
        const TV vv = v;
        if (TRI && fe >= ntri) return;
#pragma unroll
        for (int q = 0; q < BPL; ++q) *reinterpret_cast<TV *>(reinterpret_cast<unsigned char *>(W) + foff[q]) = vv;
    };

    // Every global load of the round loop is unconditional, with clamped in-bounds indices:
    // a load issued on one path only makes the compiler wait for ALL outstanding loads
    // (vmcnt(0)) where the paths meet, which would expose each prefetch's full latency.
    //
    // W values of connection ca (the sign is applied when they are stored, so the loads stay
    // in flight through the compute phase): row (pq|..) of a single excitation, or the
    // diagonal row sum_{p in Ia} (pp|..) from dw.  Entries past n2 run into the next row or
    // the zero padding after the table and land in Wr slots no lane reads.
    auto load_w = [&](T &pre, bool &neg, int ca) {
        const bool single = ca < nsxA;
        const u32 e = conn[single ? ca : 0];
        const T *er = single ? eri + (i64)(sx_r(e) * n + sx_s(e)) * n2 : dw + Ia * NW;
        neg = single && sx_neg(e);
        pre = er[fsrc];
    };
    // source elements of the round (chunk, connection ca); rounds are tracked by incremental
    // counters: a 64-bit division here would be expanded into vector code in the FMA phase.
    // Lanes past the last source string (never scattered) and rounds past the last one
    // (never consumed) read a valid element instead.
    auto load_x = [&](T (&xv)[NVS], i64 chunk, int ca) {
        i64 Jb = (chunk < nchunks ? chunk : nchunks - 1) * AB_SRC_NJ + jl;
        Jb = Jb < NB ? Jb : NB - 1;
        const i64 Ja = ca < nsxA ? (i64)sx_J(conn[ca < nsxA ? ca : 0]) : Ia;
#pragma unroll
        for (int v = 0; v < NVS; ++v) xv[v] = xt[(Ja * NB + Jb) * nvst + v0 + v];
    };

    T pre;
    bool pneg = false;
    load_w(pre, pneg, 0);
    int wca = 0;  // connection whose W values are in pre
    if (DB == 2) {
        // round 0's row goes into buffer 0 now; pre then carries round 1's
        store_w(Wr, pneg ? -pre : pre);
        wca = 1 % nca;
        load_w(pre, pneg, wca);
    }
    T xq[XPF][NVS];
    i64 pchunk = 0;  // round whose source elements are loaded next
    int pca = 0;
    auto advance = [&](i64 &c, int &a) {
        if (++a == nca) {
            a = 0;
            ++c;
        }
    };
#pragma unroll
    for (int d = 0; d < XPF; ++d) {
#pragma unroll
        for (int v = 0; v < NVS; ++v) xq[d][v] = T(0);
        load_x(xq[d], pchunk, pca);
        advance(pchunk, pca);
    }
    T acc[KH][VM][NVS], accd[NVS];
    u32 oo[KH], vv[VM], dd[KH];
    i64 Jb = 0;
    bool live = false;
    // slots beyond (kB, vB) of the instantiated KM x VM grid read valid LDS words into
    // accumulators that are never scattered; their diagonal terms read the zeros Z
    const u32 zero_off = (u32)(DB * NE * R * sizeof(T));
    // TRI: byte offset of entry (a, b), a <= b, copy copy_off
    auto tri_off = [&](u32 a, u32 b) {
        return (u32)((a * (u32)n - (a * (a - 1u)) / 2u + (b - a)) * (u32)(R * sizeof(T))) + copy_off;
    };
    // TRI: the slot addresses, formed once per chunk, two 16-bit offsets per register (the
    // triangle with its copies spans 34,816 bytes)
    constexpr int VP = TRI ? (VM + 1) / 2 : 1;
    u32 sa[TRI ? KH : 1][VP];
    auto slot_addr = [&](int i, int iv) -> u32 {
        const u32 w = sa[i % (TRI ? KH : 1)][(iv / 2) % VP];
        return (iv & 1) ? w >> 16 : w & 0xffffu;
    };
    u32 op[KH];
    // per chunk (at its first round): the thread's source string, its slot orbitals and
    // addresses, zeroed sums
    auto setup = [&](i64 c) {
        Jb = c * AB_SRC_NJ + jl;
        live = Jb < NB;
        const u64 sl = strsB[live ? Jb : NB - 1];
        const u64 s = live ? sl : 0ull;
        u64 mo = s, mv = ~s & ((n < 64) ? ((1ull << n) - 1) : ~0ull);
        for (int i = 0; i < half * KH; ++i) mo &= mo - 1;  // skip the other shares
#pragma unroll
        for (int i = 0; i < KH; ++i) {
            const u32 p = mo ? (u32)__builtin_ctzll(mo) : 0u;
            mo &= mo - 1;
            op[i] = p;
            if constexpr (!TRI) oo[i] = p * (u32)(n * R * sizeof(T));
            dd[i] = half * KH + i < kB ? (TRI ? tri_off(p, p) : p * (u32)((n + 1) * R * sizeof(T)) + copy_off)
                                       : zero_off;
        }
#pragma unroll
        for (int iv = 0; iv < VM; ++iv) {
            const u32 q = mv ? (u32)__builtin_ctzll(mv) : 0u;
            mv &= mv - 1;
            if constexpr (!TRI) vv[iv] = q * (u32)(R * sizeof(T)) + copy_off;
            if constexpr (TRI) {
#pragma unroll
                for (int i = 0; i < KH; ++i) {
                    const u32 a = op[i] < q ? tri_off(op[i], q) : tri_off(q, op[i]);
                    sa[i][iv / 2] = (iv & 1) ? (sa[i][iv / 2] | (a << 16)) : a;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < KH; ++i)
#pragma unroll
            for (int iv = 0; iv < VM; ++iv)
#pragma unroll
                for (int v = 0; v < NVS; ++v) acc[i][iv][v] = T(0);
#pragma unroll
        for (int v = 0; v < NVS; ++v) accd[v] = T(0);
    };
    i64 chunk = 0;
    int ca = 0;
    for (i64 it = 0; it < nit; ++it, advance(chunk, ca)) {
        if (ca == 0) setup(chunk);
        u32 wr_off = 0;  // byte offset of this round's replicated row
        if (DB == 2) {
            // this round's row (filled last round) is complete and the other buffer's last
            // readers are done: fill it with the next round's row, load the one after
            __syncthreads();
            wr_off = (u32)((it & 1) * NE * R * sizeof(T));
            if (it + 1 < nit) {
                store_w(Wr + ((it + 1) & 1) * NE * R, pneg ? -pre : pre);
            }
            wca = wca + 1 == nca ? 0 : wca + 1;
            load_w(pre, pneg, wca);
        } else {
            __syncthreads();  // everyone is done reading the previous W
            store_w(Wr, pneg ? -pre : pre);
            __syncthreads();
            load_w(pre, pneg, ca + 1 == nca ? 0 : ca + 1);
        }
        T xnew[NVS];
#pragma unroll
        for (int v = 0; v < NVS; ++v) xnew[v] = T(0);
        load_x(xnew, pchunk, pca);
        advance(pchunk, pca);
        {
            // a zero read from LDS after this round's barrier makes the row offsets opaque per
            // round, so the KH*VM slot addresses are formed in the loop (one add each) instead
            // of being hoisted into VGPRs (inline asm would do it too, but forces full waits)
            // HOIST: the KH x VM slot addresses may stay in VGPRs across rounds (single-vector,
            // single-buffer kernels, where the registers allow); otherwise a zero read from LDS
            // after this round's barrier makes the row offsets opaque per round, so the
            // addresses are formed in the loop (one add each) instead of being hoisted
            constexpr bool hoist = AB_SRC_HOIST && NVS == 1 && DB == 1 && KM * VM <= 64;
            const u32 lz = hoist ? 0u : *reinterpret_cast<const u32 *>(smem_raw + zero_off);
            u32 ob[KH];
#pragma unroll
            for (int i = 0; i < KH; ++i) ob[i] = TRI ? 0u : oo[i] + wr_off + lz;
            // software pipeline over the slot rows: row i+PD's VM gathers are issued before
            // row i's FMAs, so the LDS latency hides behind the FMAs of PD rows; the
            // scheduling barrier keeps the compiler from hoisting more rows of gathers (their
            // registers would push the accumulators out)
            constexpr int PD = AB_SRC_PD < KH ? AB_SRC_PD : KH;
            T wv[PD + 1][VM];
#pragma unroll
            for (int d = 0; d < PD; ++d)
#pragma unroll
                for (int iv = 0; iv < VM; ++iv)
                    wv[d][iv] = *reinterpret_cast<const T *>(smem_raw + (TRI ? slot_addr(d, iv) : ob[d] + vv[iv]));
#pragma unroll
            for (int i = 0; i < KH; ++i) {
                if (i + PD < KH) {
#pragma unroll
                    for (int iv = 0; iv < VM; ++iv)
                        wv[(i + PD) % (PD + 1)][iv] = *reinterpret_cast<const T *>(
                            smem_raw + (TRI ? slot_addr(i + PD, iv) : ob[i + PD] + vv[iv]));
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int iv = 0; iv < VM; ++iv)
#pragma unroll
                    for (int v = 0; v < NVS; ++v) acc[i][iv][v] += wv[i % (PD + 1)][iv] * xq[0][v];
            }
            T wd = 0;
#pragma unroll
            for (int i = 0; i < KH; ++i)
                wd += *reinterpret_cast<const T *>(smem_raw + dd[i] + (half * KH + i < kB ? wr_off : 0u));
#pragma unroll
            for (int v = 0; v < NVS; ++v) accd[v] += wd * xq[0][v];
        }
        // advance the source-element queue
#pragma unroll
        for (int d = 0; d + 1 < XPF; ++d)
#pragma unroll
            for (int v = 0; v < NVS; ++v) xq[d][v] = xq[d + 1][v];
#pragma unroll
        for (int v = 0; v < NVS; ++v) xq[XPF - 1][v] = xnew[v];
        if (ca == nca - 1) {
            // scatter this chunk's contributions, one wave after the other.  The target of
            // slot (io, iv) is computed, not loaded: with p_0 < .. < p_{k-1} the occupied and
            // t_iv the virtual orbitals of Jb, m = t - iv occupied orbitals lie below t, and
            // in colex order (rank = sum_j C(p_j, j+1))
            //   t > r = p_io:  rank(Jb - r + t) = Jb - C(r, io+1) + C(t, m)   + sum_{io<j<m}  [C(p_j, j)   - C(p_j, j+1)]
            //   t < r:         rank(Jb - r + t) = Jb - C(r, io+1) + C(t, m+1) + sum_{m<=j<io} [C(p_j, j+2) - C(p_j, j+1)]
            // split into a per-io and a per-iv term; the sign is the parity of the occupied
            // orbitals strictly between r and t: m - io - 1 (t > r) or io - m (t < r).
            constexpr int nwaves = BD / 64;
            const int wave = tid >> 6;
            // ordered: one wave after the other, so the LDS adds into each destination
            // element happen in a fixed order (bitwise reproducible); otherwise all waves at
            // once and the rounding of the sum depends on their timing
            for (int w = 0; w < (ordered ? nwaves : 1); ++w) {
                if ((!ordered || wave == w) && live) {
                    // (computed inside the turn so the temporaries are short-lived)
                    const u64 s = strsB[Jb];
                    // occupied orbitals and, per own occupied slot, -C(r,io+1) - A_{io+1} and
                    // -C(r,io+1) + B_io with A_j = sum_{i<j} C(p_i,i)-C(p_i,i+1),
                    // B_j = sum_{i<j} C(p_i,i+2)-C(p_i,i+1)
                    int pa[KM], up[KH], dn[KH];
                    {
                        u64 mo = s;
                        int A = 0, Bp = 0;
#pragma unroll
                        for (int j = 0; j < KM; ++j) {
                            const int p = mo ? __builtin_ctzll(mo) : 0;
                            mo &= mo - 1;
                            pa[j] = p;
                            const int c0 = binom[p * BW + j], c1 = binom[p * BW + j + 1], c2 = binom[p * BW + j + 2];
                            const int a = j < kB ? c0 - c1 : 0, b = j < kB ? c2 - c1 : 0;
                            const int i = j - half * KH;
                            if (i >= 0 && i < KH) {
                                up[i] = -c1 - (A + a);
                                dn[i] = -c1 + Bp;
                            }
                            A += a;
                            Bp += b;
                        }
                    }
                    u64 mv = ~s & ((n < 64) ? ((1ull << n) - 1) : ~0ull);
#pragma unroll
                    for (int iv = 0; iv < VM; ++iv) {
                        const int t = mv ? __builtin_ctzll(mv) : 0;
                        mv &= mv - 1;
                        if (iv < vB) {
                            // C(t, m) + A_m and C(t, m+1) - B_m
                            const int m = t - iv;
                            int F = binom[t * BW + m], G = binom[t * BW + m + 1];
#pragma unroll
                            for (int j = 0; j < KM; ++j) {
                                if (j < kB && pa[j] < t) {
                                    const int c1 = binom[pa[j] * BW + j + 1];
                                    F += binom[pa[j] * BW + j] - c1;
                                    G -= binom[pa[j] * BW + j + 2] - c1;
                                }
                            }
#pragma unroll
                            for (int i = 0; i < KH; ++i) {
                                const int io = half * KH + i;
                                if (io < kB) {
                                    const bool above = m > io;
                                    const i64 Ip = Jb + (above ? up[i] + F : dn[i] + G) - lo;
                                    const bool neg = (m + io + (above ? 1 : 0)) & 1;
                                    if (!PARTS || (u64)Ip < (u64)nbp) {
#pragma unroll
                                        for (int v = 0; v < NVS; ++v)
                                            atomicAdd(&S[Ip * NVS + v], (ST)(neg ? -acc[i][iv][v] : acc[i][iv][v]));
                                    }
                                }
                            }
                        }
                    }
                    if (!PARTS || (u64)(Jb - lo) < (u64)nbp) {
#pragma unroll
                        for (int v = 0; v < NVS; ++v) atomicAdd(&S[(Jb - lo) * NVS + v], (ST)accd[v]);
                    }
                }
                __syncthreads();
            }
        }
    }
    __syncthreads();
    for (i64 t = tid; t < nbp * NVS; t += BD) {
        const i64 Ip = t / NVS;
        T &o = s1[(Ia * NB + lo + Ip) * nvst + v0 + (t - Ip * NVS)];
        o = (T)((ST)o + S[t]);
    }
}
