// Sigma build  sigma = H x  on gfx950, excitation-driven (no Pauli network scan).
//
// Replaces cuda_hvec / hvec_gather_*_kernel (src/include/cuda/hvec.cuh:6-446) and the CPU
// contract_network_otf (src/include/op/hvec.hpp:481-496).  Pipeline per block of nv vectors:
//   K1 k_gauge_in   x -> x~ = E x, written alpha-major (xt) and beta-major (xtT), T precision
//   K2 k_rowspmm    s1  = G_alpha x~          rows = alpha strings, coalesced row streams
//   K3 k_rowspmm    s2T = G_beta  x~^T        same kernel on the beta-major copy
//   K4 k_ab         s1 += H_ab x~              one workgroup per destination alpha row; each
//                                               source row staged in LDS, beta connectivity
//                                               read per lane, integrals (pq|rs) row in LDS
//   K5 k_finalize   sigma = E (s1 + s2T^T + E_core x~)   tiled transpose, FP64 output
// T = double (drop-in FP64) or float (mixed precision).  Sum order differs from the
// reference, so FP64 results agree to rounding (~1e-15 relative), not bit-for-bit.
#include "hipotf.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

namespace hipotf {

constexpr int MAXNV = 8;
struct SrcPack { const double *p[MAXNV]; };
struct DstPack { double *p[MAXNV]; };

__device__ __forceinline__ int parity64(u64 v) { return __popcll(v) & 1; }

// v with its sign bit xor-ed with bit 31 of s (s is 0 or 0x80000000)
__device__ __forceinline__ double flip_sign(double v, u32 s) {
    return __longlong_as_double(__double_as_longlong(v) ^ ((long long)s << 32));
}
__device__ __forceinline__ float flip_sign(float v, u32 s) { return __uint_as_float(__float_as_uint(v) ^ s); }

// ---------------------------------------------------------------------------------------
// K1: gauge + dual layout (32x32 tiles through LDS)
// ---------------------------------------------------------------------------------------
template <typename T, int NV>
__global__ __launch_bounds__(256) void k_gauge_in(SrcPack src, T *__restrict__ xt, T *__restrict__ xtT,
                                                  const u64 *__restrict__ strsB,
                                                  const u64 *__restrict__ M, i64 NA, i64 NB) {
    __shared__ T tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const i64 b0 = (i64)blockIdx.x * 32, a0 = (i64)blockIdx.y * 32;
#pragma unroll 1
    for (int v = 0; v < NV; ++v) {
        for (int i = 0; i < 4; ++i) {
            const i64 a = a0 + ty + 8 * i, b = b0 + tx;
            if (a < NA && b < NB) {
                double val = src.p[v][a * NB + b];
                if (parity64(strsB[b] & M[a])) val = -val;
                const T tv = (T)val;
                xt[(a * NB + b) * NV + v] = tv;
                tile[ty + 8 * i][tx] = tv;
            }
        }
        __syncthreads();
        for (int i = 0; i < 4; ++i) {
            const i64 b = b0 + ty + 8 * i, a = a0 + tx;
            if (a < NA && b < NB) xtT[(b * NA + a) * NV + v] = tile[tx][ty + 8 * i];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// K2/K3: Y[I, :] (+)= sum_l C[I,l] X[J[I,l], :]   (CSR rows sorted by source, uniform per wave)
// One wave = one destination row x 64 consecutive (flattened) columns.  Blocks are mapped so
// that each XCD walks all rows of one column tile before the next.  A launch covers only the
// sources in ranges [g0, g1) (entries P[I][g0] .. P[I][g1]): the tile's source columns then
// span 1/npass of the rows and stay resident in the XCD's 4 MB L2; npass launches sum up.
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_rowspmm(const T *__restrict__ X, T *__restrict__ Y,
                                                 const int32_t *__restrict__ J,
                                                 const int32_t *__restrict__ P,
                                                 const T *__restrict__ C, int L, i64 nrows,
                                                 i64 ncols, i64 ntiles, i64 bpt, int g0, int g1,
                                                 int accumulate) {
    const int lane = threadIdx.x & 63;
    const i64 b = blockIdx.x;
    const i64 xcd = b & 7, j = b >> 3;
    const i64 tile = (j / bpt) * 8 + xcd;
    const i64 rb = j % bpt;
    if (tile >= ntiles) return;
    const i64 row = __builtin_amdgcn_readfirstlane((int)(rb * 4 + (threadIdx.x >> 6)));
    if (row >= nrows) return;
    const i64 col = tile * 64 + lane;
    const bool ok = col < ncols;
    const int32_t *Jr = J + row * (i64)L;
    const T *Cr = C + row * (i64)L;
    const T *Xc = X + col;
    const int lb = P[row * (SS_RANGES + 1) + g0], le = P[row * (SS_RANGES + 1) + g1];
    T acc0 = 0, acc1 = 0;
    int l = lb;
    if (ok) {
        // U gathers in flight per wave; the next batch's (uniform) indices are loaded while
        // the current batch's gathers are outstanding
        constexpr int U = 16;
        int jn[U];
        if (l + U <= le) {
#pragma unroll
            for (int u = 0; u < U; ++u) jn[u] = Jr[l + u];
        }
        for (; l + U <= le; l += U) {
            int jc[U];
#pragma unroll
            for (int u = 0; u < U; ++u) jc[u] = jn[u];
            T xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = Xc[(i64)jc[u] * ncols];
            if (l + 2 * U <= le) {
#pragma unroll
                for (int u = 0; u < U; ++u) jn[u] = Jr[l + U + u];
            }
#pragma unroll
            for (int u = 0; u < U; u += 2) {
                acc0 += Cr[l + u] * xv[u];
                acc1 += Cr[l + u + 1] * xv[u + 1];
            }
        }
        for (; l < le; ++l) acc0 += Cr[l] * Xc[(i64)Jr[l] * ncols];
        T *y = Y + row * ncols + col;
        *y = accumulate ? *y + (acc0 + acc1) : acc0 + acc1;
    }
}

// ---------------------------------------------------------------------------------------
// K2/K3 on row blocks: Y[I0 + r, :] (+)= sum_e C[e][r] X[J[e], :], r < RB, over the union J of
// the sources of the RB consecutive destination rows I0 .. I0 + RB - 1 (RowBlockTable).  Each
// gathered source slice feeds RB accumulators, so the gather volume drops to the union's size
// (at 16 orbitals, 8 alpha: 0.43 of the per-row volume for RB = 8) at the price of RB FMAs per
// gather, zeros included.  Workgroup = one row block x 256 consecutive columns (4 waves x 64);
// the source index and the RB coefficients of an entry are uniform (scalar loads) and shared
// by the 4 waves.  Work items (tile-major: all row blocks of a column tile, then the next
// tile) are dealt to the 8 XCDs in contiguous ranges, so each XCD sweeps whole column tiles.
// ---------------------------------------------------------------------------------------
// CPL consecutive columns per lane, up to 16 bytes (2 FP64 or 4 FP32 columns) when the row
// length allows: the gathers are bound by the texture pipeline's per-lane work (TA ~85%, TD
// ~92% busy at ~26 B/clk/CU with one FP64 column per lane), so wider lanes move more bytes per
// cycle (FP64 sigma at 16 orbitals: 218 -> 209 ms)
template <typename T>
static int rs_cpl(i64 ncols) {
    constexpr int wide = 16 / (int)sizeof(T);
    if (ncols % wide == 0) return wide;
    return ncols % 2 == 0 ? 2 : 1;
}

template <typename T, int RB, int CPL>
__global__ __launch_bounds__(256) void k_rbspmm(const T *__restrict__ X, T *__restrict__ Y,
                                                const int32_t *__restrict__ J,
                                                const int32_t *__restrict__ P,
                                                const T *__restrict__ C, i64 nrows, i64 ncols,
                                                i64 nblk, i64 nwork, i64 per_xcd, int g0, int g1,
                                                int accumulate) {
    typedef T TV __attribute__((ext_vector_type(CPL)));
    const i64 b = blockIdx.x;
    const i64 w = (b & 7) * per_xcd + (b >> 3);
    if (w >= nwork) return;
    const i64 tile = w / nblk, blk = w - (w / nblk) * nblk;
    const i64 col = (tile * 256 + threadIdx.x) * CPL;  // ncols is a multiple of CPL
    if (col >= ncols) return;
    const int eb = P[blk * (SS_RANGES + 1) + g0], ee = P[blk * (SS_RANGES + 1) + g1];
    TV acc[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) acc[r] = 0;
    const T *Xc = X + col;
    auto gather = [&](int e) { return *reinterpret_cast<const TV *>(Xc + (i64)J[e] * ncols); };
    // batches of U gathers in flight per wave.  (Measured at 16 orbitals: issuing batch i + 1
    // before batch i's FMAs, or buffer loads with a scalar row base, are 2-7% slower.)
    constexpr int U = 8;
    int e = eb;
    for (; e + U <= ee; e += U) {
        TV xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = gather(e + u);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const T *c = C + (i64)(e + u) * RB;
#pragma unroll
            for (int r = 0; r < RB; ++r) acc[r] += c[r] * xv[u];
        }
    }
    for (; e < ee; ++e) {
        const TV x = gather(e);
        const T *c = C + (i64)e * RB;
#pragma unroll
        for (int r = 0; r < RB; ++r) acc[r] += c[r] * x;
    }
    const i64 I0 = blk * RB;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        if (I0 + r < nrows) {
            TV *y = reinterpret_cast<TV *>(Y + (I0 + r) * ncols + col);
            *y = accumulate ? *y + acc[r] : acc[r];
        }
    }
}

// ---------------------------------------------------------------------------------------
// K4: alpha-beta part.  Workgroup = destination alpha row Ia.  For each alpha connection
// (its nsxA single excitations Ja = Ia - p + q, then Ia itself) the source row x~[Ja, :] is
// staged in LDS (in chunks if it does not fit) together with the integral row
//   W[rs] = s_a (pq|rs)            (single)   or   W[rs] = sum_{p in Ia} (pp|rs)   (diagonal)
// and every lane Ib accumulates  sum_{singles Ib<-Jb} s_b W[rs] x~[Ja,Jb] + (sum_{r in Ib} W[rr]) x~[Ja,Ib].
// ---------------------------------------------------------------------------------------
template <typename T, int NVS, int EPT>
__global__ __launch_bounds__(1024) void k_ab(const T *__restrict__ xt, T *__restrict__ s1, int nvst,
                                             int v0, const u32 *__restrict__ sxA,
                                             const u32 *__restrict__ sxB,
                                             const u64 *__restrict__ strsA,
                                             const u64 *__restrict__ strsB, const T *__restrict__ eri,
                                             int n, i64 NA, i64 NB, int nsxA, int nsxB, int ept,
                                             i64 chunk, int nchunks, int nsplit, i64 per_split) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *W = reinterpret_cast<T *>(smem_raw);
    const int n2 = n * n;
    T *xrow = W + ((n2 + 15) & ~15);
    const int tid = threadIdx.x, bd = blockDim.x;
    // blocks are dispatched round-robin over the 8 XCDs: give each XCD a contiguous range of
    // destination rows, and put the nsplit beta-slices of one row on the same XCD back to back
    // so the staged source rows are served from that XCD's L2
    const i64 per = (NA + 7) / 8;
    const i64 jx = blockIdx.x >> 3;
    const i64 Ia = (blockIdx.x & 7) * per + jx / nsplit;
    const i64 ib0 = (jx % nsplit) * per_split;
    if (Ia >= NA) return;
    const i64 ib1 = min(NB, ib0 + per_split);
    const u64 astr = strsA[Ia];

    T acc[EPT][NVS];
#pragma unroll
    for (int e = 0; e < EPT; ++e)
#pragma unroll
        for (int v = 0; v < NVS; ++v) acc[e][v] = 0;

    for (int ca = 0; ca <= nsxA; ++ca) {
        i64 Ja;
        __syncthreads();
        if (ca < nsxA) {
            const u32 d = sxA[(i64)ca * NA + Ia];
            Ja = sx_J(d);
            const T sg = sx_neg(d) ? T(-1) : T(1);
            const T *er = eri + (i64)(sx_r(d) * n + sx_s(d)) * n2;
            for (int t = tid; t < n2; t += bd) W[t] = sg * er[t];
        } else {
            Ja = Ia;
            for (int t = tid; t < n2; t += bd) {
                T v = 0;
                u64 m = astr;
                while (m) {
                    const int p = __builtin_ctzll(m);
                    m &= m - 1;
                    v += eri[(i64)(p * n + p) * n2 + t];
                }
                W[t] = v;
            }
        }
        for (int c = 0; c < nchunks; ++c) {
            const i64 base = (i64)c * chunk;
            const i64 len = min(chunk, NB - base);
            if (c > 0) __syncthreads();
            const T *srow = xt + (Ja * NB + base) * nvst + v0;
            if (NVS == 1 && nvst == 1) {
                for (i64 t = tid; t < len; t += bd) xrow[t] = srow[t];
            } else {
                for (i64 t = tid; t < len * NVS; t += bd) {
                    const i64 el = t / NVS;
                    xrow[t] = srow[el * nvst + (t - el * NVS)];
                }
            }
            __syncthreads();
#pragma unroll
            for (int e = 0; e < EPT; ++e) {
                const i64 Ib = ib0 + tid + (i64)e * bd;
                if (e >= ept || Ib >= ib1) continue;
                T a[NVS];
#pragma unroll
                for (int v = 0; v < NVS; ++v) a[v] = 0;
                for (int cb = 0; cb < nsxB; ++cb) {
                    const u32 d = sxB[(i64)cb * NB + Ib];
                    const i64 off = (i64)sx_J(d) - base;
                    if (nchunks > 1 && (off < 0 || off >= len)) continue;
                    T w = W[sx_r(d) * n + sx_s(d)];
                    if (sx_neg(d)) w = -w;
#pragma unroll
                    for (int v = 0; v < NVS; ++v) a[v] += w * xrow[off * NVS + v];
                }
                const i64 offd = Ib - base;
                if (offd >= 0 && offd < len) {
                    T wd = 0;
                    u64 m = strsB[Ib];
                    while (m) {
                        const int r = __builtin_ctzll(m);
                        m &= m - 1;
                        wd += W[r * n + r];
                    }
#pragma unroll
                    for (int v = 0; v < NVS; ++v) a[v] += wd * xrow[offd * NVS + v];
                }
#pragma unroll
                for (int v = 0; v < NVS; ++v) acc[e][v] += a[v];
            }
        }
    }
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const i64 Ib = ib0 + tid + (i64)e * bd;
        if (e >= ept || Ib >= ib1) continue;
#pragma unroll
        for (int v = 0; v < NVS; ++v) s1[(Ia * NB + Ib) * nvst + v0 + v] += acc[e][v];
    }
}

// ---------------------------------------------------------------------------------------
// K4 (register variant): one beta element per lane, its nsxB single-excitation descriptors
// held in VGPRs for the whole sweep over the alpha connections, so the connectivity table is
// read once per workgroup instead of once per staged source row.  A destination row is
// covered by ceil(NB / blockDim) workgroups that run back to back on one XCD, so the source
// rows they all stage come from that XCD's L2.
// ---------------------------------------------------------------------------------------
template <typename T, int NVS, int SXMAX>
__global__ __launch_bounds__(1024) void k_ab_reg(const T *__restrict__ xt, T *__restrict__ s1,
                                                 int nvst, int v0, const u32 *__restrict__ sxA,
                                                 const u32 *__restrict__ sxB,
                                                 const u64 *__restrict__ strsA,
                                                 const u64 *__restrict__ strsB,
                                                 const T *__restrict__ eri, int n, i64 NA, i64 NB,
                                                 int nsxA, int nsxB, int nsplit) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *W = reinterpret_cast<T *>(smem_raw);
    const int n2 = n * n;
    T *Wd = W + n2;                       // W[r*n + r], r < n
    T *xrow = W + ((n2 + n + 15) & ~15);
    const int tid = threadIdx.x, bd = blockDim.x;
    const i64 per = (NA + 7) / 8;
    const i64 jx = blockIdx.x >> 3;
    const i64 Ia = (blockIdx.x & 7) * per + jx / nsplit;
    if (Ia >= NA) return;
    const i64 Ib = (jx % nsplit) * (i64)bd + tid;
    const bool live = Ib < NB;
    const u64 astr = strsA[Ia];

    // lane-ready descriptors in LDS byte offsets: bits 0-17 the source element x~[Ja, J] inside
    // the staged row (whole frame < 160 KB < 2^18), bits 18-30 the integral W[r*n+s]
    // (n <= 32: < 8 KB), bit 31 the sign of <Ib|E_rs|J>
    const u32 xoff = (u32)(((n2 + n + 15) & ~15) * sizeof(T));
    u32 d[SXMAX];
#pragma unroll
    for (int c = 0; c < SXMAX; ++c) {
        u32 e = (live && c < nsxB) ? sxB[(i64)c * NB + Ib] : 0u;
        d[c] = (xoff + sx_J(e) * (u32)(NVS * sizeof(T))) |
               ((u32)((sx_r(e) * n + sx_s(e)) * sizeof(T)) << 18) | (e & 0x80000000u);
    }
    const u64 bocc = live ? strsB[Ib] : 0ull;

    T acc[NVS];
#pragma unroll
    for (int v = 0; v < NVS; ++v) acc[v] = 0;

    for (int ca = 0; ca <= nsxA; ++ca) {
        i64 Ja;
        __syncthreads();
        if (ca < nsxA) {
            const u32 e = sxA[(i64)ca * NA + Ia];
            Ja = sx_J(e);
            const T sg = sx_neg(e) ? T(-1) : T(1);
            const T *er = eri + (i64)(sx_r(e) * n + sx_s(e)) * n2;
            for (int t = tid; t < n2; t += bd) W[t] = sg * er[t];
            for (int t = tid; t < n; t += bd) Wd[t] = sg * er[t * n + t];
        } else {
            Ja = Ia;
            for (int t = tid; t < n2 + n; t += bd) {
                const int tt = t < n2 ? t : (t - n2) * (n + 1);
                T v = 0;
                u64 m = astr;
                while (m) {
                    const int p = __builtin_ctzll(m);
                    m &= m - 1;
                    v += eri[(i64)(p * n + p) * n2 + tt];
                }
                W[t] = v;  // t >= n2 lands in Wd[t - n2]
            }
        }
        const T *srow = xt + Ja * NB * nvst + v0;
        if (NVS == 1 && nvst == 1) {
            for (i64 t = tid; t < NB; t += bd) xrow[t] = srow[t];
        } else {
            for (i64 t = tid; t < NB * NVS; t += bd) {
                const i64 el = t / NVS;
                xrow[t] = srow[el * nvst + (t - el * NVS)];
            }
        }
        __syncthreads();
        if (!live) continue;
        T a[NVS];
#pragma unroll
        for (int v = 0; v < NVS; ++v) a[v] = 0;
#pragma unroll
        for (int c = 0; c < SXMAX; ++c) {
            if (c < nsxB) {
                // The empty asm ties the decode to this alpha connection (input ca), so it is
                // not hoisted out of the ca loop (two VGPRs per slot would spill); it has no
                // side effects, so the LDS loads still schedule freely around it.
                u32 e = d[c];
                asm("" : "+v"(e) : "s"(ca));
                const T w = flip_sign(*reinterpret_cast<const T *>(smem_raw + ((e >> 18) & 0x1FFFu)),
                                      e & 0x80000000u);
                const T *xp = reinterpret_cast<const T *>(smem_raw + (e & 0x3FFFFu));
#pragma unroll
                for (int v = 0; v < NVS; ++v) a[v] += w * xp[v];
            }
        }
        T wd = 0;
        u64 m = bocc;
        while (m) {
            const int r = __builtin_ctzll(m);
            m &= m - 1;
            wd += Wd[r];
        }
        const T *xd = xrow + Ib * NVS;
#pragma unroll
        for (int v = 0; v < NVS; ++v) acc[v] += a[v] + wd * xd[v];
    }
    if (live) {
#pragma unroll
        for (int v = 0; v < NVS; ++v) s1[(Ia * NB + Ib) * nvst + v0 + v] += acc[v];
    }
}

#include "ab_src.hpp"
#include "ab_cls.hpp"
#include "ab_mma.hpp"

// ---------------------------------------------------------------------------------------
// K5: sigma = E (s1 + s2T^T + E_core x~), FP64 output vectors
// ---------------------------------------------------------------------------------------
template <typename T, int NV>
__global__ __launch_bounds__(256) void k_finalize(const T *__restrict__ s1, const T *__restrict__ s2T,
                                                  const T *__restrict__ xt, double ecore, DstPack dst,
                                                  const u64 *__restrict__ strsB,
                                                  const u64 *__restrict__ M, i64 NA, i64 NB) {
    __shared__ T tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const i64 b0 = (i64)blockIdx.x * 32, a0 = (i64)blockIdx.y * 32;
#pragma unroll 1
    for (int v = 0; v < NV; ++v) {
        for (int i = 0; i < 4; ++i) {
            const i64 b = b0 + ty + 8 * i, a = a0 + tx;
            if (a < NA && b < NB) tile[ty + 8 * i][tx] = s2T[(b * NA + a) * NV + v];
        }
        __syncthreads();
        for (int i = 0; i < 4; ++i) {
            const i64 a = a0 + ty + 8 * i, b = b0 + tx;
            if (a < NA && b < NB) {
                const i64 k = (a * NB + b) * NV + v;
                double val = (double)s1[k] + (double)tile[tx][ty + 8 * i] + ecore * (double)xt[k];
                if (parity64(strsB[b] & M[a])) val = -val;
                dst.p[v][a * NB + b] = val;
            }
        }
        __syncthreads();
    }
}

// diag[i] += H_ii  (accumulates, as get_diags_elements does: src/include/op/hvec.hpp:498-554)
__global__ __launch_bounds__(256) void k_diag(double *__restrict__ diags, const double *__restrict__ gdA,
                                              const double *__restrict__ gdB,
                                              const double *__restrict__ jab,
                                              const u64 *__restrict__ strsB, int n, i64 NA, i64 NB,
                                              double ecore) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NA * NB) return;
    const i64 a = i / NB, b = i - a * NB;
    double v = gdA[a] + gdB[b] + ecore;
    u64 m = strsB[b];
    const double *ja = jab + a * n;
    while (m) {
        const int r = __builtin_ctzll(m);
        m &= m - 1;
        v += ja[r];
    }
    diags[i] += v;
}

// ---------------------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------------------
static size_t lds_budget() {
    static size_t cached = 0;
    if (!cached) {
        int dev = 0, v = 0;
        HIPOTF_CHECK(hipGetDevice(&dev));
        HIPOTF_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        cached = (size_t)v;
    }
    return cached;
}

template <typename T, int RB, int CPL>
static void launch_rbspmm_c(const T *X, T *Y, const RowBlockTable &rt, i64 nrows, i64 ncols, int npass,
                            hipStream_t st) {
    const i64 ntiles = (ncols + 256 * CPL - 1) / (256 * CPL);
    const i64 nwork = ntiles * rt.nblk;
    const i64 per_xcd = (nwork + 7) / 8;
    HIPOTF_REQUIRE(per_xcd * 8 < (1ll << 31), "row-block SpMM grid too large");
    const T *C = sizeof(T) == 4 ? (const T *)rt.c32.p : (const T *)rt.c64.p;
    for (int p = 0; p < npass; ++p)
        hipLaunchKernelGGL((k_rbspmm<T, RB, CPL>), dim3((unsigned)(per_xcd * 8)), dim3(256), 0, st, X, Y,
                           rt.J.as<int32_t>(), rt.P.as<int32_t>(), C, nrows, ncols, rt.nblk, nwork,
                           per_xcd, p * SS_RANGES / npass, (p + 1) * SS_RANGES / npass, p > 0 ? 1 : 0);
    HIPOTF_CHECK(hipGetLastError());
}

template <typename T, int RB>
static void launch_rbspmm(const T *X, T *Y, const RowBlockTable &rt, i64 nrows, i64 ncols, int npass,
                          hipStream_t st) {
    // (the buffers are 256-byte aligned and ncols is a multiple of CPL, so every lane's columns
    // are aligned to their size)
    const int cpl = rs_cpl<T>(ncols);
    if constexpr (sizeof(T) == 4) {
        if (cpl == 4) return launch_rbspmm_c<T, RB, 4>(X, Y, rt, nrows, ncols, npass, st);
    }
    if (cpl == 2) launch_rbspmm_c<T, RB, 2>(X, Y, rt, nrows, ncols, npass, st);
    else launch_rbspmm_c<T, RB, 1>(X, Y, rt, nrows, ncols, npass, st);
}

// source-range passes of a row SpMM: the tuned value, or (0) the fewest passes (1, 2, 4, 8) for
// which the gathered slice of one column tile, nrows x tile_cols values, fits an XCD's 4 MB L2
template <typename T>
static int rs_npass(i64 nrows, int tile_cols) {
    const int tp = tuning(sizeof(T) == 4).rs_passes;
    if (tp == 1 || tp == 2 || tp == 4 || tp == 8) return tp;
    const double slice = (double)nrows * tile_cols * sizeof(T), l2 = 4.0 * 1024 * 1024;
    int p = 1;
    while (p < 8 && slice > l2 * p) p *= 2;
    return p;
}

template <typename T>
static void launch_rowspmm(const T *X, T *Y, SameSpinTable &tab, const T *C, i64 nrows,
                           i64 ncols, hipStream_t st) {
    if (nrows == 0 || ncols == 0) return;
    HIPOTF_REQUIRE(nrows == tab.N, "internal: same-spin table does not match the row count");
    const int rb = tuning(sizeof(T) == 4).rs_block;
    if (rb == 4 || rb == 8 || rb == 16) {
        const int npass = rs_npass<T>(nrows, 256 * rs_cpl<T>(ncols));
        const RowBlockTable &rt = row_blocks(tab, rb);
        if (rb == 4) launch_rbspmm<T, 4>(X, Y, rt, nrows, ncols, npass, st);
        else if (rb == 8) launch_rbspmm<T, 8>(X, Y, rt, nrows, ncols, npass, st);
        else launch_rbspmm<T, 16>(X, Y, rt, nrows, ncols, npass, st);
        return;
    }
    const i64 ntiles = (ncols + 63) / 64;
    const i64 bpt = (nrows + 3) / 4;
    const i64 ntiles_pad = (ntiles + 7) / 8 * 8;
    const i64 nblocks = ntiles_pad * bpt;
    HIPOTF_REQUIRE(nblocks < (1ll << 31), "row-SpMM grid too large");
    // passes over source ranges (each pass's source columns of a tile span 1/npass of the
    // rows).  Measured at 16 orbitals: 1, 2 and 4 passes take the same time here -- these
    // gathers are bound by the per-CU load pipeline (~31 B/clk), not by L2 misses
    const int npass = rs_npass<T>(nrows, 64);
    for (int p = 0; p < npass; ++p)
        hipLaunchKernelGGL(k_rowspmm<T>, dim3((unsigned)nblocks), dim3(256), 0, st, X, Y,
                           tab.J.as<int32_t>(), tab.P.as<int32_t>(), C, tab.L, nrows, ncols, ntiles,
                           bpt, p * SS_RANGES / npass, (p + 1) * SS_RANGES / npass, p > 0 ? 1 : 0);
    HIPOTF_CHECK(hipGetLastError());
}

template <typename T, int NVS, int EPT>
static void launch_ab_ept(DevHam *H, const T *xt, T *s1, int nvst, int v0, const T *eri, int bd,
                          int ept, i64 chunk, int nchunks, int nsplit, i64 per_split, size_t smem,
                          hipStream_t st) {
    DevBasis *B = H->basis;
    const DevSpin &A = *B->pA, &Bs = *B->pB;
    auto kern = k_ab<T, NVS, EPT>;
    HIPOTF_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)smem));
    const i64 per = (A.N + 7) / 8;
    const i64 nblocks = per * nsplit * 8;
    HIPOTF_REQUIRE(nblocks < (1ll << 31), "alpha-beta grid too large");
    hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(bd), smem, st, xt, s1, nvst, v0,
                       A.sx.as<u32>(), Bs.sx.as<u32>(), A.strs.as<u64>(), Bs.strs.as<u64>(), eri,
                       H->n, A.N, Bs.N, A.nsx, Bs.nsx, ept, chunk, nchunks, nsplit, per_split);
    HIPOTF_CHECK(hipGetLastError());
}

// accumulator budget per lane: EPT_MAX * NVS values of T within ~64 VGPRs (no spills at
// 1024 threads / 128 VGPRs); a destination row wider than bd * EPT_MAX is split over blocks
template <typename T, int NVS>
constexpr int ab_eptmax() { return (sizeof(T) == 8 && NVS >= 4) ? 4 : 8; }

// alpha-beta kernel family (SigmaTuning::ab_kernel): 0 takes the first of matrix-core (14 to 16
// orbitals), class-split (FP32 vector pairs, ab_cls.hpp), source-lane, register, chunked that
// applies; 1..3 start the search at source-lane, register, chunked; 4 is 1 with FP32 vector
// pairs on the class-split kernel; 5 puts the matrix-core kernel (ab_mma.hpp, one vector per
// launch, n <= 16) first.  The matrix-core kernel's default range is where it was measured
// faster: at 12 orbitals its FP64 sigma is 8% slower (0.66 vs 0.61 ms), at 14 and 16 it is 1.4
// and 1.6 times faster.
static bool ab_mma_first(int pref, int n) { return pref == 5 || (pref == 0 && n >= 14 && n <= 16); }
static int ab_kernel_pref(int prec) {
    const int p = tuning(prec).ab_kernel;
    return p >= 0 && p <= 5 ? p : 0;
}

AbTiming &ab_timing() {
    static AbTiming t;
    return t;
}

// bitwise reproducible sigma (the alpha-beta scatter runs wave by wave, ~15% slower); by
// default its FP summation order depends on scheduling.  The initial value comes from the
// environment (HIPOTF_DETERMINISTIC=1).
SigmaTuning &tuning(int precision) {
    static SigmaTuning t[2] = {[] {
        SigmaTuning d;
        const char *s = std::getenv("HIPOTF_DETERMINISTIC");
        d.deterministic = s && std::atoi(s) != 0;
        return d;
    }()};
    static bool init = [] { t[1] = t[0]; return true; }();
    (void)init;
    return t[precision ? 1 : 0];
}
static bool deterministic(int prec) { return tuning(prec).deterministic != 0; }

template <typename T, int NVS, int R, int NW, bool TRI = false>
static bool try_ab_src_r(DevHam *H, const T *xt, T *s1, int nvst, int v0, const T *eri, hipStream_t st);

// entries per integral row of the source-lane kernel for n orbitals (n*n <= 512)
static int ab_src_nw(int n) { return n * n <= 256 ? 256 : 512; }

// dw[Ia][rt] = sum_{p in Ia} (pp|rt) for rt < n*n, 0 up to the row width NW: the integral row
// of the diagonal alpha connection, read by the source-lane kernel like the (pq|..) rows
template <typename T>
__global__ __launch_bounds__(512) void k_diag_w(const u64 *__restrict__ strsA,
                                                const double *__restrict__ eri, int n, int nw,
                                                T *__restrict__ dw) {
    const i64 Ia = blockIdx.x;
    const int t = threadIdx.x, n2 = n * n;
    double v = 0.0;
    if (t < n2) {
        u64 m = strsA[Ia];
        while (m) {
            const int p = __builtin_ctzll(m);
            m &= m - 1;
            v += eri[(i64)(p * n + p) * n2 + t];
        }
    }
    if (t < nw) dw[Ia * nw + t] = (T)v;
}

template <typename T>
static const T *diag_w_table(DevHam *H, hipStream_t st) {
    DevBuf &b = sizeof(T) == 8 ? H->dw64 : H->dw32;
    const DevSpin &A = *H->basis->pA;
    if (!b.p) {
        const int nw = ab_src_nw(H->n);
        HIPOTF_REQUIRE(A.N > 0 && A.N < (1ll << 31), "alpha string count out of range");
        b.alloc((size_t)A.N * nw * sizeof(T));
        hipLaunchKernelGGL(k_diag_w<T>, dim3((unsigned)A.N), dim3(nw), 0, st, A.strs.as<u64>(),
                           H->eri64.as<double>(), H->n, nw, b.as<T>());
        HIPOTF_CHECK(hipGetLastError());
    }
    return b.as<T>();
}

template <typename T, int NVS>
static bool try_ab_src(DevHam *H, const T *xt, T *s1, int nvst, int v0, const T *eri, hipStream_t st) {
    if constexpr (NVS > 2 || (sizeof(T) == 8 && NVS == 2)) {
        return false;  // accumulators would not fit the register file
    } else {
        // R copies of the integral row: 32 lanes of 4-byte reads are served together, so 32
        // copies make them conflict-free; 8-byte reads get 32 copies of the r <= t triangle in
        // the single-buffer single-vector form (TRI), else 16 of the full row (LDS space).
        // (FP32 with 16 copies at 16 orbitals: the same for two vectors, 9% slower for one; 64
        // copies 11% slower for one.  FP64 with 8 copies: sigma 208 -> 216 ms.)
        if (H->n * H->n <= 256) {
            if constexpr (sizeof(T) == 8 && NVS == 1)
                if (try_ab_src_r<T, NVS, 32, 256, true>(H, xt, s1, nvst, v0, eri, st)) return true;
            return try_ab_src_r<T, NVS, sizeof(T) == 8 ? 16 : 32, 256>(H, xt, s1, nvst, v0, eri, st);
        }
        // 17-22 orbitals: 512-entry rows, single-vector kernels only, 8 copies of the integral
        // row (FP32 at 18 orbitals / 7 electrons: 8 copies 4% faster than 16).  FP32: the 127 KB
        // destination row stays whole.  FP64: the 254 KB row is done in two halves, and at most
        // 7 x 11 slots per string (77 accumulators)
        if constexpr (NVS == 1) return try_ab_src_r<T, NVS, 8, 512>(H, xt, s1, nvst, v0, eri, st);
        return false;
    }
}

template <typename T, int NVS, int R, int NW, bool TRI>
static bool try_ab_src_r(DevHam *H, const T *xt, T *s1, int nvst, int v0, const T *eri, hipStream_t st) {
    {
        DevBasis *B = H->basis;
        const DevSpin &A = *B->pA, &Bs = *B->pB;
        const int n = H->n, n2 = n * n;
        const int kB = Bs.k, vB = Bs.n - Bs.k;
        // threads per source string: 1 (512-thread blocks, 2 waves per SIMD, accumulators in
        // up to 256 VGPRs) or 2 (1024-thread blocks with half the slots per thread)
        constexpr int prec = sizeof(T) == 4;
        const int split = NW == 256 && tuning(prec).ab_split == 2 ? 2 : 1;
        const i64 NB = Bs.N;
        const int bd = AB_SRC_NJ * split;
        if (n2 > NW || ab_src_nw(n) != NW || A.nsx > AB_SRC_CONN || n * (8 + 2) > AB_SRC_BINOM) return false;
        // a W fill reads NW entries from the start of one (pq|..) row
        HIPOTF_REQUIRE(NW - n2 <= ERI_PAD, "internal: integral table padding too short for the row width");
        // + the binomial table C(p, j), p < n <= 18, j < KM + 2 <= 10; two replicated rows
        // (one barrier per connection) if tuned so and they fit
        constexpr int NE = TRI ? AB_SRC_NTRI : NW;  // entries stored per integral row
        auto lds_for = [&](int db, i64 nbp) {
            return ((size_t)db * NE * R + 16) * sizeof(T) + (size_t)nbp * NVS * sizeof(AbSrcRow<T, NVS, NW>) +
                   (AB_SRC_BINOM + AB_SRC_CONN) * 4;
        };
        const size_t lds_max = std::min<size_t>(lds_budget(), 160 * 1024);
        const int db = (tuning(prec).ab_db == 2 && lds_for(2, NB) <= lds_max) ? 2 : 1;
        // the triangular row (TRI) serves the single-buffer form with one lane per source string
        if (TRI && (db != 1 || split != 1 || n * (n + 1) / 2 > AB_SRC_NTRI)) return false;
        // destination-row parts: one, or (FP64 512-entry form) as few as fit the LDS
        int parts = 1;  // (the kernel checks scatter targets against its part in this form only)
        if (NW == 512 && sizeof(T) == 8)
            while (parts < 4 && lds_for(db, (NB + parts - 1) / parts) > lds_max) ++parts;
        const i64 nbp = (NB + parts - 1) / parts;
        const size_t smem = lds_for(db, nbp);
        if (smem > lds_max) return false;
        const i64 per = (A.N + 7) / 8;
        const i64 nblocks = per * 8;
        HIPOTF_REQUIRE(nblocks < (1ll << 31), "alpha-beta grid too large");
        const T *dw = diag_w_table<T>(H, st);
        auto go = [&](auto kern) {
            HIPOTF_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)smem));
            for (int part = 0; part < parts; ++part) {
                const i64 lo = part * nbp, cnt = std::min<i64>(nbp, NB - lo);
                hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(bd), smem, st, xt, s1, nvst, v0,
                                   A.sx.as<u32>(), Bs.sx.as<u32>(), A.strs.as<u64>(), Bs.strs.as<u64>(), eri,
                                   dw, n, A.N, NB, A.nsx, kB, vB, deterministic(prec) ? 1 : 0, lo, cnt);
                HIPOTF_CHECK(hipGetLastError());
            }
        };
        if constexpr (NW == 512) {
            // 17-18 orbitals: up to 8 occupied and 12 virtual orbitals per beta string (FP32),
            // 7 and 11 (FP64)
            if constexpr (sizeof(T) == 8) {
                if (kB <= 7 && vB <= 11 && db == 1) {
                    go(k_ab_src<T, NVS, 7, 11, R, 1, 1, NW>);
                    return true;
                }
            } else if (kB <= 7 && vB <= 11 && db == 1) {
                go(k_ab_src<T, NVS, 7, 11, R, 1, 1, NW>);  // 18 orbitals / 7 electrons: no empty slots
                return true;
            } else if (kB <= 8 && vB <= 12 && db == 1) {
                go(k_ab_src<T, NVS, 8, 12, R, 1, 1, NW>);
                return true;
            }
            return false;
        } else {
            auto pick = [&](auto split_c, auto db_c) {
                constexpr int SPLIT = decltype(split_c)::value, DB = decltype(db_c)::value;
                if (kB <= 4 && vB <= 4) go(k_ab_src<T, NVS, 4, 4, R, SPLIT, DB, NW, TRI>);
                else if (kB <= 6 && vB <= 6) go(k_ab_src<T, NVS, 6, 6, R, SPLIT, DB, NW, TRI>);
                else if constexpr (sizeof(T) == 8 && NVS == 2) return false;  // 128 FP64 accumulators: spills
                else if (kB <= 8 && vB <= 8) go(k_ab_src<T, NVS, 8, 8, R, SPLIT, DB, NW, TRI>);
                else return false;
                return true;
            };
            using I1 = std::integral_constant<int, 1>;
            using I2 = std::integral_constant<int, 2>;
            if constexpr (TRI) {
                return pick(I1{}, I1{});
            } else {
                if (split == 2) return db == 2 ? pick(I2{}, I2{}) : pick(I2{}, I1{});
                return db == 2 ? pick(I1{}, I2{}) : pick(I1{}, I1{});
            }
        }
    }
}

// the matrix-core kernel (ab_mma.hpp) for vector v0: up to 16 orbitals (16-bit strings), W and
// the FP64 destination elements in LDS -- FP32 the whole row, FP64 one class of destinations
// per launch (split by the top orbital); not in the deterministic mode (its LDS adds come from
// all waves at once)
template <typename T>
static bool try_ab_mma(DevHam *H, const T *xt, T *s1, int nvst, int v0, const T *eri, hipStream_t st) {
    constexpr int prec = sizeof(T) == 4;
    if (deterministic(prec)) return false;
    DevBasis *B = H->basis;
    const DevSpin &A = *B->pA, &Bs = *B->pB;
    const int n = H->n;
    constexpr bool CLS = sizeof(T) == 8;
    constexpr int MT = 9, KSP = 17, NWV = CLS ? AB_MMA_NWV64 : AB_MMA_NWV32;
    if (n > 16 || n < 1 || A.nsx + 1 > 4 * KSP || A.nsx + 1 > AB_SRC_CONN || Bs.N < 1) return false;
    const i64 NB = Bs.N;
    if ((u64)A.N * (u64)NB * (u64)nvst >= (1ull << 32)) return false;  // 32-bit element offsets
    // class N: the beta strings without orbital n - 1, the first C(n - 1, k) in colex order
    i64 NN = NB;
    if (CLS) {
        NN = 1;
        for (int i = 0; i < Bs.k; ++i) NN = NN * (n - 1 - i) / (i + 1);
        if (Bs.k > n - 1) NN = 0;
    }
    const size_t smem = ab_mma_lds<T, MT, KSP, NWV>(CLS ? std::max(NN, NB - NN) : NB);
    if (smem > std::min<size_t>(lds_budget(), 160 * 1024)) return false;
    const i64 nblocks = (A.N + 7) / 8 * 8;
    HIPOTF_REQUIRE(nblocks < (1ll << 31), "alpha-beta grid too large");
    const T *dw = diag_w_table<T>(H, st);
    const int nks = (A.nsx + 1 + 3) / 4;
    auto kern = k_ab_mma<T, MT, KSP, NWV, CLS>;
    HIPOTF_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    for (int cls = 0; cls < (CLS ? 2 : 1); ++cls) {
        if (CLS && (cls ? NB - NN : NN) == 0) continue;
        hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(NWV * 64), smem, st, xt, s1, nvst, v0, A.sx.as<u32>(),
                           Bs.strs.as<u64>(), eri, dw, ab_src_nw(n), n, A.N, (int)NB, A.nsx, nks, cls, (int)NN);
        HIPOTF_CHECK(hipGetLastError());
        ++ab_timing().mma_launches;
    }
    return true;
}

// the class-split FP32 kernel for vectors v0, v0 + 1 (ab_cls.hpp): 256-entry integral rows
// (n <= 16), 1..8 beta electrons and holes, one class's FP64 destination elements of both
// vectors in LDS; two launches (class T, class N)
static bool try_ab_cls(DevHam *H, const float *xt, float *s1, int nvst, int v0, const float *eri,
                       hipStream_t st) {
    DevBasis *B = H->basis;
    const DevSpin &A = *B->pA, &Bs = *B->pB;
    const int n = H->n, kB = Bs.k, vB = Bs.n - Bs.k;
    if (n * n > 256 || kB < 1 || vB < 1 || kB > 8 || vB > 8 || A.nsx > AB_SRC_CONN || n * 10 > AB_SRC_BINOM)
        return false;
    HIPOTF_REQUIRE(256 - n * n <= ERI_PAD, "internal: integral table padding too short for the row width");
    constexpr int R = 32, NVS = 2;
    const i64 NB = Bs.N, nN = NB * vB / n, nT = NB - nN;
    const size_t smem = (256 * R + 16) * sizeof(float) + (size_t)std::max(nN, nT) * NVS * sizeof(double) +
                        (AB_SRC_BINOM + AB_SRC_CONN) * 4;
    if (smem > std::min<size_t>(lds_budget(), 160 * 1024)) return false;
    const i64 nblocks = (A.N + 7) / 8 * 8;
    HIPOTF_REQUIRE(nblocks < (1ll << 31), "alpha-beta grid too large");
    const float *dw = diag_w_table<float>(H, st);
    auto go = [&](auto kern, int ko, int vo, int kf) {
        HIPOTF_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(AB_SRC_NJ), smem, st, xt, s1, nvst, v0,
                           A.sx.as<u32>(), Bs.strs.as<u64>(), eri, dw, n, A.N, (int)NB, A.nsx, kB, ko, vo, kf);
        HIPOTF_CHECK(hipGetLastError());
    };
    // class T: own slots (k-1) x (n-k), foreign k; class N: own k x (n-k-1), foreign n-k
    if (kB == 8 && vB == 8) {
        go(k_ab_cls<NVS, 7, 8, 8, R, true, true>, 7, 8, 8);
        go(k_ab_cls<NVS, 8, 7, 8, R, false, true>, 8, 7, 8);
    } else {
        go(k_ab_cls<NVS, 8, 8, 8, R, true, false>, kB - 1, vB, kB);
        go(k_ab_cls<NVS, 8, 8, 8, R, false, false>, kB, vB - 1, vB);
    }
    return true;
}

template <typename T, int NVS>
static void launch_ab_nvs(DevHam *H, const T *xt, T *s1, int nvst, int v0, const T *eri,
                          hipStream_t st) {
    DevBasis *B = H->basis;
    const i64 NB = B->pB->N;
    const int n2 = H->n * H->n;
    // register-descriptor kernel whenever a whole source row fits LDS
    if (ab_kernel_pref(sizeof(T) == 4) != 3) {
        const DevSpin &A = *B->pA, &Bs = *B->pB;
        const size_t wb = (size_t)((n2 + H->n + 15) & ~15) * sizeof(T);
        const size_t smem = wb + (size_t)NB * NVS * sizeof(T);
        const size_t budget = std::min<size_t>(lds_budget(), 160 * 1024) - 1024;
        if (smem <= budget && Bs.nsx <= 80) {
            const int bd = NB >= 1024 ? 1024 : (int)std::max<i64>(64, (NB + 63) / 64 * 64);
            const int nsplit = (int)((NB + bd - 1) / bd);
            const i64 per = (A.N + 7) / 8;
            const i64 nblocks = per * nsplit * 8;
            HIPOTF_REQUIRE(nblocks < (1ll << 31), "alpha-beta grid too large");
            auto go = [&](auto kern) {
                HIPOTF_CHECK(hipFuncSetAttribute((const void *)kern,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
                hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(bd), smem, st, xt, s1, nvst, v0,
                                   A.sx.as<u32>(), Bs.sx.as<u32>(), A.strs.as<u64>(), Bs.strs.as<u64>(),
                                   eri, H->n, A.N, Bs.N, A.nsx, Bs.nsx, nsplit);
                HIPOTF_CHECK(hipGetLastError());
            };
            if (Bs.nsx <= 16) go(k_ab_reg<T, NVS, 16>);
            else if (Bs.nsx <= 40) go(k_ab_reg<T, NVS, 40>);
            else if (Bs.nsx <= 64) go(k_ab_reg<T, NVS, 64>);
            else go(k_ab_reg<T, NVS, 80>);
            return;
        }
    }
    const size_t wbytes = (size_t)((n2 + 15) & ~15) * sizeof(T);
    const size_t budget = std::min<size_t>(lds_budget(), 160 * 1024) - 1024;
    HIPOTF_REQUIRE(budget > wbytes + 64 * NVS * sizeof(T), "LDS too small for the alpha-beta kernel");
    i64 chunk = (i64)((budget - wbytes) / (NVS * sizeof(T)));
    if (chunk >= NB) chunk = NB;
    const int nchunks = (int)((NB + chunk - 1) / chunk);
    const size_t smem = wbytes + (size_t)chunk * NVS * sizeof(T);
    const int bd = NB >= 4096 ? 1024 : (NB >= 1024 ? 512 : 256);
    constexpr int EM = ab_eptmax<T, NVS>();
    const int nsplit = (int)((NB + (i64)bd * EM - 1) / ((i64)bd * EM));
    const i64 per_split = (NB + nsplit - 1) / nsplit;
    const int ept = (int)((per_split + bd - 1) / bd);
    if (ept <= 4) launch_ab_ept<T, NVS, 4>(H, xt, s1, nvst, v0, eri, bd, ept, chunk, nchunks, nsplit, per_split, smem, st);
    else launch_ab_ept<T, NVS, EM>(H, xt, s1, nvst, v0, eri, bd, ept, chunk, nchunks, nsplit, per_split, smem, st);
}

// processes the nv interleaved vectors in sub-blocks of NVS so that a staged source row fits
// LDS and the per-lane accumulators stay within ~64 VGPRs
template <typename T>
static void launch_ab(DevHam *H, const T *xt, T *s1, int nv, const T *eri, hipStream_t st) {
    DevBasis *B = H->basis;
    const i64 NB = B->pB->N;
    if (B->pA->N == 0 || NB == 0) return;
    const size_t budget = std::min<size_t>(lds_budget(), 160 * 1024) - 4096;
    auto fits = [&](int s) { return (size_t)NB * s * sizeof(T) <= budget; };
    int v0 = 0;
    const int pref = ab_kernel_pref(sizeof(T) == 4);
    if (pref <= 1 || pref >= 4) {
        while (v0 < nv) {
            if (ab_mma_first(pref, H->n) && try_ab_mma<T>(H, xt, s1, nv, v0, eri, st)) {
                v0 += 1;
                continue;
            }
            if constexpr (sizeof(T) == 4) {
                if ((pref == 4 || pref == 0) && nv - v0 >= 2 && try_ab_cls(H, xt, s1, nv, v0, eri, st)) {
                    v0 += 2;
                    continue;
                }
            }
            const bool one = tuning(sizeof(T) == 4).ab_nvs == 1;
            if (!one && nv - v0 >= 2 && try_ab_src<T, 2>(H, xt, s1, nv, v0, eri, st)) { v0 += 2; continue; }
            if (try_ab_src<T, 1>(H, xt, s1, nv, v0, eri, st)) { v0 += 1; continue; }
            break;
        }
    }
    int nvs = 1;
    if (nv >= 4 && fits(4)) nvs = 4;
    else if (nv >= 2 && fits(2)) nvs = 2;
    while (v0 < nv) {
        const int m = std::min(nvs, nv - v0);
        if (m >= 4) { launch_ab_nvs<T, 4>(H, xt, s1, nv, v0, eri, st); v0 += 4; }
        else if (m >= 2) { launch_ab_nvs<T, 2>(H, xt, s1, nv, v0, eri, st); v0 += 2; }
        else { launch_ab_nvs<T, 1>(H, xt, s1, nv, v0, eri, st); v0 += 1; }
    }
}

// work: NULL (the basis's own work buffers, grown as needed) or 4 consecutive 256-byte aligned
// regions of round_up(dim * NV * sizeof(T), 256) bytes each
template <typename T, int NV>
static void sigma_group(DevHam *H, const SrcPack &src, const DstPack &dst, hipStream_t st,
                        unsigned char *work = nullptr) {
    DevBasis *B = H->basis;
    const DevSpin &A = *B->pA, &Bs = *B->pB;
    const i64 NA = A.N, NB = Bs.N, dim = NA * NB;
    if (dim == 0) return;
    const size_t bytes = (size_t)dim * NV * sizeof(T);
    T *xt, *xtT, *s1, *s2T;
    if (work) {
        const size_t stride = (bytes + 255) / 256 * 256;
        xt = reinterpret_cast<T *>(work);
        xtT = reinterpret_cast<T *>(work + stride);
        s1 = reinterpret_cast<T *>(work + 2 * stride);
        s2T = reinterpret_cast<T *>(work + 3 * stride);
    } else {
        B->xt.ensure(bytes);
        B->xtT.ensure(bytes);
        B->s1.ensure(bytes);
        B->s2T.ensure(bytes);
        xt = B->xt.as<T>(), xtT = B->xtT.as<T>(), s1 = B->s1.as<T>(), s2T = B->s2T.as<T>();
    }
    dim3 g2((unsigned)((NB + 31) / 32), (unsigned)((NA + 31) / 32));
    HIPOTF_REQUIRE(g2.y < 65536, "too many alpha strings for the tiled kernels");
    hipLaunchKernelGGL((k_gauge_in<T, NV>), g2, dim3(256), 0, st, src, xt, xtT, Bs.strs.as<u64>(),
                       B->gaugeM.as<u64>(), NA, NB);
    HIPOTF_CHECK(hipGetLastError());
    const bool f32 = sizeof(T) == 4;
    const T *CA = f32 ? (const T *)H->ptA->c32.p : (const T *)H->ptA->c64.p;
    const T *CB = f32 ? (const T *)H->ptB->c32.p : (const T *)H->ptB->c64.p;
    launch_rowspmm<T>(xt, s1, *H->ptA, CA, NA, NB * NV, st);
    launch_rowspmm<T>(xtT, s2T, *H->ptB, CB, NB, NA * NV, st);
    const T *eri = f32 ? (const T *)H->eri32.p : (const T *)H->eri64.p;
    AbTiming &tm = ab_timing();
    if (tm.on) HIPOTF_CHECK(hipEventRecord(tm.e0, st));
    launch_ab<T>(H, xt, s1, NV, eri, st);
    if (tm.on) {
        HIPOTF_CHECK(hipEventRecord(tm.e1, st));
        HIPOTF_CHECK(hipEventSynchronize(tm.e1));
        float ms = 0.f;
        HIPOTF_CHECK(hipEventElapsedTime(&ms, tm.e0, tm.e1));
        tm.ms += ms;
        ++tm.sections;
    }
    hipLaunchKernelGGL((k_finalize<T, NV>), g2, dim3(256), 0, st, s1, s2T, xt, H->ecore, dst,
                       Bs.strs.as<u64>(), B->gaugeM.as<u64>(), NA, NB);
    HIPOTF_CHECK(hipGetLastError());
}

template <typename T>
static void sigma_t(DevHam *H, int nv, const double *const *src, double *const *dst, hipStream_t st,
                    int maxnv) {
    for (int g0 = 0; g0 < nv; g0 += maxnv) {
        const int m = std::min(maxnv, nv - g0);
        SrcPack sp{};
        DstPack dp{};
        for (int i = 0; i < m; ++i) {
            HIPOTF_REQUIRE(src[g0 + i] && dst[g0 + i], "NULL vector pointer");
            sp.p[i] = src[g0 + i];
            dp.p[i] = dst[g0 + i];
        }
        switch (m) {
            case 1: sigma_group<T, 1>(H, sp, dp, st); break;
            case 2: sigma_group<T, 2>(H, sp, dp, st); break;
            case 3: sigma_group<T, 3>(H, sp, dp, st); break;
            case 4: sigma_group<T, 4>(H, sp, dp, st); break;
            case 5: sigma_group<T, 5>(H, sp, dp, st); break;
            case 6: sigma_group<T, 6>(H, sp, dp, st); break;
            case 7: sigma_group<T, 7>(H, sp, dp, st); break;
            default: sigma_group<T, 8>(H, sp, dp, st); break;
        }
    }
}

void sigma(DevHam *H, int nv, const double *const *src, double *const *dst, int prec, hipStream_t st,
           int maxnv) {
    HIPOTF_REQUIRE(H != nullptr, "Hamiltonian handle is NULL");
    HIPOTF_REQUIRE(nv >= 1, "need at least one vector");
    HIPOTF_REQUIRE(maxnv >= 1 && maxnv <= MAXNV, "vectors per sigma pass out of range");
    if (prec == 0) sigma_t<double>(H, nv, src, dst, st, maxnv);
    else sigma_t<float>(H, nv, src, dst, st, maxnv);
}

// ---------------------------------------------------------------------------------------
// allocation-free sigma: the work buffers live in a static device array (64 MiB per GPU, held
// until the process ends); host code reaches it only through the address hipGetSymbolAddress
// returns for the current device, looked up once per device on first use
// ---------------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static unsigned char *scratch_address() {
    static std::mutex mu;
    static void *addr[256] = {};
    int dev = 0;
    HIPOTF_CHECK(hipGetDevice(&dev));
    HIPOTF_REQUIRE(dev >= 0 && dev < 256, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        HIPOTF_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        HIPOTF_REQUIRE(p != nullptr, "scratch symbol has no device address");
        addr[dev] = p;
    }
    return static_cast<unsigned char *>(addr[dev]);
}

void sigma_noalloc(DevHam *H, const double *src, double *dst, hipStream_t st) {
    HIPOTF_REQUIRE(H != nullptr && src && dst, "NULL argument");
    const i64 dim = H->basis->host.dim;
    HIPOTF_REQUIRE(dim <= NOALLOC_MAX_DIM, "allocation-free sigma: " + std::to_string(dim) +
                                               " determinants exceed its scratch limit of " +
                                               std::to_string(NOALLOC_MAX_DIM));
    const size_t stride = ((size_t)dim * sizeof(double) + 255) / 256 * 256;
    HIPOTF_REQUIRE(4 * stride <= NOALLOC_SCRATCH_BYTES, "internal: scratch layout exceeds the array");
    unsigned char *work = scratch_address();
    ForbidAlloc guard;
    SrcPack sp{};
    DstPack dp{};
    sp.p[0] = src;
    dp.p[0] = dst;
    sigma_group<double, 1>(H, sp, dp, st, work);
}

void diag_elements(DevHam *H, double *diags, hipStream_t st) {
    HIPOTF_REQUIRE(H && diags, "NULL argument");
    DevBasis *B = H->basis;
    const i64 NA = B->pA->N, NB = B->pB->N, dim = NA * NB;
    if (dim == 0) return;
    hipLaunchKernelGGL(k_diag, dim3((unsigned)((dim + 255) / 256)), dim3(256), 0, st, diags,
                       H->ptA->gdiag.as<double>(), H->ptB->gdiag.as<double>(), H->jab.as<double>(),
                       B->pB->strs.as<u64>(), H->n, NA, NB, H->ecore);
    HIPOTF_CHECK(hipGetLastError());
}

}  // namespace hipotf
