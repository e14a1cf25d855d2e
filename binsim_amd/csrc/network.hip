// Generic Pauli-group network operator: the sigma build and the diagonal of an operator given
// only as the reference's SVD network (src/include/ham/otf.hpp:114-169, math.hpp:37-124), for
// the networks that are not a spin-free two-body Hamiltonian (spin models, S^2 penalties,
// three-body terms, ...).  The integral form (sigma.hip) is the fast path; this one evaluates
// the network itself, as the reference's CUDA kernel does (src/include/cuda/hvec.cuh:6-446):
//
//   y[(a, b)] = sum_g f_g(a ^ ax_g, b ^ bx_g) x[(a ^ ax_g, b ^ bx_g)],
//   f_g(s_a, s_b) = sum_r Pa_{g,r}(s_a) Pb_{g,r}(s_b),  P_{g,r}(s) = sum_k w_{g,r,k} (-1)^popc(s & z_{g,k})
//
// with sources outside the basis (a flip that changes an electron count) skipped.  The phase
// tables P_{g,r} over the basis strings are built once per basis; the sum runs with one
// workgroup per destination alpha string, the groups outermost (their alpha source and its
// colex rank are uniform across the workgroup), the beta destinations on the lanes, and the
// destination row accumulated in LDS.
#include "hipotf.hpp"

#include <algorithm>
#include <map>
#include <memory>
#include <vector>

namespace hipotf {

struct GGroup {
    u32 ax, bx;
    int rank, nza, nzb;
    i64 zaoff, zboff, waoff, wboff;  // into the flat z / w arrays
    i64 paoff, pboff;                // into the per-basis phase tables
};

struct GenBasisTables {
    DevBuf groups;  // GGroup[ngs] with this basis's phase-table offsets
    DevBuf pa, pb;  // phases over the alpha / beta strings
    i64 diag_group = -1;
};

struct DevGenNet {
    int n = 0;
    std::vector<GGroup> host;  // offsets into the flat arrays (phase offsets per basis)
    std::vector<u32> fza, fzb;
    std::vector<double> fwa, fwb;
    DevBuf zas, zbs, was, wbs;  // the flat arrays on the device, uploaded at first use
    std::map<i64, std::unique_ptr<GenBasisTables>> tables;
};

constexpr int GEN_BW = 18;  // binomial table: C(p, j), p < 33, j < 18

__device__ __forceinline__ i64 colex_rank(u64 s, const i64 *binom) {
    i64 r = 0;
    int j = 1;
    while (s) {
        const int p = __builtin_ctzll(s);
        s &= s - 1;
        r += binom[p * GEN_BW + j];
        ++j;
    }
    return r;
}

__device__ void fill_binom(i64 *binom) {
    for (int t = threadIdx.x; t < 33 * GEN_BW; t += blockDim.x) {
        const int p = t / GEN_BW, j = t - p * GEN_BW;
        u64 c = j <= p ? 1ull : 0ull;
        for (int i = 0; i < j && c; ++i) c = c * (u64)(p - i) / (u64)(i + 1);
        binom[t] = (i64)c;
    }
}

// P[poff_g + r * Ns + i] for group g = blockIdx.y over the strings i of one spin
template <bool BETA>
__global__ __launch_bounds__(256) void k_gen_phase(const GGroup *__restrict__ G, const u32 *__restrict__ zs,
                                                   const double *__restrict__ ws, const u64 *__restrict__ strs,
                                                   i64 Ns, double *__restrict__ P) {
    const GGroup g = G[blockIdx.y];
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Ns) return;
    const u64 s = strs[i];
    const int nz = BETA ? g.nzb : g.nza;
    const u32 *z = zs + (BETA ? g.zboff : g.zaoff);
    const double *w = ws + (BETA ? g.wboff : g.waoff);
    const i64 poff = BETA ? g.pboff : g.paoff;
    for (int r = 0; r < g.rank; ++r) {
        double v = 0.0;
        for (int k = 0; k < nz; ++k) v += (__builtin_popcountll(s & (u64)z[k]) & 1) ? -w[r * nz + k] : w[r * nz + k];
        P[poff + (i64)r * Ns + i] = v;
    }
}

// one workgroup per destination alpha string; ROW: the destination row accumulates in LDS
template <bool ROW>
__global__ __launch_bounds__(256) void k_gen_hvec(const GGroup *__restrict__ G, int ngs, const double *__restrict__ PA,
                                                  const double *__restrict__ PB, const u64 *__restrict__ strsA,
                                                  const u64 *__restrict__ strsB, i64 NA, i64 NB, int n, int na,
                                                  int nb, const double *__restrict__ x, double *__restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    i64 *binom = reinterpret_cast<i64 *>(smem_raw);
    double *row = reinterpret_cast<double *>(binom + 33 * GEN_BW);
    const i64 a = blockIdx.x;
    const int tid = threadIdx.x, bd = blockDim.x;
    fill_binom(binom);
    for (i64 ib = tid; ib < NB; ib += bd) {
        if (ROW) row[ib] = 0.0;
        else y[a * NB + ib] = 0.0;
    }
    __syncthreads();
    const u64 sA = strsA[a];
    for (int gi = 0; gi < ngs; ++gi) {
        const GGroup g = G[gi];
        const u64 ta = sA ^ (u64)g.ax;
        // uniform: the alpha source is not in the basis (an electron count changed, or a flip
        // beyond the n orbitals)
        if (__builtin_popcountll(ta) != na || (ta >> n)) continue;
        const i64 ra = colex_rank(ta, binom);
        const double *pa = PA + g.paoff + ra;
        const double *pb = PB + g.pboff;
        for (i64 ib = tid; ib < NB; ib += bd) {
            const u64 tb = strsB[ib] ^ (u64)g.bx;
            if (__builtin_popcountll(tb) != nb || (tb >> n)) continue;
            const i64 rb = colex_rank(tb, binom);
            double f = 0.0;
            for (int r = 0; r < g.rank; ++r) f += pa[(i64)r * NA] * pb[(i64)r * NB + rb];
            const double v = f * x[ra * NB + rb];
            if (ROW) row[ib] += v;
            else y[a * NB + ib] += v;
        }
    }
    if (ROW) {
        __syncthreads();
        for (i64 ib = tid; ib < NB; ib += bd) y[a * NB + ib] = row[ib];
    }
}

// diag[(a, b)] += f_g(a, b) of the first group with no flips (src/include/op/hvec.hpp:498-554)
__global__ __launch_bounds__(256) void k_gen_diag(const GGroup *__restrict__ G, i64 gd, const double *__restrict__ PA,
                                                  const double *__restrict__ PB, i64 NA, i64 NB,
                                                  double *__restrict__ diag) {
    const GGroup g = G[gd];
    const i64 dim = NA * NB;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += (i64)gridDim.x * blockDim.x) {
        const i64 a = i / NB, b = i - a * NB;
        double f = 0.0;
        for (int r = 0; r < g.rank; ++r) f += PA[g.paoff + (i64)r * NA + a] * PB[g.pboff + (i64)r * NB + b];
        diag[i] += f;
    }
}

DevGenNet *make_dev_gennet(int n, const std::vector<u32> &ax, const std::vector<u32> &bx,
                           const std::vector<int> &rank, const std::vector<std::vector<u32>> &zas,
                           const std::vector<std::vector<u32>> &zbs, const std::vector<std::vector<double>> &was,
                           const std::vector<std::vector<double>> &wbs) {
    auto D = std::make_unique<DevGenNet>();
    D->n = n;
    for (size_t g = 0; g < ax.size(); ++g) {
        GGroup G{};
        G.ax = ax[g];
        G.bx = bx[g];
        G.rank = rank[g];
        G.nza = (int)zas[g].size();
        G.nzb = (int)zbs[g].size();
        G.zaoff = (i64)D->fza.size();
        G.zboff = (i64)D->fzb.size();
        G.waoff = (i64)D->fwa.size();
        G.wboff = (i64)D->fwb.size();
        D->fza.insert(D->fza.end(), zas[g].begin(), zas[g].end());
        D->fzb.insert(D->fzb.end(), zbs[g].begin(), zbs[g].end());
        D->fwa.insert(D->fwa.end(), was[g].begin(), was[g].end());
        D->fwb.insert(D->fwb.end(), wbs[g].begin(), wbs[g].end());
        D->host.push_back(G);
    }
    return D.release();  // host data only: the device copies are made with the first basis
}

void destroy_dev_gennet(DevGenNet *D) { delete D; }

static GenBasisTables &tables_for(DevGenNet *D, DevBasis *B, hipStream_t st) {
    auto it = D->tables.find(B->uid);
    if (it != D->tables.end()) return *it->second;
    HIPOTF_REQUIRE(D->n == B->host.norb, "orbital count of the network does not match the basis");
    const i64 NA = B->pA->N, NB = B->pB->N;
    if (!D->zas.p) {
        auto up = [](DevBuf &b, const void *p, size_t bytes) {
            b.alloc(std::max<size_t>(bytes, 16));
            if (bytes) HIPOTF_CHECK(hipMemcpy(b.p, p, bytes, hipMemcpyHostToDevice));
        };
        up(D->zas, D->fza.data(), D->fza.size() * sizeof(u32));
        up(D->zbs, D->fzb.data(), D->fzb.size() * sizeof(u32));
        up(D->was, D->fwa.data(), D->fwa.size() * sizeof(double));
        up(D->wbs, D->fwb.data(), D->fwb.size() * sizeof(double));
    }
    auto T = std::make_unique<GenBasisTables>();
    std::vector<GGroup> g = D->host;
    i64 pa = 0, pb = 0;
    for (size_t i = 0; i < g.size(); ++i) {
        g[i].paoff = pa;
        g[i].pboff = pb;
        pa += (i64)g[i].rank * NA;
        pb += (i64)g[i].rank * NB;
        if (T->diag_group < 0 && g[i].ax == 0 && g[i].bx == 0) T->diag_group = (i64)i;
    }
    T->groups.alloc(std::max<size_t>(g.size(), 1) * sizeof(GGroup));
    if (!g.empty()) HIPOTF_CHECK(hipMemcpy(T->groups.p, g.data(), g.size() * sizeof(GGroup), hipMemcpyHostToDevice));
    T->pa.alloc(std::max<i64>(pa, 1) * sizeof(double));
    T->pb.alloc(std::max<i64>(pb, 1) * sizeof(double));
    const unsigned ng = (unsigned)g.size();
    HIPOTF_REQUIRE(g.size() < 65536, "network has too many groups for the phase-table launch");
    if (ng && NA) {
        hipLaunchKernelGGL(k_gen_phase<false>, dim3((unsigned)((NA + 255) / 256), ng), dim3(256), 0, st,
                           T->groups.as<GGroup>(), D->zas.as<u32>(), D->was.as<double>(), B->pA->strs.as<u64>(), NA,
                           T->pa.as<double>());
        HIPOTF_CHECK(hipGetLastError());
    }
    if (ng && NB) {
        hipLaunchKernelGGL(k_gen_phase<true>, dim3((unsigned)((NB + 255) / 256), ng), dim3(256), 0, st,
                           T->groups.as<GGroup>(), D->zbs.as<u32>(), D->wbs.as<double>(), B->pB->strs.as<u64>(), NB,
                           T->pb.as<double>());
        HIPOTF_CHECK(hipGetLastError());
    }
    auto &ref = *T;
    D->tables[B->uid] = std::move(T);
    return ref;
}

void gennet_hvec(DevGenNet *D, DevBasis *B, const double *x, double *y, hipStream_t st) {
    HIPOTF_REQUIRE(x && y, "NULL vector");
    const i64 NA = B->pA->N, NB = B->pB->N;
    if (NA * NB == 0) return;
    GenBasisTables &T = tables_for(D, B, st);
    const int na = B->pA->k, nb = B->pB->k;
    const size_t head = 33 * GEN_BW * sizeof(i64);
    const size_t row = (size_t)NB * sizeof(double);
    const size_t lds_max = 150 * 1024;
    const int ngs = (int)D->host.size();
    if (head + row <= lds_max) {
        const size_t smem = head + row;
        HIPOTF_CHECK(hipFuncSetAttribute((const void *)k_gen_hvec<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)smem));
        hipLaunchKernelGGL(k_gen_hvec<true>, dim3((unsigned)NA), dim3(256), smem, st, T.groups.as<GGroup>(), ngs,
                           T.pa.as<double>(), T.pb.as<double>(), B->pA->strs.as<u64>(), B->pB->strs.as<u64>(), NA,
                           NB, D->n, na, nb, x, y);
    } else {
        hipLaunchKernelGGL(k_gen_hvec<false>, dim3((unsigned)NA), dim3(256), head, st, T.groups.as<GGroup>(), ngs,
                           T.pa.as<double>(), T.pb.as<double>(), B->pA->strs.as<u64>(), B->pB->strs.as<u64>(), NA,
                           NB, D->n, na, nb, x, y);
    }
    HIPOTF_CHECK(hipGetLastError());
}

void gennet_diag(DevGenNet *D, DevBasis *B, double *diag, hipStream_t st) {
    HIPOTF_REQUIRE(diag, "NULL vector");
    const i64 NA = B->pA->N, NB = B->pB->N;
    if (NA * NB == 0) return;
    GenBasisTables &T = tables_for(D, B, st);
    if (T.diag_group < 0) return;  // no diagonal group: nothing to add
    const i64 dim = NA * NB;
    const int nbk = (int)std::min<i64>((dim + 255) / 256, 4096);
    hipLaunchKernelGGL(k_gen_diag, dim3(nbk), dim3(256), 0, st, T.groups.as<GGroup>(), T.diag_group,
                       T.pa.as<double>(), T.pb.as<double>(), NA, NB, diag);
    HIPOTF_CHECK(hipGetLastError());
}

}  // namespace hipotf
