// Davidson inner-loop kernels (libdiag on the GPU).
//
// Same entry points and arithmetic as src/include/diag/davidson.hpp:31-272 (C ABI
// src/include/diag/declare.hpp:4-75); state vectors are device pointers, the small arrays
// (Ritz coefficients, heff) stay on the host as in jl/davidson.jl.  Reductions are two-stage
// (per-block partials, then one block) so results are deterministic run to run.
#include "hipotf.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace hipotf {

constexpr int DMAX = 64;  // subspace cap, as local_V[64] in davidson.hpp:41-43

struct VecPack {
    const double *v[DMAX];
    double c[DMAX];
};

__device__ __forceinline__ double block_sum_d(double v) {
    __shared__ double red[16];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)((blockDim.x + 63) >> 6); ++i) t += red[i];
    return t;
}

static int nblk(i64 n) { return (int)std::max<i64>(1, std::min<i64>((n + 255) / 256, 2048)); }

// f(std::integral_constant<int, c>) for c = k in 1..8 (k > 8 -> 8)
template <typename F>
static void dispatch_m8(int k, F &&f) {
    switch (k) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    default: f(std::integral_constant<int, 8>{}); break;
    }
}

// Streaming kernels keep several independent loads in flight per thread: the vector loops load
// 4 vectors' elements before using them, the single-vector loops 4 elements (256 apart) per step.
constexpr int VU = 4;

// y = sum_j c_j V_j ; with RES: r = sum_j c_j AV_j - e y (written to y), partial |r|^2
// (the sum runs over j in order, as before the loads were batched)
template <bool RES>
__global__ __launch_bounds__(256) void k_combine(VecPack V, VecPack AV, int dim, double e, double *__restrict__ out,
                                                 double *__restrict__ partial, i64 n) {
    double acc = 0.0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        double x = 0.0, ax = 0.0;
        int j = 0;
        for (; j + VU <= dim; j += VU) {
            double xv[VU], av[VU];
#pragma unroll
            for (int u = 0; u < VU; ++u) {
                xv[u] = V.v[j + u][i];
                if (RES) av[u] = AV.v[j + u][i];
            }
#pragma unroll
            for (int u = 0; u < VU; ++u) {
                x += V.c[j + u] * xv[u];
                if (RES) ax += V.c[j + u] * av[u];
            }
        }
        for (; j < dim; ++j) {
            x += V.c[j] * V.v[j][i];
            if (RES) ax += V.c[j] * AV.v[j][i];
        }
        const double r = RES ? ax - e * x : x;
        out[i] = r;
        acc += r * r;
    }
    const double t = block_sum_d(acc);
    if (threadIdx.x == 0 && partial) partial[blockIdx.x] = t;
}

// Calls f(i) for every i < n, VU indices per thread and step: the full steps are unrolled so
// their loads issue together, the last partial step is checked per index.
template <typename F>
__device__ __forceinline__ void stream_for(i64 n, F &&f) {
    const i64 step = (i64)gridDim.x * blockDim.x * VU;
    for (i64 base = (i64)blockIdx.x * blockDim.x * VU + threadIdx.x; base < n; base += step) {
        if (base + (i64)(VU - 1) * blockDim.x < n) {
#pragma unroll
            for (int u = 0; u < VU; ++u) f(base + (i64)u * blockDim.x);
        } else {
            for (int u = 0; u < VU; ++u)
                if (base + (i64)u * blockDim.x < n) f(base + (i64)u * blockDim.x);
        }
    }
}

__global__ __launch_bounds__(256) void k_reduce(const double *__restrict__ partial, int n,
                                                double *__restrict__ out, int mode) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
    const double t = block_sum_d(acc);
    if (threadIdx.x == 0) *out = mode == 1 ? sqrt(t) : t;
}

__global__ __launch_bounds__(256) void k_precond(double *__restrict__ axt, const double *__restrict__ diags,
                                                 double e, double shift, i64 n) {
    const double as = fabs(shift);
    stream_for(n, [&](i64 i) {
        double d = diags[i] - e;
        if (fabs(d) < as) d = d >= 0 ? as : -as;
        axt[i] = axt[i] / d;
    });
}

__global__ __launch_bounds__(256) void k_dotp(const double *__restrict__ a, const double *__restrict__ b,
                                              double *__restrict__ partial, i64 n) {
    double acc = 0.0;
    stream_for(n, [&](i64 i) { acc += a[i] * b[i]; });
    const double t = block_sum_d(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// y -= (*proj) x
__global__ __launch_bounds__(256) void k_axpy_dev(double *__restrict__ y, const double *__restrict__ x,
                                                  const double *__restrict__ proj, i64 n) {
    const double p = *proj;
    stream_for(n, [&](i64 i) { y[i] -= p * x[i]; });
}

// y *= 1 / *nrm
__global__ __launch_bounds__(256) void k_scale_inv(double *__restrict__ y, const double *__restrict__ nrm,
                                                   i64 n) {
    const double s = 1.0 / *nrm;
    stream_for(n, [&](i64 i) { y[i] *= s; });
}

// out[j] = <V_j, a> for j < M (M <= 8): each block accumulates M partials
template <int M>
__global__ __launch_bounds__(256) void k_multidot(VecPack V, const double *__restrict__ a,
                                                  double *__restrict__ partial, int ldp, i64 n) {
    double acc[M];
#pragma unroll
    for (int j = 0; j < M; ++j) acc[j] = 0.0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const double av = a[i];
        double v[M];
#pragma unroll
        for (int j = 0; j < M; ++j) v[j] = V.v[j][i];
#pragma unroll
        for (int j = 0; j < M; ++j) acc[j] += v[j] * av;
    }
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const double t = block_sum_d(acc[j]);
        if (threadIdx.x == 0) partial[j * ldp + blockIdx.x] = t;
    }
}

struct DiagScratch {
    DevBuf partial, scal;
    DiagScratch() {
        partial.alloc(8 * 2048 * sizeof(double));
        scal.alloc(128 * sizeof(double));
    }
};

static DiagScratch &scratch() {
    static thread_local std::unique_ptr<DiagScratch> s;
    if (!s) s.reset(new DiagScratch());
    return *s;
}

static double d2h(const double *d, hipStream_t st) {
    double h = 0.0;
    HIPOTF_CHECK(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPOTF_CHECK(hipStreamSynchronize(st));
    return h;
}

static void pack(VecPack &P, const double *const *V, const double *coeffs, int dim) {
    HIPOTF_REQUIRE(dim >= 0 && dim <= DMAX, "subspace dimension must be in [0, 64]");
    for (int j = 0; j < dim; ++j) {
        HIPOTF_REQUIRE(V[j] != nullptr, "NULL subspace vector");
        P.v[j] = V[j];
        P.c[j] = coeffs ? coeffs[j] : 0.0;
    }
}

double ritz_and_residual(i64 N, int dim, const double *coeffs, double min_e, const double *const *V,
                         const double *const *AV, double *axt, hipStream_t st) {
    VecPack P{}, Q{};
    pack(P, V, coeffs, dim);
    pack(Q, AV, coeffs, dim);
    auto &S = scratch();
    const int nb = nblk(N);
    hipLaunchKernelGGL(k_combine<true>, dim3(nb), dim3(256), 0, st, P, Q, dim, min_e, axt,
                       S.partial.as<double>(), N);
    HIPOTF_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, st, S.partial.as<double>(), N > 0 ? nb : 0,
                       S.scal.as<double>(), 1);
    HIPOTF_CHECK(hipGetLastError());
    return d2h(S.scal.as<double>(), st);
}

void ritz_vector(i64 N, int dim, const double *coeffs, const double *const *V, double *xt,
                 hipStream_t st) {
    VecPack P{}, Q{};
    pack(P, V, coeffs, dim);
    hipLaunchKernelGGL(k_combine<false>, dim3(nblk(N)), dim3(256), 0, st, P, Q, dim, 0.0, xt,
                       (double *)nullptr, N);
    HIPOTF_CHECK(hipGetLastError());
}

void precondition(i64 N, double *axt, const double *diags, double min_e, double shift, hipStream_t st) {
    hipLaunchKernelGGL(k_precond, dim3(nblk(N)), dim3(256), 0, st, axt, diags, min_e, shift, N);
    HIPOTF_CHECK(hipGetLastError());
}

// dot into device scalar dres
void dot_dev(i64 N, const double *a, const double *b, double *dres, hipStream_t st) {
    auto &S = scratch();
    const int nb = nblk(N);
    hipLaunchKernelGGL(k_dotp, dim3(nb), dim3(256), 0, st, a, b, S.partial.as<double>(), N);
    HIPOTF_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, st, S.partial.as<double>(), N > 0 ? nb : 0,
                       dres, 0);
    HIPOTF_CHECK(hipGetLastError());
}

double mgs(i64 N, int dim, const double *const *V, double *vt, hipStream_t st) {
    auto &S = scratch();
    double *proj = S.scal.as<double>() + 8;
    for (int j = 0; j < dim; ++j) {
        HIPOTF_REQUIRE(V[j] != nullptr, "NULL subspace vector");
        dot_dev(N, V[j], vt, proj, st);
        hipLaunchKernelGGL(k_axpy_dev, dim3(nblk(N)), dim3(256), 0, st, vt, V[j], proj, N);
        HIPOTF_CHECK(hipGetLastError());
    }
    double *nrm = S.scal.as<double>() + 9;
    dot_dev(N, vt, vt, nrm, st);
    return std::sqrt(d2h(nrm, st));
}

double normalize(i64 N, double *vec, hipStream_t st) {
    auto &S = scratch();
    double *nsq = S.scal.as<double>() + 10;
    dot_dev(N, vec, vec, nsq, st);
    const double nrm = std::sqrt(d2h(nsq, st));
    // scale by 1/nrm computed on the host, as inplace_normalize does (davidson.hpp:150-169)
    double *dn = S.scal.as<double>() + 11;
    HIPOTF_CHECK(hipMemcpyAsync(dn, &nrm, sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_scale_inv, dim3(nblk(N)), dim3(256), 0, st, vec, dn, N);
    HIPOTF_CHECK(hipGetLastError());
    HIPOTF_CHECK(hipStreamSynchronize(st));
    return nrm;
}

double dot(i64 N, const double *a, const double *b, hipStream_t st) {
    auto &S = scratch();
    double *d = S.scal.as<double>() + 12;
    dot_dev(N, a, b, d, st);
    return d2h(d, st);
}

// out[j] = <V_j, a>, j < dim, written to host memory
void multidot(i64 N, int dim, const double *const *V, const double *a, double *out_host,
              hipStream_t st) {
    auto &S = scratch();
    const int nb = nblk(N);
    std::vector<double> res(dim);
    for (int j0 = 0; j0 < dim; j0 += 8) {
        const int m = std::min(8, dim - j0);
        VecPack P{};
        pack(P, V + j0, nullptr, m);
        dispatch_m8(m, [&](auto mc) {
            hipLaunchKernelGGL(k_multidot<decltype(mc)::value>, dim3(nb), dim3(256), 0, st, P, a,
                               S.partial.as<double>(), 2048, N);
        });
        HIPOTF_CHECK(hipGetLastError());
        for (int j = 0; j < m; ++j) {
            hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, st, S.partial.as<double>() + j * 2048,
                               N > 0 ? nb : 0, S.scal.as<double>() + 16 + j, 0);
            HIPOTF_CHECK(hipGetLastError());
        }
        HIPOTF_CHECK(hipMemcpyAsync(res.data() + j0, S.scal.as<double>() + 16, m * sizeof(double),
                                    hipMemcpyDeviceToHost, st));
        HIPOTF_CHECK(hipStreamSynchronize(st));
    }
    std::copy(res.begin(), res.end(), out_host);
}

// ------------------------------------------------------------------------------------------
// Block subspace operations of the mixed-precision block Davidson (no reference counterpart:
// jl/davidson.jl works one vector at a time).  Each reads every subspace vector once per call
// instead of once per (root, vector) pair; reductions stay two-stage and deterministic.
// ------------------------------------------------------------------------------------------
constexpr int BK = 8;  // outputs per launch

struct Ptr64 { const double *p[DMAX]; };
struct Out8 { double *p[BK]; };

// out_r = sum_j C[j*ldc + r] X_j  (r < nout <= BK); with Y: out_r = sum_j C[j*ldc+r] (Y_j - w_r X_j)
// and partial[r * nblk + b] = block sums of out_r^2
// (the sums run over j in order; the loads of VU vectors are issued together)
template <int KO, bool HY>
__global__ __launch_bounds__(256) void k_block_combo(Ptr64 X, Ptr64 Y, int m, const double *__restrict__ C, int ldc,
                                                     const double *__restrict__ w, Out8 out,
                                                     double *__restrict__ partial, i64 n) {
    double nrm[KO];
#pragma unroll
    for (int r = 0; r < KO; ++r) nrm[r] = 0.0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        double x[KO], y[KO];
#pragma unroll
        for (int r = 0; r < KO; ++r) x[r] = y[r] = 0.0;
        auto step = [&](int j, double xv, double yv) {
#pragma unroll
            for (int r = 0; r < KO; ++r) {
                const double c = C[j * ldc + r];
                x[r] += c * xv;
                if (HY) y[r] += c * yv;
            }
        };
        int j = 0;
        for (; j + VU <= m; j += VU) {
            double xv[VU], yv[VU];
#pragma unroll
            for (int u = 0; u < VU; ++u) {
                xv[u] = X.p[j + u][i];
                yv[u] = HY ? Y.p[j + u][i] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < VU; ++u) step(j + u, xv[u], yv[u]);
        }
        for (; j < m; ++j) step(j, X.p[j][i], HY ? Y.p[j][i] : 0.0);
#pragma unroll
        for (int r = 0; r < KO; ++r) {
            const double o = HY ? y[r] - w[r] * x[r] : x[r];
            out.p[r][i] = o;
            nrm[r] += o * o;
        }
    }
    if (partial) {
#pragma unroll
        for (int r = 0; r < KO; ++r) {
            const double t = block_sum_d(nrm[r]);
            if (threadIdx.x == 0) partial[r * gridDim.x + blockIdx.x] = t;
        }
    }
}

// T_j -= sum_i P[i*ldp + j] V_i   (j < nt <= BK, i < m <= DMAX)
template <int KO>
__global__ __launch_bounds__(256) void k_block_subtract(Ptr64 V, int m, const double *__restrict__ P, int ldp,
                                                        Out8 T, i64 n) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        double acc[KO];
#pragma unroll
        for (int r = 0; r < KO; ++r) acc[r] = 0.0;
        int j = 0;
        for (; j + VU <= m; j += VU) {
            double v[VU];
#pragma unroll
            for (int u = 0; u < VU; ++u) v[u] = V.p[j + u][i];
#pragma unroll
            for (int u = 0; u < VU; ++u)
#pragma unroll
                for (int r = 0; r < KO; ++r) acc[r] += P[(j + u) * ldp + r] * v[u];
        }
        for (; j < m; ++j) {
            const double v = V.p[j][i];
#pragma unroll
            for (int r = 0; r < KO; ++r) acc[r] += P[j * ldp + r] * v;
        }
#pragma unroll
        for (int r = 0; r < KO; ++r) T.p[r][i] -= acc[r];
    }
}

// partial[(a * NB + b) * nblk + blk] = block sums of A_a . B_b   (a < NA, b < NB, NA * NB <= 32);
// every pointer is loaded (the host repeats a vector to fill a partial group)
constexpr int DOT_ACC = 32;
template <int NA, int NB>
__global__ __launch_bounds__(256) void k_block_dots(Ptr64 A, Ptr64 B, double *__restrict__ partial, i64 n) {
    static_assert(NA * NB <= DOT_ACC, "accumulator budget");
    double acc[NA][NB];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = 0.0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        double bv[NB], av[NA];
#pragma unroll
        for (int b = 0; b < NB; ++b) bv[b] = B.p[b][i];
#pragma unroll
        for (int a = 0; a < NA; ++a) av[a] = A.p[a][i];
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b) acc[a][b] += av[a] * bv[b];
    }
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const double t = block_sum_d(acc[a][b]);
            if (threadIdx.x == 0) partial[(a * NB + b) * gridDim.x + blockIdx.x] = t;
        }
}

// out[o] = sum of partial[o * nparts .. + nparts) for o = blockIdx.x
__global__ __launch_bounds__(256) void k_multi_reduce(const double *__restrict__ partial, int nparts,
                                                      double *__restrict__ out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc += partial[(i64)blockIdx.x * nparts + i];
    const double t = block_sum_d(acc);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

struct BlockScratch {
    DevBuf partial, coef, small;
    BlockScratch() {
        partial.alloc((size_t)DOT_ACC * 2048 * sizeof(double));
        coef.alloc((size_t)DMAX * DMAX * sizeof(double));
        small.alloc((size_t)4 * DMAX * BK * sizeof(double));
    }
};
static BlockScratch &bscratch() {
    static thread_local std::unique_ptr<BlockScratch> s;
    if (!s) s.reset(new BlockScratch());
    return *s;
}
static int nblk_block(i64 n) { return (int)std::max<i64>(1, std::min<i64>((n + 255) / 256, 2048)); }

template <int KO>
static void launch_combo(i64 N, const double *const *X, const double *const *Y, int m, const double *Cdev, int ldc,
                         const double *wdev, double *const *out, double *partial, hipStream_t st) {
    Ptr64 PX{}, PY{};
    for (int j = 0; j < m; ++j) {
        PX.p[j] = X[j];
        PY.p[j] = Y ? Y[j] : nullptr;
    }
    Out8 O{};
    for (int r = 0; r < KO; ++r) O.p[r] = out[r];
    if (Y)
        hipLaunchKernelGGL((k_block_combo<KO, true>), dim3(nblk_block(N)), dim3(256), 0, st, PX, PY, m, Cdev, ldc,
                           wdev, O, partial, N);
    else
        hipLaunchKernelGGL((k_block_combo<KO, false>), dim3(nblk_block(N)), dim3(256), 0, st, PX, PY, m, Cdev, ldc,
                           wdev, O, partial, N);
    HIPOTF_CHECK(hipGetLastError());
}

// out_r = sum_j C[j][r] X_j (Y == nullptr) or residuals sum_j C[j][r] (Y_j - w_r X_j) with their
// norms in norms[r] (host); C is host row-major [m][k]; k may exceed BK (done in slices)
void block_combine(i64 N, int m, int k, const double *C, const double *w, const double *const *X,
                   const double *const *Y, double *const *out, double *norms, hipStream_t st) {
    HIPOTF_REQUIRE(m >= 1 && m <= DMAX && k >= 1 && k <= DMAX, "block_combine: bad sizes");
    auto &S = bscratch();
    std::vector<double> cw((size_t)m * BK + BK);
    const int nb = nblk_block(N);
    for (int r0 = 0; r0 < k; r0 += BK) {
        const int kk = std::min(BK, k - r0);
        for (int j = 0; j < m; ++j)
            for (int r = 0; r < BK; ++r) cw[(size_t)j * BK + r] = r < kk ? C[(size_t)j * k + r0 + r] : 0.0;
        for (int r = 0; r < BK; ++r) cw[(size_t)m * BK + r] = (w && r < kk) ? w[r0 + r] : 0.0;
        HIPOTF_CHECK(hipMemcpyAsync(S.coef.p, cw.data(), cw.size() * sizeof(double), hipMemcpyHostToDevice, st));
        const double *Cd = S.coef.as<double>();
        const double *wd = Cd + (size_t)m * BK;
        double *part = norms ? S.partial.as<double>() : nullptr;
        dispatch_m8(kk, [&](auto kc) {
            constexpr int KO = decltype(kc)::value;
            launch_combo<KO>(N, X, Y, m, Cd, BK, wd, out + r0, part, st);
        });
        if (norms) {
            hipLaunchKernelGGL(k_multi_reduce, dim3(kk), dim3(256), 0, st, S.partial.as<double>(), nb,
                               S.small.as<double>());
            HIPOTF_CHECK(hipGetLastError());
            std::vector<double> h(kk);
            HIPOTF_CHECK(hipMemcpyAsync(h.data(), S.small.p, kk * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPOTF_CHECK(hipStreamSynchronize(st));
            for (int r = 0; r < kk; ++r) norms[r0 + r] = std::sqrt(h[r]);
        }
        // the next slice rewrites the coefficient buffer: let this one finish reading it
        HIPOTF_CHECK(hipStreamSynchronize(st));
    }
}

// In place X_r <- sum_j C[j][r] X_j for r < keep <= m <= DMAX: the subspace restart without a
// second copy of the kept vectors.  Each thread loads all m elements of its index before it
// writes any, and no two threads share an index, so reading and writing the same vectors is
// safe.  C sits in LDS (uniform reads, broadcast).  M = m rounded up to a multiple of 8.
struct PtrMut64 { double *p[DMAX]; };

template <int M>
__global__ __launch_bounds__(256) void k_rotate_inplace(PtrMut64 X, int m, int keep, const double *__restrict__ C,
                                                        i64 n) {
    __shared__ double cs[M * DMAX];
    for (int t = threadIdx.x; t < m * keep; t += blockDim.x) cs[(t / keep) * DMAX + t % keep] = C[t];
    __syncthreads();
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        double x[M];
#pragma unroll
        for (int j = 0; j < M; ++j) x[j] = j < m ? X.p[j][i] : 0.0;
        for (int r = 0; r < keep; ++r) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < M; ++j)
                if (j < m) acc += cs[j * DMAX + r] * x[j];
            X.p[r][i] = acc;
        }
    }
}

// X_r <- sum_j C[j*keep + r] X_j in place (C host row-major [m][keep])
void rotate_inplace(i64 N, int m, int keep, const double *C, double *const *X, hipStream_t st) {
    HIPOTF_REQUIRE(m >= 1 && m <= DMAX && keep >= 1 && keep <= m, "rotate_inplace: bad sizes");
    auto &S = bscratch();
    HIPOTF_CHECK(hipMemcpyAsync(S.coef.p, C, (size_t)m * keep * sizeof(double), hipMemcpyHostToDevice, st));
    PtrMut64 P{};
    for (int j = 0; j < m; ++j) P.p[j] = X[j];
    const int mm = (m + 7) / 8 * 8;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(nblk_block(N)), dim3(256), 0, st, P, m, keep, S.coef.as<double>(), N);
    };
    switch (mm) {
    case 8: go(k_rotate_inplace<8>); break;
    case 16: go(k_rotate_inplace<16>); break;
    case 24: go(k_rotate_inplace<24>); break;
    case 32: go(k_rotate_inplace<32>); break;
    case 40: go(k_rotate_inplace<40>); break;
    case 48: go(k_rotate_inplace<48>); break;
    case 56: go(k_rotate_inplace<56>); break;
    default: go(k_rotate_inplace<64>); break;
    }
    HIPOTF_CHECK(hipGetLastError());
    // the coefficient buffer is rewritten by the next block call: let this one finish reading it
    HIPOTF_CHECK(hipStreamSynchronize(st));
}

__global__ __launch_bounds__(256) void k_scatter(double *__restrict__ v, const i64 *__restrict__ idx,
                                                 const double *__restrict__ val, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) v[idx[i]] = val[i];
}

// v[idx[i]] = val[i], i < count (host arrays)
void scatter_values(double *v, const i64 *idx, const double *val, int count, hipStream_t st) {
    if (count <= 0) return;
    DevBuf di((size_t)count * sizeof(i64)), dv((size_t)count * sizeof(double));
    HIPOTF_CHECK(hipMemcpyAsync(di.p, idx, (size_t)count * sizeof(i64), hipMemcpyHostToDevice, st));
    HIPOTF_CHECK(hipMemcpyAsync(dv.p, val, (size_t)count * sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_scatter, dim3((count + 255) / 256), dim3(256), 0, st, v, di.as<i64>(), dv.as<double>(),
                       count);
    HIPOTF_CHECK(hipGetLastError());
    HIPOTF_CHECK(hipStreamSynchronize(st));
}

// G[a][b] = A_a . B_b into host row-major [na][nb] (na <= DMAX, nb <= BK)
void block_dots(i64 N, int na, const double *const *A, int nbv, const double *const *B, double *G,
                hipStream_t st) {
    HIPOTF_REQUIRE(na >= 0 && na <= DMAX && nbv >= 1 && nbv <= BK, "block_dots: bad sizes");
    if (na == 0) return;
    auto &S = bscratch();
    const int nb = nblk_block(N);
    Ptr64 PB{};
    for (int b = 0; b < nbv; ++b) PB.p[b] = B[b];
    // A vectors per launch: a power of two with NA * nbv <= 32 (at most 16); the groups are
    // balanced, and a group smaller than its NA repeats its first vector (results dropped)
    int cap = 16;
    while (cap * nbv > DOT_ACC) cap /= 2;
    const int ngroups = (na + cap - 1) / cap, per = (na + ngroups - 1) / ngroups;
    int NA = 1;
    while (NA < per) NA *= 2;
    HIPOTF_REQUIRE(NA * nbv <= DOT_ACC, "block_dots: group too large");
    std::vector<double> h((size_t)DOT_ACC);
    for (int a0 = 0; a0 < na; a0 += per) {
        const int ka = std::min(per, na - a0);
        Ptr64 PA{};
        for (int a = 0; a < NA; ++a) PA.p[a] = A[a0 + (a < ka ? a : 0)];
        double *part = S.partial.as<double>();
        dispatch_m8(nbv, [&](auto bc) {
            constexpr int NB = decltype(bc)::value;
            switch (NA) {
            case 1: hipLaunchKernelGGL((k_block_dots<1, NB>), dim3(nb), dim3(256), 0, st, PA, PB, part, N); break;
            case 2: hipLaunchKernelGGL((k_block_dots<2, NB>), dim3(nb), dim3(256), 0, st, PA, PB, part, N); break;
            case 4: hipLaunchKernelGGL((k_block_dots<4, (NB <= 8 ? NB : 8)>), dim3(nb), dim3(256), 0, st, PA, PB,
                                       part, N); break;
            case 8:
                if constexpr (NB <= 4)
                    hipLaunchKernelGGL((k_block_dots<8, NB>), dim3(nb), dim3(256), 0, st, PA, PB, part, N);
                break;
            default:
                if constexpr (NB <= 2)
                    hipLaunchKernelGGL((k_block_dots<16, NB>), dim3(nb), dim3(256), 0, st, PA, PB, part, N);
                break;
            }
        });
        HIPOTF_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_multi_reduce, dim3(NA * nbv), dim3(256), 0, st, part, nb, S.small.as<double>());
        HIPOTF_CHECK(hipGetLastError());
        HIPOTF_CHECK(hipMemcpyAsync(h.data(), S.small.p, (size_t)NA * nbv * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPOTF_CHECK(hipStreamSynchronize(st));
        for (int a = 0; a < ka; ++a)
            for (int b = 0; b < nbv; ++b) G[(size_t)(a0 + a) * nbv + b] = h[(size_t)a * nbv + b];
    }
}

// T_j -= sum_i P[i][j] V_i  (P host row-major [m][nt], nt <= BK)
void block_subtract(i64 N, int m, const double *const *V, const double *P, int nt, double *const *T,
                    hipStream_t st) {
    HIPOTF_REQUIRE(m >= 0 && m <= DMAX && nt >= 1 && nt <= BK, "block_subtract: bad sizes");
    if (m == 0) return;
    auto &S = bscratch();
    std::vector<double> pw((size_t)m * BK, 0.0);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < nt; ++j) pw[(size_t)i * BK + j] = P[(size_t)i * nt + j];
    HIPOTF_CHECK(hipMemcpyAsync(S.coef.p, pw.data(), pw.size() * sizeof(double), hipMemcpyHostToDevice, st));
    Ptr64 PV{};
    for (int i = 0; i < m; ++i) PV.p[i] = V[i];
    Out8 O{};
    for (int j = 0; j < nt; ++j) O.p[j] = T[j];
    dispatch_m8(nt, [&](auto kc) {
        constexpr int KO = decltype(kc)::value;
        hipLaunchKernelGGL(k_block_subtract<KO>, dim3(nblk_block(N)), dim3(256), 0, st, PV, m,
                           S.coef.as<double>(), BK, O, N);
    });
    HIPOTF_CHECK(hipGetLastError());
    HIPOTF_CHECK(hipStreamSynchronize(st));
}

}  // namespace hipotf
