// Davidson inner-loop kernels (libdiag on the GPU).
//
// Same entry points and arithmetic as src/include/diag/davidson.hpp:31-272 (C ABI
// src/include/diag/declare.hpp:4-75); state vectors are device pointers, the small arrays
// (Ritz coefficients, heff) stay on the host as in jl/davidson.jl.  Reductions are two-stage
// (per-block partials, then one block) so results are deterministic run to run.
#include "hipotf.hpp"

#include <algorithm>
#include <cmath>

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

// y = sum_j c_j V_j ; if AV given: r = sum_j c_j AV_j - e y (written to y), partial |r|^2
__global__ __launch_bounds__(256) void k_combine(VecPack V, VecPack AV, int dim, int with_res,
                                                 double e, double *__restrict__ out,
                                                 double *__restrict__ partial, i64 n) {
    double acc = 0.0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        double x = 0.0, ax = 0.0;
        for (int j = 0; j < dim; ++j) {
            x += V.c[j] * V.v[j][i];
            if (with_res) ax += V.c[j] * AV.v[j][i];
        }
        const double r = with_res ? ax - e * x : x;
        out[i] = r;
        acc += r * r;
    }
    const double t = block_sum_d(acc);
    if (threadIdx.x == 0 && partial) partial[blockIdx.x] = t;
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
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        double d = diags[i] - e;
        if (fabs(d) < as) d = d >= 0 ? as : -as;
        axt[i] = axt[i] / d;
    }
}

__global__ __launch_bounds__(256) void k_dotp(const double *__restrict__ a, const double *__restrict__ b,
                                              double *__restrict__ partial, i64 n) {
    double acc = 0.0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        acc += a[i] * b[i];
    const double t = block_sum_d(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// y -= (*proj) x
__global__ __launch_bounds__(256) void k_axpy_dev(double *__restrict__ y, const double *__restrict__ x,
                                                  const double *__restrict__ proj, i64 n) {
    const double p = *proj;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        y[i] -= p * x[i];
}

// y *= 1 / *nrm
__global__ __launch_bounds__(256) void k_scale_inv(double *__restrict__ y, const double *__restrict__ nrm,
                                                   i64 n) {
    const double s = 1.0 / *nrm;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        y[i] *= s;
}

// out[j] = <V_j, a> for j < m (m <= 8): each block accumulates 8 partials
__global__ __launch_bounds__(256) void k_multidot(VecPack V, int m, const double *__restrict__ a,
                                                  double *__restrict__ partial, int ldp, i64 n) {
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const double av = a[i];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (j < m) acc[j] += V.v[j][i] * av;
    }
    for (int j = 0; j < m; ++j) {
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
    hipLaunchKernelGGL(k_combine, dim3(nb), dim3(256), 0, st, P, Q, dim, 1, min_e, axt,
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
    hipLaunchKernelGGL(k_combine, dim3(nblk(N)), dim3(256), 0, st, P, Q, dim, 0, 0.0, xt,
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
        hipLaunchKernelGGL(k_multidot, dim3(nb), dim3(256), 0, st, P, m, a, S.partial.as<double>(), 2048, N);
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

}  // namespace hipotf
