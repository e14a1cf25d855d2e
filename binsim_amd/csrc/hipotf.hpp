// Internal declarations of libhipotf (not part of the C ABI; see include/hipotf.h).
//
// Data layout in HBM (one symmetry block, |a,b> at index ia*Nb + ib as in
// src/include/basis/basis.hpp:327):
//   * state vectors: float64[dim], alpha-string-major (a row = one alpha string, contiguous
//     beta strings), interleaved Jordan-Wigner sign convention of the reference
//     (spin-orbital 2p = alpha_p, 2p+1 = beta_p; src/include/ham/ham.hpp:419-451).
//   * sigma work buffers: T[dim*nv] with the nv vectors of a block interleaved per
//     determinant, in the alpha-major gauge  x~ = E x,  E_ab = (-1)^popc(b & M(a)),
//     M(a) = XOR_{p in a} ((1<<p)-1); both alpha-major and beta-major copies.
//   * string tables per spin: colex-ordered strings, single-excitation table SX[slot][I]
//     (packed J|r|s|sign), and the same-spin Hamiltonian G as fixed-length CSR rows.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace hipotf {

using i64 = int64_t;
using u32 = uint32_t;
using u64 = uint64_t;

// ------------------------------------------------------------------------------------------
// errors: every C entry point catches, records a thread-local message and returns a sentinel
// ------------------------------------------------------------------------------------------
void set_last_error(const std::string &msg);
void clear_last_error();

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPOTF_CHECK(expr)                                                                   \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess)                                                                \
            throw ::hipotf::Error(std::string(#expr) + " failed: " + hipGetErrorString(_e)); \
    } while (0)

#define HIPOTF_REQUIRE(cond, msg)                   \
    do {                                            \
        if (!(cond)) throw ::hipotf::Error(msg);    \
    } while (0)

// Set while an allocation-free entry point runs: any device allocation attempted underneath
// (a table that would be built on first use) throws instead of calling hipMalloc.
extern thread_local bool g_forbid_alloc;
struct ForbidAlloc {
    bool prev;
    ForbidAlloc() : prev(g_forbid_alloc) { g_forbid_alloc = true; }
    ~ForbidAlloc() { g_forbid_alloc = prev; }
};

// RAII device allocation
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        if (g_forbid_alloc)
            throw Error("a device table this call needs is not built yet (an allocation-free entry point "
                        "never allocates): call hipotf_prepare or the allocating entry point once first");
        release();
        if (n == 0) return;
        HIPOTF_CHECK(hipMalloc(&p, n));
        bytes = n;
    }
    void ensure(size_t n) { if (n > bytes) alloc(n); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
    template <typename T> void upload(const std::vector<T> &v) {
        alloc(v.size() * sizeof(T));
        if (!v.empty()) HIPOTF_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
};

// ------------------------------------------------------------------------------------------
// host-side basis (mirror of BasisManager, src/include/basis/basis.hpp:29-110)
// ------------------------------------------------------------------------------------------
struct SpinSpace {
    int n = 0, k = 0;
    i64 N = 0;                       // C(n,k)
    std::vector<u64> strs;           // ascending numeric (= colex) order
    int nsx = 0;                     // k*(n-k) single excitations per string
    int ndx = 0;                     // C(k,2)*C(n-k,2) double excitations per string
};

struct HostBasis {
    i64 norb = 0, na = 0, nb = 0, total_sym = 0, num_irreps = 0;
    std::vector<i64> orbsym;
    i64 dim = 0;
    i64 num_blocks = 0;
    SpinSpace A, B;
};

HostBasis *make_host_basis(i64 norb, i64 na, i64 nb, i64 total_sym, const i64 *orbsym, i64 num_irreps);

// single-excitation table entry: J (20 bits) | r (5) | s (5) | unused | sign (bit 31)
// meaning: dest string I has orbital r occupied, s empty; J = I - r + s; <I|E_rs|J> = sign.
__host__ __device__ inline u32 sx_pack(u32 J, int r, int s, bool neg) {
    return J | (u32(r) << 20) | (u32(s) << 25) | (neg ? 0x80000000u : 0u);
}
__host__ __device__ inline u32 sx_J(u32 e) { return e & 0xFFFFFu; }
__host__ __device__ inline int sx_r(u32 e) { return int((e >> 20) & 31u); }
__host__ __device__ inline int sx_s(u32 e) { return int((e >> 25) & 31u); }
__host__ __device__ inline bool sx_neg(u32 e) { return (e >> 31) != 0u; }

// ------------------------------------------------------------------------------------------
// device basis (counterpart of BasisViewDev, src/include/cuda/basis.cuh:8-186)
// ------------------------------------------------------------------------------------------
struct DevSpin {
    int n = 0, k = 0, nsx = 0, ndx = 0;
    i64 N = 0;
    DevBuf strs;      // u64[N]
    DevBuf sx;        // u32[nsx][N]
};

struct DevBasis {
    HostBasis host;           // own copy of the host description
    DevSpin A, B;             // B aliases A's tables when na == nb (shared pointers below)
    DevSpin *pA = nullptr, *pB = nullptr;
    DevBuf gaugeM;            // u64[Na]: M(a) = XOR_{p in a} ((1<<p)-1)
    i64 uid = 0;              // unique id (operators cache per-basis tables under it)
    // scratch shared by sigma calls on this basis
    DevBuf xt, xtT, s1, s2T;
    DevBuf red;               // reduction partials
};

DevBasis *make_dev_basis(const HostBasis &hb);

// ------------------------------------------------------------------------------------------
// Hamiltonian in integral form (replaces the SVD network of src/include/ham/otf.hpp:114-169
// for the 2-body H; the network constructor decodes a network into this form)
// ------------------------------------------------------------------------------------------
constexpr int SS_RANGES = 8;  // source-index ranges the same-spin rows are split into

// the same-spin rows grouped in blocks of rb consecutive destination strings: per block the
// ascending union of its rows' sources, each with rb coefficients (0 where a row of the block
// has no such source), so that one gathered source row serves rb destination rows
struct RowBlockTable {
    int rb = 0;
    i64 nblk = 0;             // ceil(N / rb)
    i64 entries = 0;
    DevBuf J;                 // int32[entries]: sources of block 0, then block 1, ...
    DevBuf P;                 // int32[nblk][SS_RANGES + 1]: entry offset (absolute) of the
                              // first source >= g * N / SS_RANGES; P[.][SS_RANGES] = block end
    DevBuf c64, c32;          // double / float [entries][rb]
};

struct SameSpinTable {
    i64 N = 0;
    int L = 0;                // 1 + nsx + ndx entries per dest string, ascending source index
    DevBuf J;                 // int32[N][L]
    DevBuf P;                 // int32[N][SS_RANGES + 1]: first entry of each source range
                              // g (sources >= g * N / SS_RANGES); P[.][SS_RANGES] = L
    DevBuf c64, c32;          // double / float [N][L]
    DevBuf gdiag;             // double[N]  same-spin diagonal G[I,I]
    std::vector<int32_t> hJ;  // host copies of J and the FP64 coefficients when the table was
    std::vector<double> hC;   // built on the host (row blocks are then built from them)
    bool dev_built = false;   // built on the device (tables.hip): row blocks too
    std::unique_ptr<RowBlockTable> blocks[3];  // rb = 4, 8, 16
};

// row-block table of T for rb in {4, 8, 16}, built (and uploaded) on first use
const RowBlockTable &row_blocks(SameSpinTable &T, int rb);
// device construction (tables.hip): false when a row is too long for its in-LDS sort (then the
// host build applies)
struct DevSpin;
bool build_same_spin_dev(SameSpinTable &T, const DevSpin &S, const DevBuf &h64, const DevBuf &eri64, int n);
void row_blocks_dev(SameSpinTable &T, RowBlockTable &R, int rb);
void build_jab_dev(DevBuf &jab, const DevSpin &A, int n, const DevBuf &eri64);
void build_sx_dev(DevSpin &D);  // single-excitation table of one spin (n <= 32)

constexpr int ERI_PAD = 256;  // zero entries appended to the device integral tables

struct DevHam {
    DevBasis *basis = nullptr;
    int n = 0;
    double ecore = 0.0;
    std::vector<double> h, eri;   // host copies: h[p*n+q], eri[((p*n+q)*n+r)*n+s] = (pq|rs)
    DevBuf h64;                   // h on device
    DevBuf eri64, eri32;          // (pq|rs) on device
    DevBuf jab;                   // double[Na][n]: sum_{p in a} (pp|rr)
    DevBuf dw64, dw32;            // [Na][256]: sum_{p in a} (pp|rt) at r*n + t (n*n <= 256),
                                  // zero beyond n*n; built on first use by the alpha-beta kernel
    SameSpinTable tA, tB;
    SameSpinTable *ptA = nullptr, *ptB = nullptr;
};

// determinants as interleaved spin-orbital occupations (bit 2p = alpha p, 2p+1 = beta p)
u64 interleave_occ(u32 a, u32 b);
// one ladder operator on occupation x; adds the Jordan-Wigner parity to par; false if zero
bool ladder_apply(u64 &x, int &par, int pos, bool create);
// <bra|H|ket> of the two-body Hamiltonian (h, (pq|rs), ecore) by Slater-Condon rules with
// interleaved Jordan-Wigner signs (the reference's convention)
double slater_condon(int n, const double *h, const double *eri, double ecore, u64 bra, u64 ket);
void integrals_from_reference_layout(int n, const double *h_colmajor, const double *g_ref_colmajor,
                                     std::vector<double> &h, std::vector<double> &eri);
DevHam *make_dev_ham(DevBasis *b, int n, const std::vector<double> &h, const std::vector<double> &eri,
                     double ecore);

// launch choices of the sigma kernels (process-wide; set by hipotf_set_tuning, chosen by the
// autotuner in binsim_amd/autotune.py)
struct SigmaTuning {
    int ab_kernel = 0;      // alpha-beta kernel: 0 first that applies (source-lane, register,
                            // chunked), 1 source-lane, 2 register-descriptor, 3 chunked
    int ab_split = 1;       // source-lane: threads per source string (1: 8 waves/CU, 2: 16)
    int ab_db = 1;          // source-lane: integral-row buffers (1, or 2 where LDS allows)
    int ab_nvs = 2;         // source-lane, FP32: vectors per launch (1 or 2)
    int rs_passes = 0;      // row SpMM: launches over source ranges (1, 2, 4, 8; 0 = fewest
                            // for which a column tile's gathered slice fits an XCD's L2)
    int deterministic = 0;  // 1: alpha-beta scatter wave by wave (bitwise reproducible)
    int rs_block = 8;       // row SpMM: destination rows sharing each gathered source row
                            // (4, 8, 16), or 0 for one row per wave
    int sweep_mode = 4;     // VQE rotation sweeps (FP64 table): 0 fused segments with as many
                            // rows (and the leading part of one more) in LDS as fit, the
                            // others in global memory; 1 fused, rows
                            // always in global memory; 2 one launch per generator; 3 fused,
                            // at most one row and half of the next per block in LDS; +4 (4..7):
                            // the same with the adjoint sweep's pair segments on interleaved
                            // (l, r) rows (the default: half the scattered accesses where rows
                            // miss the LDS)
};
SigmaTuning &tuning(int precision);  // 0: FP64 kernels, 1: FP32 kernels
// timing of the sigma builds' alpha-beta sections (hipotf_ab_timing): while on, HIP events on
// the sigma stream around each section, synchronized and accumulated
struct AbTiming {
    bool on = false;
    double ms = 0.0;
    long long sections = 0, mma_launches = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
};
AbTiming &ab_timing();
constexpr int TUNING_FIELDS = 8;

// sigma: dst_v = H src_v for nv vectors; prec 0 = float64 work, 1 = float32 work; at most
// maxnv (1..8) vectors share one pass (its work buffers hold dim * maxnv words each)
void sigma(DevHam *H, int nv, const double *const *src, double *const *dst, int prec,
           hipStream_t st, int maxnv = 8);
void diag_elements(DevHam *H, double *diags, hipStream_t st);  // accumulates
// one FP64 sigma with the four work buffers carved from the library's static device scratch
// array (no device allocation; every table must already exist); at most
// NOALLOC_MAX_DIM determinants
constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;
constexpr i64 NOALLOC_MAX_DIM = (i64)(NOALLOC_SCRATCH_BYTES / 32);
void sigma_noalloc(DevHam *H, const double *src, double *dst, hipStream_t st);

// ------------------------------------------------------------------------------------------
// generator pool (UCC excitations G = T - T^dagger, jl/ansatz.jl:230-276)
// ------------------------------------------------------------------------------------------
struct PairList {            // pattern within one spin: strings u with ann subset, cre disjoint
    u64 cre = 0, ann = 0;
    i64 len = 0;
    DevBuf uw;               // int32[2*len]: (index of u, index of w = u^cre^ann)
    bool identity = false;   // (x, x) for every string
    int spin = 0;            // 0 alpha, 1 beta
};

struct Generator {
    int nops = 0;
    int pos[8];              // interleaved spin-orbital positions, in APPLICATION order
    int cre[8];              // 1 = creation
    int kind = 0;            // 0 identity, 1 alpha only, 2 beta only, 3 mixed
    int pa = -1, pb = -1;    // pattern ids (into DevPool::lists)
};

struct DevPool {
    DevBasis *basis = nullptr;
    std::vector<Generator> gens;
    std::vector<PairList> lists;   // lists[0] = alpha identity, lists[1] = beta identity
    DevBuf partial;                // double partial sums for reductions
    DevBuf scal;                   // a few device doubles
    DevBuf segdesc, segpart;       // fused sweeps: per-position descriptors, per-block partials
    DevBuf tl, tr;                 // fused sweeps: beta-major copies of the two state vectors
    DevBuf tz;                     // fused adjoint sweep on interleaved rows: (l, r) per element
    DevBuf vqe_res;                // fused VQE objective: per-parameter results + norms (grown, kept)
    DevBuf zeros;                  // batch tran: a zero angle per generator
};

// forward expm sweep and/or adjoint backgrad sweep over positions idxs[0..n) with fused
// row-local segments; dres[i] = backgrad value of position i (device)
void fused_sweeps(DevPool *P, i64 n, const i64 *idxs, const double *thetas, double *psi, double *lam,
                  bool forward, bool backward, double *dres, hipStream_t st);

// gradient-only pass over every pool generator (jl/vqe.jl's batch gradient for ADAPT):
// grads[g] = <lp| G_g e^{theta_g G_g} |rp>, the generators grouped into the sweeps' fused
// segments; thetas and grads are device arrays
void fused_grads(DevPool *P, const double *thetas, const double *lp, const double *rp, double *grads,
                 hipStream_t st);
// trans[g] = the gradient of generator g at theta = 0 (op/batchtran.hpp:517-528), device arrays
void batch_tran(DevPool *P, const double *lp, const double *rp, double *trans, hipStream_t st);

DevPool *make_dev_pool(DevBasis *b, i64 nops, const int32_t *offsets, const int32_t *orbs);

// generic Pauli-group network (network.hip): any operator given as the reference's SVD network,
// evaluated group by group; group g: flips (ax, bx), rank, z masks and weights w[r * nz + k]
struct DevGenNet;
DevGenNet *make_dev_gennet(int n, const std::vector<u32> &ax, const std::vector<u32> &bx,
                           const std::vector<int> &rank, const std::vector<std::vector<u32>> &zas,
                           const std::vector<std::vector<u32>> &zbs, const std::vector<std::vector<double>> &was,
                           const std::vector<std::vector<double>> &wbs);
void destroy_dev_gennet(DevGenNet *D);
void gennet_hvec(DevGenNet *D, DevBasis *B, const double *x, double *y, hipStream_t st);  // y = H x
void gennet_diag(DevGenNet *D, DevBasis *B, double *diag, hipStream_t st);               // accumulates

void rot_expm(DevPool *P, i64 idx, double theta, double *vec, hipStream_t st);
// dst = G_idx src (dst zeroed first)
void rot_tvec(DevPool *P, i64 idx, const double *src, double *dst, hipStream_t st);
// lp, rp <- exp(theta G_idx) lp, rp; bp = G_idx lp (bp zeroed first)
void rot_backtran(DevPool *P, i64 idx, double theta, double *lp, double *rp, double *bp, hipStream_t st);
// exp(theta G_idx) on columns 0..nvec-1 of the det-major matrix m[i * ld + k]
void rot_batch_expm(DevPool *P, i64 idx, double theta, double *m, i64 ld, i64 nvec, hipStream_t st);
// grad: returns device pointer to the scalar result (written asynchronously)
void rot_grad(DevPool *P, i64 idx, double theta, const double *lp, const double *rp, double *dres,
              hipStream_t st);
void rot_backgrad(DevPool *P, i64 idx, double theta, double *lp, double *rp, double *dres,
                  hipStream_t st);

// device reductions used by Davidson / VQE
void dev_dot(const double *a, const double *b, i64 n, double *dres, DevBuf &scratch, hipStream_t st);

int num_cus();

}  // namespace hipotf
