// Determinant space: colex-ordered alpha/beta strings and their single-excitation tables.
//
// Follows create_basis_manager_tmpl (src/include/basis/basis.hpp:118-296): strings are
// enumerated with Gosper's hack (next_combination, src/include/core/types.hpp:78-86) in
// ascending numeric order and the determinant |a,b> sits at index ia*Nb + ib
// (basis.hpp:327).  Only the one-block case (every orbsym entry 0 and total_sym 0) is
// supported; other symmetry settings are rejected with an error.
#include "hipotf.hpp"

#include <algorithm>
#include <thread>

namespace hipotf {

static thread_local std::string g_last_error;
thread_local bool g_forbid_alloc = false;
void set_last_error(const std::string &msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }
const char *last_error_cstr() { return g_last_error.c_str(); }

static inline int popc64(u64 v) { return __builtin_popcountll(v); }

static inline u64 next_comb(u64 v) {
    u64 c = v & (~v + 1);
    u64 r = v + c;
    return (((r ^ v) >> 2) / c) | r;
}

static void build_spin(SpinSpace &S, int n, int k) {
    S.n = n;
    S.k = k;
    S.strs.clear();
    if (k < 0 || k > n) {
        S.N = 0;
        return;
    }
    u64 s = (k == 0) ? 0 : ((k == 64) ? ~0ull : ((1ull << k) - 1));
    const u64 lim = (n == 64) ? ~0ull : (1ull << n);
    while (true) {
        S.strs.push_back(s);
        if (k == 0) break;
        u64 t = next_comb(s);
        if (t >= lim || t <= s) break;
        s = t;
    }
    S.N = (i64)S.strs.size();
    S.nsx = k * (n - k);
    S.ndx = (k * (k - 1) / 2) * ((n - k) * (n - k - 1) / 2);
}

HostBasis *make_host_basis(i64 norb, i64 na, i64 nb, i64 total_sym, const i64 *orbsym,
                           i64 num_irreps) {
    HIPOTF_REQUIRE(norb >= 1 && norb <= 32, "norb must be in [1, 32] per spin");
    HIPOTF_REQUIRE(na >= 0 && na <= norb && nb >= 0 && nb <= norb, "electron count out of range");
    HIPOTF_REQUIRE(num_irreps >= 1, "num_irreps must be >= 1");
    auto hb = std::make_unique<HostBasis>();
    hb->norb = norb;
    hb->na = na;
    hb->nb = nb;
    hb->total_sym = total_sym;
    hb->num_irreps = num_irreps;
    hb->orbsym.assign(orbsym ? orbsym : nullptr, orbsym ? orbsym + norb : nullptr);
    if (hb->orbsym.empty()) hb->orbsym.assign(norb, 0);
    bool trivial = std::all_of(hb->orbsym.begin(), hb->orbsym.end(), [](i64 v) { return v == 0; });
    HIPOTF_REQUIRE(trivial, "only the single-block basis (all orbsym == 0) is supported");
    build_spin(hb->A, (int)norb, (int)na);
    build_spin(hb->B, (int)norb, (int)nb);
    // one block iff total_sym ^ 0 == 0 (basis.hpp:173-180); otherwise the space is empty
    if (total_sym == 0 && hb->A.N > 0 && hb->B.N > 0) {
        hb->num_blocks = 1;
        hb->dim = hb->A.N * hb->B.N;
    } else {
        hb->num_blocks = 0;
        hb->dim = 0;
    }
    return hb.release();
}

static i64 find_str(const SpinSpace &S, u64 s) {
    auto it = std::lower_bound(S.strs.begin(), S.strs.end(), s);
    if (it != S.strs.end() && *it == s) return it - S.strs.begin();
    return -1;
}

static inline u64 between_mask(int r, int s) {
    int lo = r < s ? r : s, hi = r < s ? s : r;
    return ((1ull << hi) - 1) & ~((1ull << (lo + 1)) - 1);
}

template <typename F>
static void parallel_for(i64 n, F f) {
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (n < 4096 || nt == 1) {
        for (i64 i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([=, &f]() {
            for (i64 i = t; i < n; i += nt) f(i);
        });
    for (auto &x : th) x.join();
}

// SX[slot][I], slot = io*(n-k) + iv; r = io-th occupied, s = iv-th empty orbital of I.
static std::vector<u32> build_sx(const SpinSpace &S) {
    const int n = S.n, k = S.k, nv = n - k;
    HIPOTF_REQUIRE(S.N < (1 << 20), "too many strings per spin for the packed tables (2^20)");
    std::vector<u32> sx((size_t)S.nsx * S.N);
    parallel_for(S.N, [&](i64 I) {
        u64 s = S.strs[I];
        int occ[64], vir[64], no = 0, nvv = 0;
        for (int p = 0; p < n; ++p) {
            if ((s >> p) & 1) occ[no++] = p;
            else vir[nvv++] = p;
        }
        for (int io = 0; io < k; ++io)
            for (int iv = 0; iv < nv; ++iv) {
                int r = occ[io], t = vir[iv];
                u64 j = s ^ (1ull << r) ^ (1ull << t);
                i64 J = find_str(S, j);
                bool neg = popc64(s & between_mask(r, t)) & 1;
                sx[(size_t)(io * nv + iv) * S.N + I] = sx_pack((u32)J, r, t, neg);
            }
    });
    return sx;
}

static void upload_spin(DevSpin &D, const SpinSpace &S) {
    D.n = S.n;
    D.k = S.k;
    D.N = S.N;
    D.nsx = S.nsx;
    D.ndx = S.ndx;
    D.strs.upload(S.strs);
    // single-excitation table built on the device (tables.hip), on the host past 32 orbitals
    if (S.nsx > 0 && S.N > 0) {
        if (S.n <= 32) build_sx_dev(D);
        else D.sx.upload(build_sx(S));
    }
}

DevBasis *make_dev_basis(const HostBasis &hb) {
    auto db = std::make_unique<DevBasis>();
    db->host = hb;
    static i64 next_uid = 1;
    db->uid = next_uid++;
    upload_spin(db->A, hb.A);
    db->pA = &db->A;
    if (hb.na == hb.nb) {
        db->pB = &db->A;
    } else {
        upload_spin(db->B, hb.B);
        db->pB = &db->B;
    }
    std::vector<u64> M(hb.A.N);
    for (i64 i = 0; i < hb.A.N; ++i) {
        u64 a = hb.A.strs[i], m = 0;
        for (int p = 0; p < hb.norb; ++p)
            if ((a >> p) & 1) m ^= (1ull << p) - 1;
        M[i] = m;
    }
    db->gaugeM.upload(M);
    return db.release();
}

int num_cus() {
    static int cached = 0;
    if (!cached) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cached = prop.multiProcessorCount;
        else
            cached = 256;
    }
    return cached;
}

// exposed to the table builder of ham.cpp
i64 host_find_str(const SpinSpace &S, u64 s) { return find_str(S, s); }

}  // namespace hipotf
