// Host side of the C ABI declared in include/admp_hip.h.
#include "../../include/admp_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "admp_common.h"
#include "nblist_kernels.h"
#include "pair_kernels.h"
#include "recip_kernels.h"
#include "site_kernels.h"

using namespace admp;

namespace {

struct Grid {
    int K[3] = {0, 0, 0};
    size_t G = 0;
    FFTPlan plan[3];
    void* tw[3] = {nullptr, nullptr, nullptr};   // R precision
    void* tw64[3] = {nullptr, nullptr, nullptr}; // double twiddles (used by the FP64 grids)
    double* bmod[3] = {nullptr, nullptr, nullptr};
};

}  // namespace

struct admp_ctx {
    admp_config cfg;
    hipStream_t stream = nullptr;
    std::string err;
    int N = 0, F = 1, nh = 1;
    int cap = 0;
    bool has_topo = false, has_nbr = false;
    int use_frames = 0;
    // topology
    int *d_axis_type = nullptr, *d_axis_idx = nullptr, *d_xoff = nullptr, *d_xent = nullptr;
    int *d_rev_off = nullptr, *d_rev_ent = nullptr;
    // geometry
    BoxInfo* d_box = nullptr;
    double4* d_pos = nullptr;
    double *d_frames = nullptr, *d_Qc = nullptr;
    // neighbours
    int *d_nbr = nullptr, *d_ncnt = nullptr, *d_nhalf = nullptr, *d_maxcnt = nullptr;
    int *d_cell = nullptr, *d_ccount = nullptr, *d_cstart = nullptr, *d_cfill = nullptr, *d_catoms = nullptr;
    int ncell_alloc = 0;
    // per-site work (allocated for double, reused for float)
    void *d_site = nullptr, *d_pcoef = nullptr, *d_vR = nullptr;
    int4* d_mbase = nullptr;
    void *d_wfrac = nullptr, *d_fq = nullptr;
    // grids
    Grid eg, dg;
    void* d_grid = nullptr;     // complex, F * max(G)
    double* d_part = nullptr;   // influence partials
    int part_blocks = 0;
    double* d_rec = nullptr;    // F x 7 reduced influence sums
    // accumulators
    double *d_e_atom = nullptr, *d_gpos = nullptr, *d_gq = nullptr, *d_vir_atom = nullptr, *d_gpol = nullptr,
           *d_gth = nullptr, *d_e_self = nullptr, *d_e_pol = nullptr;
    double *d_gq_rec = nullptr, *d_gpos_rec = nullptr, *d_vir_rec = nullptr;
    double *d_contrib = nullptr, *d_vir_fr = nullptr, *d_gQl = nullptr;
    double *d_red = nullptr;  // F x 16 scratch reductions
    double *d_red9a = nullptr, *d_red9b = nullptr, *d_red9c = nullptr;
    // solver
    double *d_b = nullptr, *d_r = nullptr, *d_z = nullptr, *d_p = nullptr, *d_Ap = nullptr, *d_t1 = nullptr,
           *d_t2 = nullptr, *d_U = nullptr;
    double *d_dpart = nullptr, *d_dots = nullptr, *d_coef = nullptr;
    int* d_active = nullptr;
    int dots_blocks = 64;
    std::vector<void*> allocs;
    // per-kernel timing (HIP events around launches on the context stream)
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    double prof_ms[16] = {0};
    long prof_n[16] = {0};
};

static const char* PROF_NAMES[16] = {"real_final", "real_permfield", "real_matvec", "spline_prep", "spread",
                                     "fft", "influence", "gather", "nbr_build", "frames", "pair_short",
                                     "solver_vec", "", "", "", ""};

static hipEvent_t ev_get(admp_ctx* ctx) {
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}

// RAII bracket: records a start/stop event pair around the launches in scope
struct ProfScope {
    admp_ctx* ctx;
    int key;
    hipEvent_t a = nullptr;
    ProfScope(admp_ctx* c, int k) : ctx(c), key(k) {
        if (ctx->prof) {
            a = ev_get(ctx);
            (void)hipEventRecord(a, ctx->stream);
        }
    }
    ~ProfScope() {
        if (ctx->prof) {
            hipEvent_t b = ev_get(ctx);
            (void)hipEventRecord(b, ctx->stream);
            ctx->ev_pending.push_back({key, {a, b}});
        }
    }
};

// resolve completed event pairs (call after a stream synchronisation)
static void prof_resolve(admp_ctx* ctx) {
    for (auto& p : ctx->ev_pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
            ctx->prof_ms[p.first] += ms;
            ctx->prof_n[p.first] += 1;
        }
        ctx->ev_pool.push_back(p.second.first);
        ctx->ev_pool.push_back(p.second.second);
    }
    ctx->ev_pending.clear();
}

#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            ctx->err = std::string(#x) + ": " + hipGetErrorString(e_);             \
            return ADMP_ERR_HIP;                                                   \
        }                                                                          \
    } while (0)

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <class T> static hipError_t dalloc(admp_ctx* ctx, T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) {
        ctx->allocs.push_back(q);
        *p = reinterpret_cast<T*>(q);
    }
    return e;
}

static void dfree(admp_ctx* ctx, void* p) {
    if (!p) return;
    auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), p);
    if (it != ctx->allocs.end()) ctx->allocs.erase(it);
    (void)hipFree(p);
}

static std::vector<int> factorize(int K) {
    std::vector<int> f;
    int n = K;
    while (n % 4 == 0) { f.push_back(4); n /= 4; }
    while (n % 2 == 0) { f.push_back(2); n /= 2; }
    for (int p = 3; p * p <= n; p += 2)
        while (n % p == 0) { f.push_back(p); n /= p; }
    if (n > 1) f.push_back(n);
    return f;
}

// theta_k per dimension: sum_{m=-2..2} M6(m+3) cos(2 pi m i / K) (recip.py:440-448)
static std::vector<double> bspline_moduli(int K) {
    // M6 at integer arguments 1..5 (cardinal B-spline of order 6)
    const double m6[5] = {1.0 / 120, 26.0 / 120, 66.0 / 120, 26.0 / 120, 1.0 / 120};
    std::vector<double> out(K);
    for (int i = 0; i < K; ++i) {
        double s = 0;
        for (int m = -2; m <= 2; ++m) s += m6[m + 2] * std::cos(2.0 * M_PI * m * i / K);
        out[i] = s;
    }
    return out;
}

static admp_status setup_grid(admp_ctx* ctx, Grid& g, int K0, int K1, int K2, bool mixed) {
    g.K[0] = K0; g.K[1] = K1; g.K[2] = K2;
    g.G = (size_t)K0 * K1 * K2;
    for (int d = 0; d < 3; ++d) {
        const int K = g.K[d];
        std::vector<int> f = factorize(K);
        if ((int)f.size() > 12) { ctx->err = "FFT size has too many factors"; return ADMP_ERR_ARG; }
        g.plan[d].K = K;
        g.plan[d].nst = (int)f.size();
        for (size_t s = 0; s < f.size(); ++s) g.plan[d].fac[s] = f[s];
        std::vector<double> tw(2 * K);
        std::vector<float> twf(2 * K);
        for (int t = 0; t < K; ++t) {
            long double a = -2.0L * 3.14159265358979323846264338327950288L * t / K;
            tw[2 * t] = (double)cosl(a); tw[2 * t + 1] = (double)sinl(a);
            twf[2 * t] = (float)tw[2 * t]; twf[2 * t + 1] = (float)tw[2 * t + 1];
        }
        double* t64;
        HIPCHK(dalloc(ctx, &t64, 2 * K));
        HIPCHK(hipMemcpy(t64, tw.data(), 2 * K * sizeof(double), hipMemcpyHostToDevice));
        g.tw64[d] = t64;
        if (mixed) {
            float* t32;
            HIPCHK(dalloc(ctx, &t32, 2 * K));
            HIPCHK(hipMemcpy(t32, twf.data(), 2 * K * sizeof(float), hipMemcpyHostToDevice));
            g.tw[d] = t32;
        } else {
            g.tw[d] = t64;
        }
        std::vector<double> bm = bspline_moduli(K);
        HIPCHK(dalloc(ctx, &g.bmod[d], K));
        HIPCHK(hipMemcpy(g.bmod[d], bm.data(), K * sizeof(double), hipMemcpyHostToDevice));
    }
    return ADMP_OK;
}

extern "C" const char* admp_version(void) { return "admp-hip 0.1 (gfx950)"; }

extern "C" const char* admp_last_error(const admp_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" admp_status admp_set_stream(admp_ctx* ctx, void* stream) {
    if (!ctx) return ADMP_ERR_ARG;
    ctx->stream = reinterpret_cast<hipStream_t>(stream);
    return ADMP_OK;
}

extern "C" admp_status admp_profile(admp_ctx* ctx, int32_t enable, int32_t reset) {
    if (!ctx) return ADMP_ERR_ARG;
    if (reset) {
        for (int k = 0; k < 16; ++k) { ctx->prof_ms[k] = 0; ctx->prof_n[k] = 0; }
    }
    ctx->prof = enable != 0;
    return ADMP_OK;
}

extern "C" admp_status admp_profile_get(admp_ctx* ctx, int32_t key, double* total_ms, int64_t* count,
                                        const char** name) {
    if (!ctx || key < 0 || key >= 16) return ADMP_ERR_ARG;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return ADMP_ERR_HIP;
    prof_resolve(ctx);
    if (total_ms) *total_ms = ctx->prof_ms[key];
    if (count) *count = ctx->prof_n[key];
    if (name) *name = PROF_NAMES[key];
    return ADMP_OK;
}

extern "C" admp_status admp_ctx_destroy(admp_ctx* ctx) {
    if (!ctx) return ADMP_OK;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->ev_pending) { ctx->ev_pool.push_back(p.second.first); ctx->ev_pool.push_back(p.second.second); }
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (void* p : ctx->allocs) (void)hipFree(p);
    delete ctx;
    return ADMP_OK;
}

extern "C" admp_status admp_ctx_create(const admp_config* cfg, admp_ctx** out) {
    if (!cfg || !out) return ADMP_ERR_ARG;
    *out = nullptr;
    admp_ctx* ctx = new admp_ctx();
    ctx->cfg = *cfg;
    auto fail = [&](admp_status s) {
        fprintf(stderr, "admp_ctx_create: %s\n", ctx->err.c_str());
        admp_ctx_destroy(ctx);
        return s;
    };
    const admp_config& c = ctx->cfg;
    if (c.n_atoms <= 0 || c.n_frames <= 0 || c.lmax < 0 || c.lmax > 2) {
        ctx->err = "bad n_atoms / n_frames / lmax";
        return fail(ADMP_ERR_ARG);
    }
    if (c.lpme && c.pme_order != 6) { ctx->err = "only pme_order 6 is supported (recip.py:13)"; return fail(ADMP_ERR_ARG); }
    if (c.lpme && (c.K1 <= 0 || c.K2 <= 0 || c.K3 <= 0)) { ctx->err = "bad PME grid"; return fail(ADMP_ERR_ARG); }
    if (c.n_atoms >= (1 << NB_SHIFT)) { ctx->err = "too many atoms"; return fail(ADMP_ERR_ARG); }
    ctx->N = c.n_atoms;
    ctx->F = c.n_frames;
    ctx->nh = (c.lmax + 1) * (c.lmax + 1);
    ctx->use_frames = c.lmax > 0 ? 1 : 0;
    const size_t NF = (size_t)ctx->N * ctx->F;
    const bool mixed = c.precision == ADMP_MIXED;
    admp_status st;
    if (c.lpme) {
        if ((st = setup_grid(ctx, ctx->eg, c.K1, c.K2, c.K3, mixed)) != ADMP_OK) return fail(st);
    }
    if (c.dK1 > 0 && c.dK2 > 0 && c.dK3 > 0) {
        if ((st = setup_grid(ctx, ctx->dg, c.dK1, c.dK2, c.dK3, mixed)) != ADMP_OK) return fail(st);
    }
    size_t Gmax = std::max<size_t>(ctx->eg.G, ctx->dg.G);
#define ALLOC(p, n) do { if (dalloc(ctx, &p, n) != hipSuccess) { ctx->err = "hipMalloc failed: " #p; return fail(ADMP_ERR_HIP); } } while (0)
    ALLOC(ctx->d_box, ctx->F);
    ALLOC(ctx->d_pos, NF);
    ALLOC(ctx->d_frames, NF * 9);
    ALLOC(ctx->d_Qc, NF * 10);
    double* tmpd;
    ALLOC(tmpd, NF * SITE); ctx->d_site = tmpd;
    ALLOC(tmpd, NF * 3); ctx->d_vR = tmpd;
    ALLOC(ctx->d_mbase, NF);
    ALLOC(tmpd, NF * 3); ctx->d_wfrac = tmpd;
    ALLOC(tmpd, NF * 10); ctx->d_fq = tmpd;
    if (Gmax) {
        double* g;
        ALLOC(g, Gmax * 2 * ctx->F);
        ctx->d_grid = g;
        ctx->part_blocks = (int)nblk(Gmax, 256);
        ALLOC(ctx->d_part, (size_t)ctx->part_blocks * ctx->F * 7);
    }
    ALLOC(ctx->d_rec, ctx->F * 7);
    ALLOC(ctx->d_e_atom, NF); ALLOC(ctx->d_gpos, NF * 3); ALLOC(ctx->d_gq, NF * 10); ALLOC(ctx->d_vir_atom, NF * 9);
    ALLOC(ctx->d_gpol, NF); ALLOC(ctx->d_gth, NF); ALLOC(ctx->d_e_self, NF); ALLOC(ctx->d_e_pol, NF);
    ALLOC(ctx->d_gq_rec, NF * 10); ALLOC(ctx->d_gpos_rec, NF * 3); ALLOC(ctx->d_vir_rec, NF * 9);
    ALLOC(ctx->d_contrib, NF * 12); ALLOC(ctx->d_vir_fr, NF * 9); ALLOC(ctx->d_gQl, NF * 9);
    ALLOC(ctx->d_red, ctx->F * 16);
    ALLOC(ctx->d_red9a, ctx->F * 9); ALLOC(ctx->d_red9b, ctx->F * 9); ALLOC(ctx->d_red9c, ctx->F * 9);
    ALLOC(ctx->d_b, NF * 3); ALLOC(ctx->d_r, NF * 3); ALLOC(ctx->d_z, NF * 3); ALLOC(ctx->d_p, NF * 3);
    ALLOC(ctx->d_Ap, NF * 3); ALLOC(ctx->d_t1, NF * 3); ALLOC(ctx->d_t2, NF * 3); ALLOC(ctx->d_U, NF * 3);
    ALLOC(ctx->d_dpart, (size_t)ctx->dots_blocks * ctx->F * 3);
    ALLOC(ctx->d_dots, ctx->F * 3);
    ALLOC(ctx->d_coef, ctx->F * 4);
    ALLOC(ctx->d_active, ctx->F);
    ALLOC(ctx->d_maxcnt, 1);
    ALLOC(ctx->d_ncnt, NF);
    ALLOC(ctx->d_nhalf, NF);
    ALLOC(ctx->d_cell, NF);
    ALLOC(ctx->d_catoms, NF);
#undef ALLOC
    *out = ctx;
    return ADMP_OK;
}

extern "C" admp_status admp_set_topology(admp_ctx* ctx, const int32_t* axis_type, const int32_t* axis_idx,
                                         int64_t n_excl, const int32_t* excl_i, const int32_t* excl_j,
                                         const int32_t* excl_nbond) {
    if (!ctx) return ADMP_ERR_ARG;
    const int N = ctx->N;
    std::vector<int> at(N, 5), ai(3 * N, -1);
    if (ctx->use_frames) {
        if (!axis_type || !axis_idx) { ctx->err = "axis data required for lmax > 0"; return ADMP_ERR_ARG; }
        for (int i = 0; i < N; ++i) {
            at[i] = axis_type[i];
            if (at[i] < 0 || at[i] > 5) { ctx->err = "bad axis type"; return ADMP_ERR_ARG; }
            for (int k = 0; k < 3; ++k) {
                int v = axis_idx[3 * i + k];
                if (v >= N) { ctx->err = "axis index out of range"; return ADMP_ERR_ARG; }
                ai[3 * i + k] = v;
            }
            const int t = at[i];
            const bool needz = t != 5, needx = t != 5 && t != 4, needy = t == 2 || t == 3;
            if ((needz && ai[3 * i] < 0) || (needx && ai[3 * i + 1] < 0) || (needy && ai[3 * i + 2] < 0)) {
                ctx->err = "axis atom missing for axis type of atom " + std::to_string(i);
                return ADMP_ERR_ARG;
            }
        }
    }
    // exclusions: per-site list of (j | nbond << 28), both directions
    std::vector<std::vector<int>> ex(N);
    for (int64_t e = 0; e < n_excl; ++e) {
        int i = excl_i[e], j = excl_j[e], nb = excl_nbond[e];
        if (i < 0 || j < 0 || i >= N || j >= N || i == j) { ctx->err = "bad exclusion pair"; return ADMP_ERR_ARG; }
        if (nb < 1 || nb > 7) continue;
        ex[i].push_back(j | (nb << NB_SHIFT));
        ex[j].push_back(i | (nb << NB_SHIFT));
    }
    std::vector<int> xoff(N + 1, 0), xent;
    for (int i = 0; i < N; ++i) {
        std::sort(ex[i].begin(), ex[i].end());
        ex[i].erase(std::unique(ex[i].begin(), ex[i].end()), ex[i].end());
        xoff[i + 1] = xoff[i] + (int)ex[i].size();
        xent.insert(xent.end(), ex[i].begin(), ex[i].end());
    }
    // reverse frame map: for atom j, entries site*4+slot where site's frame uses j
    std::vector<std::vector<int>> rv(N);
    if (ctx->use_frames)
        for (int i = 0; i < N; ++i) {
            const int t = at[i];
            if (t == 5) continue;
            rv[ai[3 * i]].push_back(i * 4 + 1);
            if (t != 4) rv[ai[3 * i + 1]].push_back(i * 4 + 2);
            if (t == 2 || t == 3) rv[ai[3 * i + 2]].push_back(i * 4 + 3);
        }
    std::vector<int> roff(N + 1, 0), rent;
    for (int j = 0; j < N; ++j) {
        roff[j + 1] = roff[j] + (int)rv[j].size();
        rent.insert(rent.end(), rv[j].begin(), rv[j].end());
    }
    dfree(ctx, ctx->d_axis_type); dfree(ctx, ctx->d_axis_idx); dfree(ctx, ctx->d_xoff); dfree(ctx, ctx->d_xent);
    dfree(ctx, ctx->d_rev_off); dfree(ctx, ctx->d_rev_ent);
    HIPCHK(dalloc(ctx, &ctx->d_axis_type, N));
    HIPCHK(dalloc(ctx, &ctx->d_axis_idx, 3 * N));
    HIPCHK(dalloc(ctx, &ctx->d_xoff, N + 1));
    HIPCHK(dalloc(ctx, &ctx->d_xent, xent.size()));
    HIPCHK(dalloc(ctx, &ctx->d_rev_off, N + 1));
    HIPCHK(dalloc(ctx, &ctx->d_rev_ent, rent.size()));
    HIPCHK(hipMemcpy(ctx->d_axis_type, at.data(), N * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_axis_idx, ai.data(), 3 * N * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_xoff, xoff.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (!xent.empty()) HIPCHK(hipMemcpy(ctx->d_xent, xent.data(), xent.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_rev_off, roff.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (!rent.empty()) HIPCHK(hipMemcpy(ctx->d_rev_ent, rent.data(), rent.size() * sizeof(int), hipMemcpyHostToDevice));
    ctx->has_topo = true;
    return ADMP_OK;
}

static admp_status ensure_cap(admp_ctx* ctx, int need) {
    if (need <= ctx->cap && ctx->d_nbr) return ADMP_OK;
    int cap = std::max(64, ((int)(need * 1.25) + 63) / 64 * 64);
    dfree(ctx, ctx->d_nbr);
    dfree(ctx, ctx->d_pcoef);
    ctx->d_nbr = nullptr;
    ctx->d_pcoef = nullptr;
    const size_t NF = (size_t)ctx->N * ctx->F;
    HIPCHK(dalloc(ctx, &ctx->d_nbr, NF * cap));
    double* pc;
    HIPCHK(dalloc(ctx, &pc, NF * cap * 2));
    ctx->d_pcoef = pc;
    ctx->cap = cap;
    return ADMP_OK;
}

static admp_status upload_geometry(admp_ctx* ctx, const double* pos, const double* box) {
    const size_t NF = (size_t)ctx->N * ctx->F;
    hipLaunchKernelGGL(k_pack_pos, dim3(nblk(NF, 256)), dim3(256), 0, ctx->stream, NF, pos, ctx->d_pos);
    const admp_config& c = ctx->cfg;
    hipLaunchKernelGGL(k_box_setup, dim3(nblk(ctx->F, 64)), dim3(64), 0, ctx->stream, ctx->F, box, c.lpme ? c.K1 : 0,
                       c.lpme ? c.K2 : 0, c.lpme ? c.K3 : 0, c.dK1, c.dK2, c.dK3, ctx->d_box);
    HIPCHK(hipGetLastError());
    return ADMP_OK;
}

extern "C" admp_status admp_set_pairs(admp_ctx* ctx, int32_t frame, const void* pairs, int64_t n_pairs,
                                      int32_t pairs_is_int64) {
    if (!ctx || frame < 0 || frame >= ctx->F || (n_pairs > 0 && !pairs)) return ADMP_ERR_ARG;
    const int N = ctx->N;
    int* cnt = ctx->d_ncnt + (size_t)frame * N;
    int* bad = ctx->d_maxcnt;
    HIPCHK(hipMemsetAsync(cnt, 0, N * sizeof(int), ctx->stream));
    HIPCHK(hipMemsetAsync(bad, 0, sizeof(int), ctx->stream));
    if (n_pairs > 0) {
        if (pairs_is_int64)
            hipLaunchKernelGGL(k_upairs_count<int64_t>, dim3(nblk(n_pairs, 256)), dim3(256), 0, ctx->stream, n_pairs, N,
                               (const int64_t*)pairs, cnt, bad);
        else
            hipLaunchKernelGGL(k_upairs_count<int32_t>, dim3(nblk(n_pairs, 256)), dim3(256), 0, ctx->stream, n_pairs, N,
                               (const int32_t*)pairs, cnt, bad);
    }
    std::vector<int> h(N);
    int hbad = 0;
    HIPCHK(hipMemcpyAsync(h.data(), cnt, N * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (hbad) { ctx->err = "pair index out of range"; return ADMP_ERR_ARG; }
    int mx = 0;
    for (int v : h) mx = std::max(mx, v);
    if (mx > ctx->cap || !ctx->d_nbr) {
        admp_status s = ensure_cap(ctx, mx);
        if (s != ADMP_OK) return s;
        if (ctx->F > 1) ctx->has_nbr = false;  // other frames' lists were dropped
    }
    const int cap = ctx->cap;
    int* nbr = ctx->d_nbr + (size_t)frame * N * cap;
    int* fill = ctx->d_nhalf + (size_t)frame * N;  // scratch counters
    HIPCHK(hipMemsetAsync(fill, 0, N * sizeof(int), ctx->stream));
    if (n_pairs > 0) {
        if (pairs_is_int64)
            hipLaunchKernelGGL(k_upairs_fill<int64_t>, dim3(nblk(n_pairs, 256)), dim3(256), 0, ctx->stream, n_pairs, N,
                               cap, (const int64_t*)pairs, fill, nbr);
        else
            hipLaunchKernelGGL(k_upairs_fill<int32_t>, dim3(nblk(n_pairs, 256)), dim3(256), 0, ctx->stream, n_pairs, N,
                               cap, (const int32_t*)pairs, fill, nbr);
        int P = 1;
        while (P < mx) P <<= 1;
        hipLaunchKernelGGL(k_row_sort, dim3(N), dim3(256), P * sizeof(unsigned int), ctx->stream, N, cap, nbr, cnt);
    }
    HIPCHK(hipGetLastError());
    ctx->has_nbr = true;
    return ADMP_OK;
}

extern "C" admp_status admp_build_neighbors(admp_ctx* ctx, const double* pos, const double* box) {
    if (!ctx || !pos || !box) return ADMP_ERR_ARG;
    if (!ctx->has_topo) { ctx->err = "admp_set_topology first"; return ADMP_ERR_STATE; }
    const int N = ctx->N, F = ctx->F;
    admp_status s = upload_geometry(ctx, pos, box);
    if (s != ADMP_OK) return s;
    // cell grid from frame boxes (host copy)
    std::vector<double> hb(9 * F);
    HIPCHK(hipMemcpyAsync(hb.data(), box, 9 * F * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const double rc = ctx->cfg.rc;
    CellGrid cg;
    cg.nc[0] = cg.nc[1] = cg.nc[2] = 1 << 20;
    for (int f = 0; f < F; ++f) {
        const double* B = hb.data() + 9 * f;
        // width of the cell slab along reciprocal direction d: 1/|Binv column d|
        const double det = B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) +
                           B[2] * (B[3] * B[7] - B[4] * B[6]);
        double cof[3][3];
        cof[0][0] = B[4] * B[8] - B[5] * B[7]; cof[1][0] = B[5] * B[6] - B[3] * B[8]; cof[2][0] = B[3] * B[7] - B[4] * B[6];
        cof[0][1] = B[2] * B[7] - B[1] * B[8]; cof[1][1] = B[0] * B[8] - B[2] * B[6]; cof[2][1] = B[1] * B[6] - B[0] * B[7];
        cof[0][2] = B[1] * B[5] - B[2] * B[4]; cof[1][2] = B[2] * B[3] - B[0] * B[5]; cof[2][2] = B[0] * B[4] - B[1] * B[3];
        for (int d = 0; d < 3; ++d) {
            // Binv[c][d] = cof[c][d]/det ; |column d|
            double n2 = 0;
            for (int c = 0; c < 3; ++c) n2 += (cof[c][d] / det) * (cof[c][d] / det);
            const double width = 1.0 / std::sqrt(n2);
            int n = (int)std::floor(width / rc);
            n = std::max(1, std::min(n, 256));
            cg.nc[d] = std::min(cg.nc[d], n);
        }
    }
    cg.ncell = cg.nc[0] * cg.nc[1] * cg.nc[2];
    if (cg.ncell > ctx->ncell_alloc) {
        dfree(ctx, ctx->d_ccount); dfree(ctx, ctx->d_cstart); dfree(ctx, ctx->d_cfill);
        HIPCHK(dalloc(ctx, &ctx->d_ccount, (size_t)cg.ncell * F));
        HIPCHK(dalloc(ctx, &ctx->d_cstart, (size_t)(cg.ncell + 1) * F));
        HIPCHK(dalloc(ctx, &ctx->d_cfill, (size_t)cg.ncell * F));
        ctx->ncell_alloc = cg.ncell;
    }
    HIPCHK(hipMemsetAsync(ctx->d_ccount, 0, (size_t)cg.ncell * F * sizeof(int), ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->d_cfill, 0, (size_t)cg.ncell * F * sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_cell_assign, dim3(nblk(N, 256), F), dim3(256), 0, ctx->stream, N, ctx->d_pos, ctx->d_box, cg,
                       ctx->d_cell, ctx->d_ccount);
    hipLaunchKernelGGL(k_scan, dim3(F), dim3(1024), 0, ctx->stream, cg.ncell, ctx->d_ccount, ctx->d_cstart);
    hipLaunchKernelGGL(k_cell_fill, dim3(nblk(N, 256), F), dim3(256), 0, ctx->stream, N, cg.ncell, ctx->d_cell,
                       ctx->d_cstart, ctx->d_cfill, ctx->d_catoms);
    hipLaunchKernelGGL(k_cell_sort, dim3(nblk(cg.ncell, 256), F), dim3(256), 0, ctx->stream, cg.ncell, N, ctx->d_cstart,
                       ctx->d_catoms);
    HIPCHK(hipGetLastError());
    if (!ctx->d_nbr) {
        // first estimate from density: 1.5 x expected full-list count + margin
        double vol = 0;
        {
            const double* B = hb.data();
            vol = std::fabs(B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) +
                            B[2] * (B[3] * B[7] - B[4] * B[6]));
        }
        int est = (int)(1.5 * N / vol * 4.18879 * rc * rc * rc) + 64;
        if (ctx->cfg.max_neighbors > 0) est = ctx->cfg.max_neighbors;
        est = std::min(est, N);
        if ((s = ensure_cap(ctx, std::max(est, 1))) != ADMP_OK) return s;
    }
    for (int attempt = 0; attempt < 2; ++attempt) {
        HIPCHK(hipMemsetAsync(ctx->d_maxcnt, 0, sizeof(int), ctx->stream));
        { ProfScope ps_(ctx, 8); hipLaunchKernelGGL(k_nbr_build, dim3(nblk(N, 4), F), dim3(256), 0, ctx->stream, N, ctx->cap, ctx->d_pos,
                           ctx->d_box, cg, rc * rc, ctx->d_cell, ctx->d_cstart, ctx->d_catoms, ctx->d_xoff,
                           ctx->d_xent, ctx->d_nbr, ctx->d_ncnt, ctx->d_nhalf, ctx->d_maxcnt); }
        HIPCHK(hipGetLastError());
        int mx = 0;
        HIPCHK(hipMemcpyAsync(&mx, ctx->d_maxcnt, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (mx <= ctx->cap) break;
        if ((s = ensure_cap(ctx, mx)) != ADMP_OK) return s;
    }
    ctx->has_nbr = true;
    return ADMP_OK;
}

extern "C" admp_status admp_get_pairs(admp_ctx* ctx, int32_t frame, int32_t* out, int64_t cap, int64_t* n_out) {
    if (!ctx || frame < 0 || frame >= ctx->F || !n_out) return ADMP_ERR_ARG;
    if (!ctx->has_nbr) { ctx->err = "no neighbour list"; return ADMP_ERR_STATE; }
    const int N = ctx->N;
    // nhalf holds the j > i counts for lists built on the GPU
    int* off;
    HIPCHK(dalloc(ctx, &off, N + 1));
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, ctx->stream, N, ctx->d_nhalf + (size_t)frame * N, off);
    int total = 0;
    HIPCHK(hipMemcpyAsync(&total, off + N, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *n_out = total;
    if (out && total <= cap)
        hipLaunchKernelGGL(k_pairs_out, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, N, ctx->cap, frame, ctx->d_nbr,
                           ctx->d_ncnt, off, out);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    dfree(ctx, off);
    if (out && total > cap) { ctx->err = "output capacity too small"; return ADMP_ERR_CAPACITY; }
    return ADMP_OK;
}

// ---------------------------------------------------------------------------
// Reciprocal-space pipeline on grid g for the sources given by (Qc, U, mode).
// ---------------------------------------------------------------------------
template <class R>
static admp_status recip_run(admp_ctx* ctx, Grid& g, int use_dA, const double* Qc, const double* U, int qmode,
                             int lmax_eff, int kind, double pref, double kappa, int gather_mode, double* out3,
                             double* gq, double* gpos, double* vir, bool want_energy) {
    const int N = ctx->N, F = ctx->F;
    hipStream_t st = ctx->stream;
    Cplx<R>* grid = reinterpret_cast<Cplx<R>*>(ctx->d_grid);
    R* wfrac = reinterpret_cast<R*>(ctx->d_wfrac);
    R* fq = reinterpret_cast<R*>(ctx->d_fq);
    const size_t tot = g.G * F;
    hipLaunchKernelGGL(k_zero_cplx<R>, dim3(nblk(tot, 256)), dim3(256), 0, st, tot, grid);
    { ProfScope ps_(ctx, 3); hipLaunchKernelGGL(k_spline_prep<R>, dim3(nblk(N, 128), F), dim3(128), 0, st, N, ctx->d_pos, ctx->d_box, Qc, U, qmode,
                       use_dA, ctx->d_mbase, wfrac, fq); }
    { ProfScope ps_(ctx, 4); hipLaunchKernelGGL(k_spread_atomic<R>, dim3(nblk((size_t)N * 36, 256), F), dim3(256), 0, st, N, g.K[0], g.K[1],
                       g.K[2], ctx->d_mbase, wfrac, fq, lmax_eff, grid); }
    const int maxpts = sizeof(R) == 8 ? 2048 : 4096;
    auto fft = [&](int inverse) {
        for (int dd = 0; dd < 3; ++dd) {
            const int d = inverse ? dd : 2 - dd;
            const int K = g.K[d];
            const int nlines = (int)(g.G / K);
            int LB = std::max(1, std::min(maxpts / K, nlines));
            const size_t sm = (size_t)2 * LB * K * sizeof(Cplx<R>);
            { ProfScope ps_(ctx, 5); hipLaunchKernelGGL(k_fft_pass<R>, dim3(nblk(nlines, LB), F), dim3(256), sm, st, grid, g.K[0], g.K[1], g.K[2],
                               d, g.plan[d], reinterpret_cast<const Cplx<R>*>(g.tw[d]), inverse, LB); }
        }
    };
    fft(0);
    const int nbI = (int)nblk(g.G, 256);
    { ProfScope ps_(ctx, 6); hipLaunchKernelGGL(k_influence<R>, dim3(nbI, F), dim3(256), 0, st, g.K[0], g.K[1], g.K[2], ctx->d_box, use_dA,
                       g.bmod[0], g.bmod[1], g.bmod[2], kappa, kind, pref, grid, ctx->d_part); }
    if (want_energy)
        hipLaunchKernelGGL(k_reduce_cols, dim3(7, F), dim3(256), 0, st, nbI, 7, ctx->d_part, ctx->d_rec, 0);
    fft(1);
    { ProfScope ps_(ctx, 7); hipLaunchKernelGGL(k_gather<R>, dim3(nblk(N, 64), F), dim3(64), 0, st, N, g.K[0], g.K[1], g.K[2], ctx->d_pos,
                       ctx->d_box, use_dA, ctx->d_mbase, wfrac, fq, Qc, U, qmode, grid, gather_mode, out3, gq, gpos, vir); }
    HIPCHK(hipGetLastError());
    return ADMP_OK;
}

// A v for the induced-dipole system (everything but the stored right-hand side).
// The real-space coupling uses the per-pair coefficients stored by k_real_permfield.
template <class R>
static admp_status matvec(admp_ctx* ctx, const double* v, double* out, const double* pol, double f1) {
    const int N = ctx->N, F = ctx->F;
    const size_t NF = (size_t)N * F;
    hipStream_t st = ctx->stream;
    const R* vR;
    if (sizeof(R) == 8) {
        vR = reinterpret_cast<const R*>(v);
    } else {
        hipLaunchKernelGGL(k_cast<R>, dim3(nblk(NF * 3, 256)), dim3(256), 0, st, NF * 3, v, reinterpret_cast<R*>(ctx->d_vR));
        vR = reinterpret_cast<const R*>(ctx->d_vR);
    }
    { ProfScope ps_(ctx, 2); hipLaunchKernelGGL((k_real_matvec<R>), dim3(nblk((size_t)N * TPA, PAIR_BLOCK), F), dim3(PAIR_BLOCK), 0, st, N,
                       ctx->cap, ctx->d_pos, ctx->d_box, ctx->d_nbr, ctx->d_ncnt,
                       reinterpret_cast<const R*>(ctx->d_pcoef), vR, ctx->d_t1); }
    double* rec = nullptr;
    if (ctx->cfg.lpme) {
        admp_status s = recip_run<R>(ctx, ctx->eg, 0, nullptr, v, 2, 1, 0, DIELECTRIC, ctx->cfg.kappa, 1, ctx->d_t2,
                                     nullptr, nullptr, nullptr, false);
        if (s != ADMP_OK) return s;
        rec = ctx->d_t2;
    }
    hipLaunchKernelGGL(k_apply_diag, dim3(nblk(N, 256), F), dim3(256), 0, st, N, f1, pol, v, ctx->d_t1, rec, out);
    HIPCHK(hipGetLastError());
    return ADMP_OK;
}

static admp_status dots(admp_ctx* ctx, const double* x, const double* y, const double* z, const double* pol,
                        std::vector<double>& h) {
    const int F = ctx->F;
    hipLaunchKernelGGL(k_dots, dim3(ctx->dots_blocks, F), dim3(256), 0, ctx->stream, ctx->N, x, y, z, pol, ctx->d_dpart);
    hipLaunchKernelGGL(k_dots_final, dim3(nblk(F, 64)), dim3(64), 0, ctx->stream, F, ctx->dots_blocks, ctx->d_dpart,
                       ctx->d_dots);
    h.resize(3 * F);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->d_dots, 3 * F * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ADMP_OK;
}

struct SolveInfo {
    std::vector<int> n_iter, conv;
    std::vector<double> res;
};

// Induced-dipole solve: U (ctx->d_U) from the linear system A U = b,
// b = -dE/dU at U = 0 (real permanent field, reciprocal field, self term).
template <class R>
static admp_status solve(admp_ctx* ctx, const double* pol, const double* thole, const PairParams& pp,
                         const admp_solver_opts* o, const double* U_in, SolveInfo& info) {
    const admp_config& c = ctx->cfg;
    const int N = ctx->N, F = ctx->F;
    const size_t NF = (size_t)N * F;
    hipStream_t st = ctx->stream;
    const int lpme = c.lpme;
    const double kappa = lpme ? c.kappa : 0.0;
    const double f1 = lpme ? kappa / SQRT_PI * 2.0 * kappa * kappa / 3.0 : 0.0;
    const int lmax_eff = c.lmax == 0 ? 1 : c.lmax;
    R* site = reinterpret_cast<R*>(ctx->d_site);
    const dim3 pgrid(nblk((size_t)N * TPA, PAIR_BLOCK), F);
    double* U = ctx->d_U;
    const int kind = o ? o->kind : ADMP_SOLVER_PCG;
    const double tol = o ? o->tol : 1e-6;
    const int maxiter = o ? (kind == ADMP_SOLVER_STEPS ? o->steps : o->maxiter) : 500;
    if (o && o->warm_start && U_in)
        HIPCHK(hipMemcpyAsync(U, U_in, NF * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    else
        HIPCHK(hipMemsetAsync(U, 0, NF * 3 * sizeof(double), st));
    hipLaunchKernelGGL(k_pack_site<R>, dim3(nblk(N, 128), F), dim3(128), 0, st, N, ctx->d_Qc, pol, thole,
                       (const double*)nullptr, site);
    { ProfScope ps_(ctx, 1); hipLaunchKernelGGL((k_real_permfield<R>), pgrid, dim3(PAIR_BLOCK), 0, st, N, ctx->cap, ctx->d_pos, ctx->d_box, site,
                       ctx->d_nbr, ctx->d_ncnt, pp, ctx->d_b, reinterpret_cast<R*>(ctx->d_pcoef)); }
    HIPCHK(hipGetLastError());
    if (lpme) {
        admp_status s = recip_run<R>(ctx, ctx->eg, 0, ctx->d_Qc, nullptr, 1, lmax_eff, 0, DIELECTRIC, kappa, 1, ctx->d_t2,
                                     nullptr, nullptr, nullptr, false);
        if (s != ADMP_OK) return s;
        hipLaunchKernelGGL(k_sub, dim3(nblk(NF * 3, 256)), dim3(256), 0, st, NF * 3, ctx->d_t2, ctx->d_b);
        hipLaunchKernelGGL(k_add_self_b, dim3(nblk(N, 256), F), dim3(256), 0, st, N, f1, ctx->d_Qc, ctx->d_b);
    }
    info.n_iter.assign(F, 0);
    info.conv.assign(F, 0);
    info.res.assign(F, 0.0);
    std::vector<int> act(F, 1);
    std::vector<double> h, rz(F, 0.0);
    admp_status s;
    if (kind == ADMP_SOLVER_PCG) {
        if ((s = matvec<R>(ctx, U, ctx->d_Ap, pol, f1)) != ADMP_OK) return s;
        hipLaunchKernelGGL(k_lincomb, dim3(nblk(NF * 3, 256)), dim3(256), 0, st, NF * 3, 1.0, ctx->d_b, -1.0, ctx->d_Ap,
                           ctx->d_r);
        hipLaunchKernelGGL(k_precond, dim3(nblk(N, 256), F), dim3(256), 0, st, N, pol, ctx->d_r, ctx->d_z, ctx->d_p);
        if ((s = dots(ctx, ctx->d_r, ctx->d_z, nullptr, pol, h)) != ADMP_OK) return s;
        for (int f = 0; f < F; ++f) rz[f] = h[3 * f];
        int it = 0;
        for (it = 0; it < maxiter; ++it) {
            int nact = 0;
            for (int f = 0; f < F; ++f) {
                info.res[f] = h[3 * f + 2];
                if (act[f] && h[3 * f + 2] < tol) { act[f] = 0; info.conv[f] = 1; info.n_iter[f] = it; }
                nact += act[f];
            }
            if (!nact) break;
            HIPCHK(hipMemcpyAsync(ctx->d_active, act.data(), F * sizeof(int), hipMemcpyHostToDevice, st));
            if ((s = matvec<R>(ctx, ctx->d_p, ctx->d_Ap, pol, f1)) != ADMP_OK) return s;
            if ((s = dots(ctx, ctx->d_p, ctx->d_Ap, nullptr, pol, h)) != ADMP_OK) return s;
            std::vector<double> coef(4 * F, 0.0);
            for (int f = 0; f < F; ++f) coef[4 * f] = h[3 * f] != 0 ? rz[f] / h[3 * f] : 0.0;
            HIPCHK(hipMemcpyAsync(ctx->d_coef, coef.data(), 4 * F * sizeof(double), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_pcg_update, dim3(nblk(N, 256), F), dim3(256), 0, st, N, 0, ctx->d_coef, ctx->d_active,
                               pol, U, ctx->d_r, ctx->d_z, ctx->d_p, ctx->d_Ap);
            if ((s = dots(ctx, ctx->d_r, ctx->d_z, nullptr, pol, h)) != ADMP_OK) return s;
            for (int f = 0; f < F; ++f) {
                coef[4 * f + 1] = rz[f] != 0 ? h[3 * f] / rz[f] : 0.0;
                if (act[f]) rz[f] = h[3 * f];
            }
            HIPCHK(hipMemcpyAsync(ctx->d_coef, coef.data(), 4 * F * sizeof(double), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_pcg_update, dim3(nblk(N, 256), F), dim3(256), 0, st, N, 1, ctx->d_coef, ctx->d_active,
                               pol, U, ctx->d_r, ctx->d_z, ctx->d_p, ctx->d_Ap);
        }
        for (int f = 0; f < F; ++f)
            if (act[f]) { info.n_iter[f] = it; info.res[f] = h[3 * f + 2]; }
        // true residual of the final iterate
        if ((s = matvec<R>(ctx, U, ctx->d_Ap, pol, f1)) != ADMP_OK) return s;
        hipLaunchKernelGGL(k_lincomb, dim3(nblk(NF * 3, 256)), dim3(256), 0, st, NF * 3, 1.0, ctx->d_Ap, -1.0, ctx->d_b,
                           ctx->d_r);
        if ((s = dots(ctx, ctx->d_r, nullptr, nullptr, pol, h)) != ADMP_OK) return s;
        for (int f = 0; f < F; ++f) {
            info.res[f] = h[3 * f + 2];
            info.conv[f] = h[3 * f + 2] < tol ? 1 : 0;
        }
    } else {
        // Jacobi sweeps (pme.py:380-422)
        HIPCHK(hipMemcpyAsync(ctx->d_active, act.data(), F * sizeof(int), hipMemcpyHostToDevice, st));
        int i = 0;
        for (i = 0; i < maxiter; ++i) {
            if ((s = matvec<R>(ctx, U, ctx->d_Ap, pol, f1)) != ADMP_OK) return s;
            hipLaunchKernelGGL(k_sub, dim3(nblk(NF * 3, 256)), dim3(256), 0, st, NF * 3, ctx->d_b, ctx->d_Ap);
            if (kind == ADMP_SOLVER_REFERENCE) {
                if ((s = dots(ctx, ctx->d_Ap, nullptr, nullptr, pol, h)) != ADMP_OK) return s;
                int nact = 0;
                for (int f = 0; f < F; ++f) {
                    if (act[f]) {
                        info.res[f] = h[3 * f + 2];
                        if (h[3 * f + 2] < tol) { act[f] = 0; info.conv[f] = 1; info.n_iter[f] = i; }
                    }
                    nact += act[f];
                }
                if (!nact) break;
                HIPCHK(hipMemcpyAsync(ctx->d_active, act.data(), F * sizeof(int), hipMemcpyHostToDevice, st));
            }
            hipLaunchKernelGGL(k_jacobi, dim3(nblk(N, 256), F), dim3(256), 0, st, N, ctx->d_active, pol, ctx->d_Ap, U);
        }
        for (int f = 0; f < F; ++f) {
            if (kind == ADMP_SOLVER_STEPS) { info.n_iter[f] = maxiter; info.conv[f] = 1; }
            else if (act[f]) { info.n_iter[f] = maxiter - 1; info.conv[f] = 0; }
        }
        if (kind == ADMP_SOLVER_STEPS) {
            if ((s = matvec<R>(ctx, U, ctx->d_Ap, pol, f1)) != ADMP_OK) return s;
            hipLaunchKernelGGL(k_sub, dim3(nblk(NF * 3, 256)), dim3(256), 0, st, NF * 3, ctx->d_b, ctx->d_Ap);
            if ((s = dots(ctx, ctx->d_Ap, nullptr, nullptr, pol, h)) != ADMP_OK) return s;
            for (int f = 0; f < F; ++f) info.res[f] = h[3 * f + 2];
        }
    }
    HIPCHK(hipGetLastError());
    return ADMP_OK;
}

static PairParams make_pp(const double* mS, const double* pS, const double* dS, int n, double kappa) {
    PairParams pp;
    memset(&pp, 0, sizeof(pp));
    pp.nscale = std::max(1, std::min(n, 8));
    for (int k = 0; k < pp.nscale; ++k) {
        pp.ms[k] = mS ? mS[k] : 1.0;
        pp.ps[k] = pS ? pS[k] : 1.0;
        pp.ds[k] = dS ? dS[k] : 1.0;
    }
    pp.kappa = kappa;
    return pp;
}

template <class R>
static admp_status compute_impl(admp_ctx* ctx, const double* Q_local, const double* pol, const double* thole,
                                const PairParams& pp, double* U_inout, const admp_solver_opts* opts, int want_grad,
                                double* energy, double* dpos, double* dbox, double* dQ_local, double* dpol,
                                double* dthole, admp_stats* stats) {
    const admp_config& c = ctx->cfg;
    const int N = ctx->N, F = ctx->F;
    const size_t NF = (size_t)N * F;
    hipStream_t st = ctx->stream;
    const int lpol = c.lpol, lpme = c.lpme;
    const double kappa = lpme ? c.kappa : 0.0;
    const int lmax_eff = (lpol && c.lmax == 0) ? 1 : c.lmax;
    R* site = reinterpret_cast<R*>(ctx->d_site);
    const dim3 pgrid(nblk((size_t)N * TPA, PAIR_BLOCK), F);
    { ProfScope ps_(ctx, 9); hipLaunchKernelGGL(k_frames_fwd, dim3(nblk(N, 128), F), dim3(128), 0, st, N, ctx->nh, ctx->use_frames, ctx->d_pos,
                       ctx->d_box, ctx->d_axis_type, ctx->d_axis_idx, Q_local, ctx->d_frames, ctx->d_Qc); }
    HIPCHK(hipGetLastError());
    SolveInfo info;
    admp_status s;
    const double* U = nullptr;
    if (lpol) {
        if ((s = solve<R>(ctx, pol, thole, pp, opts, U_inout, info)) != ADMP_OK) return s;
        U = ctx->d_U;
        if (U_inout) HIPCHK(hipMemcpyAsync(U_inout, U, NF * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    // ---------------- final pass: energy and gradients with U fixed ----------------
    hipLaunchKernelGGL(k_pack_site<R>, dim3(nblk(N, 128), F), dim3(128), 0, st, N, ctx->d_Qc, pol, thole, U, site);
    if (lpol)
        { ProfScope ps_(ctx, 0); hipLaunchKernelGGL((k_real_final<R, true>), pgrid, dim3(PAIR_BLOCK), 0, st, N, ctx->cap, ctx->d_pos, ctx->d_box,
                           site, ctx->d_nbr, ctx->d_ncnt, pp, ctx->d_e_atom, ctx->d_gpos, ctx->d_gq, ctx->d_vir_atom,
                           ctx->d_gpol, ctx->d_gth); }
    else
        { ProfScope ps_(ctx, 0); hipLaunchKernelGGL((k_real_final<R, false>), pgrid, dim3(PAIR_BLOCK), 0, st, N, ctx->cap, ctx->d_pos, ctx->d_box,
                           site, ctx->d_nbr, ctx->d_ncnt, pp, ctx->d_e_atom, ctx->d_gpos, ctx->d_gq, ctx->d_vir_atom,
                           ctx->d_gpol, ctx->d_gth); }
    HIPCHK(hipGetLastError());
    if (!lpol) {
        HIPCHK(hipMemsetAsync(ctx->d_gpol, 0, NF * sizeof(double), st));
        HIPCHK(hipMemsetAsync(ctx->d_gth, 0, NF * sizeof(double), st));
    }
    if (lpme) {
        if ((s = recip_run<R>(ctx, ctx->eg, 0, ctx->d_Qc, U, lpol ? 3 : 1, lmax_eff, 0, DIELECTRIC, kappa, 0, nullptr,
                              ctx->d_gq_rec, ctx->d_gpos_rec, ctx->d_vir_rec, true)) != ADMP_OK)
            return s;
    } else {
        HIPCHK(hipMemsetAsync(ctx->d_gq_rec, 0, NF * 10 * sizeof(double), st));
        HIPCHK(hipMemsetAsync(ctx->d_gpos_rec, 0, NF * 3 * sizeof(double), st));
        HIPCHK(hipMemsetAsync(ctx->d_vir_rec, 0, NF * 9 * sizeof(double), st));
        HIPCHK(hipMemsetAsync(ctx->d_rec, 0, F * 7 * sizeof(double), st));
    }
    hipLaunchKernelGGL(k_self, dim3(nblk(N, 256), F), dim3(256), 0, st, N, kappa, lpme, lpol, ctx->d_Qc, U, pol,
                       ctx->d_e_self, ctx->d_e_pol, ctx->d_gq, lpol ? ctx->d_gpol : nullptr);
    // per-frame energy components: real, self, polarisation penalty -> d_red [0..F), [F..2F), [2F..3F)
    hipLaunchKernelGGL(k_reduce_cols, dim3(1, F), dim3(256), 0, st, N, 1, ctx->d_e_atom, ctx->d_red, 0);
    hipLaunchKernelGGL(k_reduce_cols, dim3(1, F), dim3(256), 0, st, N, 1, ctx->d_e_self, ctx->d_red + F, 0);
    hipLaunchKernelGGL(k_reduce_cols, dim3(1, F), dim3(256), 0, st, N, 1, ctx->d_e_pol, ctx->d_red + 2 * F, 0);
    std::vector<double> comp(3 * F);
    HIPCHK(hipMemcpyAsync(comp.data(), ctx->d_red, 3 * F * sizeof(double), hipMemcpyDeviceToHost, st));
    std::vector<double> rec(7 * F);
    HIPCHK(hipMemcpyAsync(rec.data(), ctx->d_rec, 7 * F * sizeof(double), hipMemcpyDeviceToHost, st));
    if (want_grad) {
        // dE/dQ_global total -> frames backward
        hipLaunchKernelGGL(k_add3, dim3(nblk(NF * 10, 256)), dim3(256), 0, st, NF * 10, ctx->d_gq, ctx->d_gq_rec,
                           (const double*)nullptr, ctx->d_gq);
        { ProfScope ps_(ctx, 9); hipLaunchKernelGGL(k_frames_bwd, dim3(nblk(N, 128), F), dim3(128), 0, st, N, ctx->nh, ctx->use_frames, ctx->d_pos,
                           ctx->d_box, ctx->d_axis_type, ctx->d_axis_idx, Q_local, ctx->d_gq, ctx->d_gQl,
                           ctx->d_contrib, ctx->d_vir_fr); }
        hipLaunchKernelGGL(k_add3, dim3(nblk(NF * 3, 256)), dim3(256), 0, st, NF * 3, ctx->d_gpos, ctx->d_gpos_rec,
                           (const double*)nullptr, ctx->d_gpos);
        hipLaunchKernelGGL(k_frames_gather, dim3(nblk(N, 256), F), dim3(256), 0, st, N, ctx->d_rev_off, ctx->d_rev_ent,
                           ctx->d_contrib, ctx->d_gpos);
        // box: direct terms (real pairs + frames), reciprocal A-terms and C_k terms
        hipLaunchKernelGGL(k_reduce_cols, dim3(9, F), dim3(256), 0, st, N, 9, ctx->d_vir_atom, ctx->d_red9a, 0);
        hipLaunchKernelGGL(k_reduce_cols, dim3(9, F), dim3(256), 0, st, N, 9, ctx->d_vir_fr, ctx->d_red9a, 1);
        hipLaunchKernelGGL(k_reduce_cols, dim3(9, F), dim3(256), 0, st, N, 9, ctx->d_vir_rec, ctx->d_red9b, 0);
        if (dbox)
            hipLaunchKernelGGL(k_box_grad, dim3(nblk(F, 64)), dim3(64), 0, st, F, ctx->d_box, ctx->d_red9a,
                               lpme ? ctx->d_red9b : nullptr, (const double*)nullptr, lpme ? ctx->d_rec : nullptr,
                               c.K1, c.K2, c.K3, 0, dbox, 0);
        if (dpos) HIPCHK(hipMemcpyAsync(dpos, ctx->d_gpos, NF * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (dQ_local)
            HIPCHK(hipMemcpyAsync(dQ_local, ctx->d_gQl, NF * ctx->nh * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (dpol) HIPCHK(hipMemcpyAsync(dpol, ctx->d_gpol, NF * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (dthole) HIPCHK(hipMemcpyAsync(dthole, ctx->d_gth, NF * sizeof(double), hipMemcpyDeviceToDevice, st));
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st));
    std::vector<double> e(F);
    for (int f = 0; f < F; ++f) e[f] = comp[f] + comp[F + f] + comp[2 * F + f] + rec[7 * f];
    if (energy) HIPCHK(hipMemcpy(energy, e.data(), F * sizeof(double), hipMemcpyHostToDevice));
    admp_status ret = ADMP_OK;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        for (int f = 0; f < F && f < 64; ++f) {
            if (lpol) {
                stats->n_iter[f] = info.n_iter[f];
                stats->converged[f] = info.conv[f];
                stats->residual[f] = info.res[f];
            } else {
                stats->converged[f] = 1;
            }
        }
        stats->e_real = comp[0];
        stats->e_self = comp[F];
        stats->e_pol = comp[2 * F];
        stats->e_recip = rec[0];
    }
    if (lpol)
        for (int f = 0; f < F; ++f)
            if (!info.conv[f]) ret = ADMP_WARN_NOT_CONVERGED;
    return ret;
}

extern "C" admp_status admp_compute(admp_ctx* ctx, const double* pos, const double* box, const double* Q_local,
                                    const double* pol, const double* thole, const double* mScales,
                                    const double* pScales, const double* dScales, int32_t n_scales, double* U_inout,
                                    const admp_solver_opts* opts, int32_t want_grad, double* energy, double* dpos,
                                    double* dbox, double* dQ_local, double* dpol, double* dthole, admp_stats* stats) {
    if (!ctx || !pos || !box || !Q_local) return ADMP_ERR_ARG;
    if (!ctx->has_topo) { ctx->err = "admp_set_topology first"; return ADMP_ERR_STATE; }
    if (!ctx->has_nbr) { ctx->err = "no pair list: admp_set_pairs or admp_build_neighbors first"; return ADMP_ERR_STATE; }
    if (ctx->cfg.lpol && (!pol || !thole)) { ctx->err = "pol and thole required"; return ADMP_ERR_ARG; }
    if (n_scales < 1 || n_scales > 8) { ctx->err = "n_scales must be 1..8"; return ADMP_ERR_ARG; }
    admp_status s = upload_geometry(ctx, pos, box);
    if (s != ADMP_OK) return s;
    PairParams pp = make_pp(mScales, pScales, dScales, n_scales, ctx->cfg.lpme ? ctx->cfg.kappa : 0.0);
    if (ctx->cfg.precision == ADMP_MIXED)
        return compute_impl<float>(ctx, Q_local, pol, thole, pp, U_inout, opts, want_grad, energy, dpos, dbox, dQ_local,
                                   dpol, dthole, stats);
    return compute_impl<double>(ctx, Q_local, pol, thole, pp, U_inout, opts, want_grad, energy, dpos, dbox, dQ_local,
                                dpol, dthole, stats);
}

// ---------------------------------------------------------------------------
// Dispersion PME and Slater short range
// ---------------------------------------------------------------------------
template <class R>
static admp_status disp_impl(admp_ctx* ctx, const double* c_list, const PairParams& pp, int want_grad, double* energy,
                             double* dpos, double* dbox, double* dc) {
    const admp_config& c = ctx->cfg;
    const int N = ctx->N, F = ctx->F;
    const size_t NF = (size_t)N * F;
    hipStream_t st = ctx->stream;
    const bool lpme = ctx->dg.G > 0;
    const double kappa = lpme ? c.disp_kappa : 0.0;
    PairParams p2 = pp;
    p2.kappa = kappa;
    { ProfScope ps_(ctx, 10); hipLaunchKernelGGL((k_pair_short<R, 0>), dim3(nblk((size_t)N * TPA, PAIR_BLOCK), F), dim3(PAIR_BLOCK), 0, st, N,
                       ctx->cap, ctx->d_pos, ctx->d_box, c_list, ctx->d_nbr, ctx->d_ncnt, p2, c.disp_pmax, ctx->d_e_atom,
                       ctx->d_gpos, ctx->d_vir_atom, ctx->d_gpos_rec /* (F,N,3) dE/dc */); }
    double* gc = ctx->d_gpos_rec;
    HIPCHK(hipMemsetAsync(ctx->d_red9b, 0, F * 9 * sizeof(double), st));
    std::vector<double> erec(F, 0.0);
    if (lpme) {
        // per-grid: dispersion coefficient as a charge, lmax 0, k = 0 included, no DIELECTRIC
        double* Qd = ctx->d_gq_rec;  // (F,N,10) scratch source
        double* gq = ctx->d_gq;      // (F,N,10) scratch gradient
        const int kinds[3] = {6, 8, 10};
        const int np = c.disp_pmax >= 10 ? 3 : (c.disp_pmax >= 8 ? 2 : 1);
        HIPCHK(hipMemsetAsync(ctx->d_red9c, 0, F * 9 * sizeof(double), st));
        std::vector<double> wsum(7 * F, 0.0);
        for (int g = 0; g < np; ++g) {
            hipLaunchKernelGGL(k_disp_source, dim3(nblk(N, 256), F), dim3(256), 0, st, N, g, c_list, Qd);
            admp_status s = recip_run<R>(ctx, ctx->dg, 1, Qd, nullptr, 1, 0, kinds[g], 1.0, kappa, 0, nullptr, gq,
                                         ctx->d_gpos_rec == gc ? ctx->d_t1 : ctx->d_gpos_rec, ctx->d_vir_rec, true);
            if (s != ADMP_OK) return s;
            // recip grads: dE/dc (charge gradient), positions (into t1 scratch), A-term virial
            hipLaunchKernelGGL(k_disp_collect, dim3(nblk(NF, 256)), dim3(256), 0, st, NF, g, gq, gc);
            hipLaunchKernelGGL(k_vadd, dim3(nblk(NF * 3, 256)), dim3(256), 0, st, NF * 3, ctx->d_t1, ctx->d_gpos);
            hipLaunchKernelGGL(k_reduce_cols, dim3(9, F), dim3(256), 0, st, N, 9, ctx->d_vir_rec, ctx->d_red9c, 1);
            std::vector<double> r7(7 * F);
            HIPCHK(hipMemcpyAsync(r7.data(), ctx->d_rec, 7 * F * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (int k = 0; k < 7 * F; ++k) wsum[k] += r7[k];
        }
        for (int f = 0; f < F; ++f) erec[f] = wsum[7 * f];
        HIPCHK(hipMemcpyAsync(ctx->d_rec, wsum.data(), 7 * F * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_disp_self, dim3(nblk(N, 256), F), dim3(256), 0, st, N, kappa, c.disp_pmax, c_list,
                           ctx->d_e_self, gc);
    } else {
        HIPCHK(hipMemsetAsync(ctx->d_e_self, 0, NF * sizeof(double), st));
    }
    hipLaunchKernelGGL(k_reduce_cols, dim3(1, F), dim3(256), 0, st, N, 1, ctx->d_e_atom, ctx->d_red9a, 0);
    hipLaunchKernelGGL(k_reduce_cols, dim3(1, F), dim3(256), 0, st, N, 1, ctx->d_e_self, ctx->d_red9a + F, 0);
    std::vector<double> comp(2 * F);
    HIPCHK(hipMemcpyAsync(comp.data(), ctx->d_red9a, 2 * F * sizeof(double), hipMemcpyDeviceToHost, st));
    if (want_grad) {
        hipLaunchKernelGGL(k_reduce_cols, dim3(9, F), dim3(256), 0, st, N, 9, ctx->d_vir_atom, ctx->d_red9b, 0);
        if (dbox)
            hipLaunchKernelGGL(k_box_grad, dim3(nblk(F, 64)), dim3(64), 0, st, F, ctx->d_box, ctx->d_red9b,
                               lpme ? ctx->d_red9c : nullptr, (const double*)nullptr, lpme ? ctx->d_rec : nullptr, c.dK1,
                               c.dK2, c.dK3, 1, dbox, 0);
        if (dpos) HIPCHK(hipMemcpyAsync(dpos, ctx->d_gpos, NF * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (dc) HIPCHK(hipMemcpyAsync(dc, gc, NF * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    std::vector<double> e(F);
    for (int f = 0; f < F; ++f) e[f] = comp[f] + comp[F + f] + erec[f];
    if (energy) HIPCHK(hipMemcpy(energy, e.data(), F * sizeof(double), hipMemcpyHostToDevice));
    return ADMP_OK;
}

extern "C" admp_status admp_disp_compute(admp_ctx* ctx, const double* pos, const double* box, const double* c_list,
                                         const double* mScales, int32_t n_scales, int32_t want_grad, double* energy,
                                         double* dpos, double* dbox, double* dc) {
    if (!ctx || !pos || !box || !c_list) return ADMP_ERR_ARG;
    if (!ctx->has_nbr) { ctx->err = "no pair list"; return ADMP_ERR_STATE; }
    if (n_scales < 1 || n_scales > 8) { ctx->err = "n_scales must be 1..8"; return ADMP_ERR_ARG; }
    admp_status s = upload_geometry(ctx, pos, box);
    if (s != ADMP_OK) return s;
    PairParams pp = make_pp(mScales, nullptr, nullptr, n_scales, 0.0);
    if (ctx->cfg.precision == ADMP_MIXED) return disp_impl<float>(ctx, c_list, pp, want_grad, energy, dpos, dbox, dc);
    return disp_impl<double>(ctx, c_list, pp, want_grad, energy, dpos, dbox, dc);
}

extern "C" admp_status admp_slater_compute(admp_ctx* ctx, const double* pos, const double* box, const double* a,
                                           const double* b, const double* mScales, int32_t n_scales, int32_t want_grad,
                                           double* energy, double* dpos, double* dbox, double* da, double* db) {
    if (!ctx || !pos || !box || !a || !b) return ADMP_ERR_ARG;
    if (!ctx->has_nbr) { ctx->err = "no pair list"; return ADMP_ERR_STATE; }
    if (n_scales < 1 || n_scales > 8) { ctx->err = "n_scales must be 1..8"; return ADMP_ERR_ARG; }
    admp_status s = upload_geometry(ctx, pos, box);
    if (s != ADMP_OK) return s;
    const int N = ctx->N, F = ctx->F;
    const size_t NF = (size_t)N * F;
    hipStream_t st = ctx->stream;
    PairParams pp = make_pp(mScales, nullptr, nullptr, n_scales, 0.0);
    // interleave (A, B) per site into t1 (N x 2 <= NF x 3)
    std::vector<double> ha(N), hb(N), ab(2 * N);
    HIPCHK(hipMemcpyAsync(ha.data(), a, N * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hb.data(), b, N * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int i = 0; i < N; ++i) { ab[2 * i] = ha[i]; ab[2 * i + 1] = hb[i]; }
    HIPCHK(hipMemcpyAsync(ctx->d_t1, ab.data(), 2 * N * sizeof(double), hipMemcpyHostToDevice, st));
    if (ctx->cfg.precision == ADMP_MIXED)
        { ProfScope ps_(ctx, 10); hipLaunchKernelGGL((k_pair_short<float, 1>), dim3(nblk((size_t)N * TPA, PAIR_BLOCK), F), dim3(PAIR_BLOCK), 0, st,
                           N, ctx->cap, ctx->d_pos, ctx->d_box, ctx->d_t1, ctx->d_nbr, ctx->d_ncnt, pp, 0, ctx->d_e_atom,
                           ctx->d_gpos, ctx->d_vir_atom, ctx->d_gpos_rec); }
    else
        { ProfScope ps_(ctx, 10); hipLaunchKernelGGL((k_pair_short<double, 1>), dim3(nblk((size_t)N * TPA, PAIR_BLOCK), F), dim3(PAIR_BLOCK), 0, st,
                           N, ctx->cap, ctx->d_pos, ctx->d_box, ctx->d_t1, ctx->d_nbr, ctx->d_ncnt, pp, 0, ctx->d_e_atom,
                           ctx->d_gpos, ctx->d_vir_atom, ctx->d_gpos_rec); }
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_cols, dim3(1, F), dim3(256), 0, st, N, 1, ctx->d_e_atom, ctx->d_red9a, 0);
    std::vector<double> e(F);
    HIPCHK(hipMemcpyAsync(e.data(), ctx->d_red9a, F * sizeof(double), hipMemcpyDeviceToHost, st));
    if (want_grad) {
        hipLaunchKernelGGL(k_reduce_cols, dim3(9, F), dim3(256), 0, st, N, 9, ctx->d_vir_atom, ctx->d_red9b, 0);
        if (dbox)
            hipLaunchKernelGGL(k_box_grad, dim3(nblk(F, 64)), dim3(64), 0, st, F, ctx->d_box, ctx->d_red9b,
                               (const double*)nullptr, (const double*)nullptr, (const double*)nullptr, 1, 1, 1, 0, dbox, 0);
        if (dpos) HIPCHK(hipMemcpyAsync(dpos, ctx->d_gpos, NF * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (da) hipLaunchKernelGGL(k_copy_cols, dim3(nblk(NF, 256)), dim3(256), 0, st, NF, 1, 3, 0, ctx->d_gpos_rec, da);
        if (db) hipLaunchKernelGGL(k_copy_cols, dim3(nblk(NF, 256)), dim3(256), 0, st, NF, 1, 3, 1, ctx->d_gpos_rec, db);
    }
    HIPCHK(hipStreamSynchronize(st));
    if (energy) HIPCHK(hipMemcpy(energy, e.data(), F * sizeof(double), hipMemcpyHostToDevice));
    return ADMP_OK;
}
