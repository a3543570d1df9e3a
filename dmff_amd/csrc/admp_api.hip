// Host side of the C ABI declared in include/admp_hip.h.
#include "../../include/admp_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <initializer_list>
#include <string>
#include <type_traits>
#include <vector>

#include "admp_common.h"
#include "nblist_kernels.h"
#include "pair_kernels.h"
#include "recip_kernels.h"
#include "recip2_kernels.h"
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
    bool plane[2] = {false, false};  // fused plane FFT pass usable ([0] float, [1] double arithmetic)
};

}  // namespace

struct admp_ctx {
    admp_config cfg;
    hipStream_t stream = nullptr;       // the context's own stream (capturable)
    hipStream_t user_stream = nullptr;  // caller's stream: ordered with events on entry and exit
    hipStream_t own = nullptr;          // the context's own stream (stream is own, or the caller's in caller mode)
    int caller_mode = 0;                // admp_set_stream_mode: work directly on the caller's stream
    hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_done = nullptr;
    hipStream_t side = nullptr;         // second context stream: real-space products beside the reciprocal chain
    PcgState* h_pcg = nullptr;          // pinned host copies of the PCG state (two slots of F)
    // neighbour build whose row-count check is still pending (resolved by nbr_settle)
    bool nbr_pending = false;
    int* h_maxcnt = nullptr;            // pinned copy of the longest row
    int* h_geom = nullptr;              // pinned copy of the geometry flags (GEOM_BAD_*)
    int* d_geom = nullptr;              // geometry flags of the last upload_geometry
    double* h_comp = nullptr;           // pinned energy components (3 F real/self/pol, 7 F reciprocal)
    hipEvent_t ev_nbr = nullptr;
    hipEvent_t ev_geom = nullptr;      // after the compute's geometry flags were copied to h_geom
    hipEvent_t ev_state[2] = {nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::string err;
    int N = 0, F = 1, nh = 1;
    int cap = 0;
    bool has_topo = false;
    std::vector<char> frame_nbr;  // per frame: its pair list is in d_nbr / d_ncnt
    bool nbr_built = false;       // lists (and d_nhalf) come from admp_build_neighbors
    CellGrid nbr_cg{};            // cell grid of the last list build
    bool cg_host = false;         // nbr_cg was sized on the host from a box read back
    bool cg_spec = false;         // the pending build took nbr_cg over unchecked (k_cell_grid_check)
    int* d_cgflag = nullptr;      // per frame: the box needs another cell grid
    int* h_cgflag = nullptr;      // pinned copy
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
    double4* d_spos = nullptr;  // positions in cell order (atom index in .w)
    int ncell_alloc = 0;
    // per-site work (allocated for double, reused for float)
    void *d_site = nullptr, *d_pcoef = nullptr, *d_vR = nullptr;
    int4* d_mbase = nullptr;
    void *d_wfrac = nullptr, *d_fq = nullptr;
    // grids
    Grid eg, dg;
    void* d_grid = nullptr;     // complex, F * max(G)   (small-grid fallback path)
    void* d_rgrid = nullptr;    // real grid, F * max(G)
    void* d_hgrid = nullptr;    // half spectrum, F * max(K0 K1 (K2/2+1)) complex
    double* d_wd = nullptr;     // per-site grid offsets w (double)
    int *d_gcell = nullptr, *d_gccount = nullptr, *d_gcstart = nullptr, *d_gcfill = nullptr, *d_gcatoms = nullptr;
    int *d_bcount = nullptr, *d_bstart = nullptr;  // per-brick candidate counts / offsets
    int2* d_blist = nullptr;                        // per-brick candidate lists
    double* d_infl = nullptr;                       // tabulated electrostatic influence function (eg)
    int infl_pm = 0;                                // its layout: plane-major (1) or line-major (0)
    bool plane_lds_ok = false;                      // plane kernels may use PLANE_LDS_MAX bytes of LDS
    long blist_cap = 0;
    int* d_scan_tmp = nullptr;  // tile totals of scan_rows
    size_t scan_cap = 0;
    const Grid* prepped = nullptr;  // grid whose site bases / sort are current
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
    double *d_dpart = nullptr, *d_dots = nullptr, *d_coef = nullptr, *d_res = nullptr, *d_dU = nullptr;
    PcgState* d_pcg = nullptr;
    double* d_s = nullptr;      // CG: s = A p by recurrence (r and s: two buffers each)
    int* d_pnear = nullptr;     // nearest coupled neighbour slot per site (-1: none)
    int* d_pcj = nullptr;       // block preconditioner partners, 3 per site (-1: unused)
    double* d_pcd = nullptr;    // the site's rows of its block inverse, 3 x 12 per site
    int* d_pcn = nullptr;       // children per pair site (k_pc_children)
    int* d_pcc = nullptr;       // the children, PC_MAXCH per site
    int pc_maxm = PC_MAXM;      // sites per preconditioner block (1: diagonal; admp_solver_opts.precond)
    CgCheck spread_chk{};       // convergence test of the next solver product (set by pcg_apply)
    int* d_active = nullptr;
    int dots_blocks = 64;
    double* d_pparts = nullptr;               // CG partials (CgParts layout)
    int cg_nw = 0;                            // z.w slots per frame
    int cg_nwu = 0;                           // z.w slots written by the last product
    double *d_swfrac = nullptr, *d_sfq = nullptr;  // solver spline coordinates / fractional dipoles (R-typed)
    int pcg_nb = 0;
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;  // parallel to allocs
    // captured PCG iteration chunks (replayed until the frames converge)
    struct ChunkGraph {
        hipGraphExec_t exec = nullptr;
        int type = -1, nrun = 0, par = 0;
        const void *rhs = nullptr, *x = nullptr;
        double tol = 0, f1 = 0;
        bool pc = true;
    };
    ChunkGraph graphs[12];
    ChunkGraph chunk_seen[16];   // chunk variants met once (captured on their second use)
    int chunk_seen_next = 0;
    // the solver set-up (right-hand side, preconditioner, CG start) and the
    // final pass, captured on their second use with the same arguments
    struct PhaseGraph {
        hipGraphExec_t exec = nullptr;
        std::string key, seen;
    };
    PhaseGraph g_setup, g_final;
    bool use_graphs = true;   // ADMP_NO_GRAPH=1 runs the chunk and phase launches eagerly
    double* d_pol = nullptr;  // context-owned copies of the caller's inputs (stable graph arguments):
    double* d_thole = nullptr;  // polarisabilities, Thole widths, local multipoles
    double* d_Qloc = nullptr;
    double* d_eout = nullptr;   // energy per frame and dE/dbox, copied to the caller's outputs
    double* d_bout = nullptr;
    // per-kernel timing (HIP events around launches on the context stream)
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    double prof_ms[16] = {0};
    long prof_n[16] = {0};
};

static const char* PROF_NAMES[16] = {"real_final", "real_permfield", "real_matvec", "spline_prep", "spread",
                                     "fft", "influence", "gather", "nbr_build", "frames", "pair_short",
                                     "solver_vec", "solver_graph", "plane_matvec", "", ""};

static void clear_chunk_graphs(admp_ctx* ctx) {
    for (auto& g : ctx->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = admp_ctx::ChunkGraph();
    }
    for (auto& g : ctx->chunk_seen) g = admp_ctx::ChunkGraph();
}

static void clear_graphs(admp_ctx* ctx) {
    clear_chunk_graphs(ctx);
    for (admp_ctx::PhaseGraph* g : {&ctx->g_setup, &ctx->g_final}) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        *g = admp_ctx::PhaseGraph();
    }
}


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
    hipStream_t sp;
    ProfScope(admp_ctx* c, int k, hipStream_t s = nullptr) : ctx(c), key(k), sp(s ? s : c->stream) {
        if (ctx->prof) {
            a = ev_get(ctx);
            (void)hipEventRecord(a, sp);
        }
    }
    ~ProfScope() {
        if (ctx->prof) {
            hipEvent_t b = ev_get(ctx);
            (void)hipEventRecord(b, sp);
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

// Run body() through a captured graph cached under key: eagerly on the first
// use of a key (its lazy allocations happen then, outside any capture), captured
// and launched on the second, replayed afterwards.  Eager whenever graphs are
// off, profiling is on, or capture is unavailable (the caller's stream being
// captured already, for instance).
template <class Body>
static admp_status run_phase(admp_ctx* ctx, admp_ctx::PhaseGraph& pg, const std::string& key, Body body) {
    if (!ctx->use_graphs || ctx->prof) return body();
    if (pg.exec && pg.key == key) {
        HIPCHK(hipGraphLaunch(pg.exec, ctx->stream));
        return ADMP_OK;
    }
    if (pg.seen != key) {
        pg.seen = key;
        return body();
    }
    if (pg.exec) {
        (void)hipGraphExecDestroy(pg.exec);
        pg.exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return body();
    }
    const admp_status s = body();
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
    hipGraphExec_t exec = nullptr;
    hipError_t ie = hipErrorUnknown;
    if (s == ADMP_OK && ce == hipSuccess && graph) ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        // nothing of the capture ran: report the body's error, or run it eagerly
        (void)hipGetLastError();
        if (s != ADMP_OK) return s;
        ctx->err.clear();
        return body();
    }
    pg.exec = exec;
    pg.key = key;
    HIPCHK(hipGraphLaunch(pg.exec, ctx->stream));
    return ADMP_OK;
}

// key bytes of a value (phase graph keys)
template <class T> static void key_add(std::string& k, const T& v) {
    k.append(reinterpret_cast<const char*>(&v), sizeof(T));
}

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }
// blocks of the gather (one z.w partial of the CG epilogue each)
static inline int gather_blocks(int N) { return (N + GATHER_SITES_PER_BLOCK - 1) / GATHER_SITES_PER_BLOCK; }

template <class T> static hipError_t dalloc(admp_ctx* ctx, T** p, size_t n) {
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    // stream-ordered allocation on the context stream: no device-wide
    // synchronisation, so contexts driven from other threads (whose streams may
    // be capturing HIP graphs) are not disturbed.  Every buffer starts zeroed.
    hipError_t e = hipMallocAsync(&q, bytes, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(q, 0, bytes, ctx->stream);
    if (e == hipSuccess) {
        ctx->allocs.push_back(q);
        ctx->alloc_bytes.push_back(bytes);
        *p = reinterpret_cast<T*>(q);
    }
    return e;
}

static void dfree(admp_ctx* ctx, void* p) {
    if (!p) return;
    auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), p);
    if (it != ctx->allocs.end()) {
        ctx->alloc_bytes.erase(ctx->alloc_bytes.begin() + (it - ctx->allocs.begin()));
        ctx->allocs.erase(it);
    }
    (void)hipFreeAsync(p, ctx->stream);
}

static bool all_frames_have_pairs(const admp_ctx* ctx) {
    if ((int)ctx->frame_nbr.size() != ctx->F) return false;
    for (char c : ctx->frame_nbr)
        if (!c) return false;
    return true;
}

static std::vector<int> factorize(int K) {
    std::vector<int> f;
    int n = K;
    while (n % 4 == 0) { f.push_back(4); n /= 4; }
    while (n % 2 == 0) { f.push_back(2); n /= 2; }
    for (int p = 3; p * p <= n; p += 2)
        while (n % p == 0) { f.push_back(p); n /= p; }
    if (n > 1) f.push_back(n);
    // a 2 and a 5 become one radix-10 stage (one LDS pass and barrier fewer)
    for (;;) {
        auto i2 = std::find(f.begin(), f.end(), 2), i5 = std::find(f.begin(), f.end(), 5);
        if (i2 == f.end() || i5 == f.end()) break;
        *i5 = 10;
        f.erase(i2);
    }
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

// dynamic LDS of k_fxy_plane: two K0 x K1 planes and the x and y twiddles
static size_t plane_smem(int K0, int K1, size_t cs) { return ((size_t)2 * K0 * K1 + K0 + K1) * cs; }
constexpr size_t PLANE_LDS_MAX = 81920;   // 80 KiB: two blocks fill a CU's 160 KiB

// The dynamic-LDS ceiling of the plane kernels is a process-wide attribute of
// each kernel: set it once to PLANE_LDS_MAX (every plane launch stays below it),
// never per grid, so that contexts with different grids cannot lower it for
// one another.
static bool plane_kernels_ready() {
    static const bool ok = [] {
        int dev = 0, lim = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
            (size_t)lim < PLANE_LDS_MAX) {
            (void)hipGetLastError();
            return false;
        }
        const int m = (int)PLANE_LDS_MAX;
        const hipError_t e[6] = {
            hipFuncSetAttribute((const void*)k_fxy_plane<double, false>, hipFuncAttributeMaxDynamicSharedMemorySize, m),
            hipFuncSetAttribute((const void*)k_fxy_plane<double, true>, hipFuncAttributeMaxDynamicSharedMemorySize, m),
            hipFuncSetAttribute((const void*)k_fxy_plane<float, false>, hipFuncAttributeMaxDynamicSharedMemorySize, m),
            hipFuncSetAttribute((const void*)k_fxy_plane<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize, m),
            hipFuncSetAttribute((const void*)k_plane_matvec<double>, hipFuncAttributeMaxDynamicSharedMemorySize, m),
            hipFuncSetAttribute((const void*)k_plane_matvec<float>, hipFuncAttributeMaxDynamicSharedMemorySize, m)};
        for (hipError_t x : e)
            if (x != hipSuccess) {
                (void)hipGetLastError();
                return false;
            }
        return true;
    }();
    return ok;
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
        HIPCHK(hipMemcpyAsync(t64, tw.data(), 2 * K * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        g.tw64[d] = t64;
        if (mixed) {
            float* t32;
            HIPCHK(dalloc(ctx, &t32, 2 * K));
            HIPCHK(hipMemcpyAsync(t32, twf.data(), 2 * K * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            g.tw[d] = t32;
        } else {
            g.tw[d] = t64;
        }
        std::vector<double> bm = bspline_moduli(K);
        HIPCHK(dalloc(ctx, &g.bmod[d], K));
        HIPCHK(hipMemcpyAsync(g.bmod[d], bm.data(), K * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    // The fused plane pass keeps two K0 x K1 complex planes and both twiddle
    // tables in LDS.  It is used up to PLANE_LDS_MAX bytes (two blocks per CU;
    // up to 50^2 in FP64, 71^2 in FP32); larger planes take the line passes.
    const char* env = std::getenv("ADMP_FFT_PLANE");
    const bool allow = !(env && env[0] == '0') && K0 >= BRX + 5 && K1 >= BRY + 5 && K2 >= BRZ + 5;
    for (int w = 0; w < 2; ++w) {
        const size_t cs = w ? sizeof(Cplx<double>) : sizeof(Cplx<float>);
        g.plane[w] = allow && plane_smem(K0, K1, cs) <= PLANE_LDS_MAX && ctx->plane_lds_ok;
    }
    return ADMP_OK;
}

extern "C" const char* admp_version(void) { return "admp-hip 0.1 (gfx950)"; }

static admp_status nbr_settle(admp_ctx* ctx);

static const char* context_free_error();
extern "C" const char* admp_last_error(const admp_ctx* ctx) { return ctx ? ctx->err.c_str() : context_free_error(); }

// The active stream: the context's own one, or in caller mode the caller's
// (when it is not the null stream, on which graphs cannot be captured).  ev_out
// marks the end of the context's last work (recorded by every exit and at
// creation); in caller mode the own stream takes that end over at every exit
// (ev_done).  A change of the active stream orders the new one after the
// context's last work only, not after unrelated work queued later on the old
// stream, and always through an event of the own stream: a wait on an event of a
// caller's stream that another thread is capturing at that moment (a server's
// worker, capturing the graphs of another context on its stream) would join that
// capture or be refused.
static admp_status apply_stream(admp_ctx* ctx) {
    hipStream_t want = (ctx->caller_mode && ctx->user_stream) ? ctx->user_stream : ctx->own;
    if (want != ctx->stream) {
        HIPCHK(hipStreamWaitEvent(want, ctx->stream == ctx->own ? ctx->ev_out : ctx->ev_done, 0));
        ctx->stream = want;
    }
    return ADMP_OK;
}

extern "C" admp_status admp_set_stream(admp_ctx* ctx, void* stream) {
    if (!ctx) return ADMP_ERR_ARG;
    ctx->user_stream = reinterpret_cast<hipStream_t>(stream);
    return apply_stream(ctx);
}

extern "C" admp_status admp_set_stream_mode(admp_ctx* ctx, int32_t caller_mode) {
    if (!ctx) return ADMP_ERR_ARG;
    ctx->caller_mode = caller_mode != 0;
    return apply_stream(ctx);
}

// work on the context stream starts after everything queued on the caller's stream
static admp_status enter(admp_ctx* ctx) {
    if (ctx->stream == ctx->user_stream) return ADMP_OK;
    HIPCHK(hipEventRecord(ctx->ev_in, ctx->user_stream));
    HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->ev_in, 0));
    return ADMP_OK;
}

// and the caller's stream continues after the context's work
static admp_status leave(admp_ctx* ctx) {
    HIPCHK(hipEventRecord(ctx->ev_out, ctx->stream));
    if (ctx->stream != ctx->own) {
        // caller mode: the own stream takes the end of this work over
        HIPCHK(hipStreamWaitEvent(ctx->own, ctx->ev_out, 0));
        HIPCHK(hipEventRecord(ctx->ev_done, ctx->own));
    }
    if (ctx->stream == ctx->user_stream) return ADMP_OK;
    HIPCHK(hipStreamWaitEvent(ctx->user_stream, ctx->ev_out, 0));
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
    // back on the own stream, which in caller mode took the end of the context's
    // last work over at its exit (never a wait on the caller's stream, which may
    // carry other work or be captured by another thread by now)
    if (ctx->own && ctx->stream != ctx->own) ctx->stream = ctx->own;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->ev_pending) { ctx->ev_pool.push_back(p.second.first); ctx->ev_pool.push_back(p.second.second); }
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    clear_graphs(ctx);
    for (void* p : ctx->allocs) (void)hipFreeAsync(p, ctx->stream);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    for (int k = 0; k < 2; ++k)
        if (ctx->ev_state[k]) (void)hipEventDestroy(ctx->ev_state[k]);
    if (ctx->h_pcg) (void)hipHostFree(ctx->h_pcg);
    if (ctx->h_maxcnt) (void)hipHostFree(ctx->h_maxcnt);
    if (ctx->h_geom) (void)hipHostFree(ctx->h_geom);
    if (ctx->h_cgflag) (void)hipHostFree(ctx->h_cgflag);
    if (ctx->h_comp) (void)hipHostFree(ctx->h_comp);
    if (ctx->ev_nbr) (void)hipEventDestroy(ctx->ev_nbr);
    if (ctx->ev_geom) (void)hipEventDestroy(ctx->ev_geom);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->ev_in) (void)hipEventDestroy(ctx->ev_in);
    if (ctx->ev_out) (void)hipEventDestroy(ctx->ev_out);
    if (ctx->ev_done) (void)hipEventDestroy(ctx->ev_done);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);   // never the caller's stream
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
    {
        const char* ng = getenv("ADMP_NO_GRAPH");
        ctx->use_graphs = !(ng && ng[0] == '1');
    }
    ctx->nh = (c.lmax + 1) * (c.lmax + 1);
    ctx->use_frames = c.lmax > 0 ? 1 : 0;
    const size_t NF = (size_t)ctx->N * ctx->F;
    const bool mixed = c.precision == ADMP_MIXED;
    admp_status st;
    if (hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking) != hipSuccess ||
        (ctx->stream = ctx->own, false) ||
        hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_state[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_state[1], hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_pcg, 2 * sizeof(PcgState) * (size_t)std::max(1, c.n_frames), 0) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_maxcnt, sizeof(int), 0) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_geom, sizeof(int), 0) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_cgflag, sizeof(int) * (size_t)std::max(1, c.n_frames), 0) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_comp, 10 * sizeof(double) * (size_t)std::max(1, c.n_frames), 0) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_nbr, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_geom, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_out, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_done, hipEventDisableTiming) != hipSuccess) {
        ctx->err = "stream/event creation failed";
        return fail(ADMP_ERR_HIP);
    }
    ctx->plane_lds_ok = plane_kernels_ready();
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
    if (Gmax) {
        double* g;
        ALLOC(g, Gmax * ctx->F);
        ctx->d_rgrid = g;
        size_t Hmax = 0, Lmax = 0;
        for (const Grid* gg : {&ctx->eg, &ctx->dg})
            if (gg->G) {
                Hmax = std::max(Hmax, (size_t)gg->K[0] * gg->K[1] * (gg->K[2] / 2 + 1));
                Lmax = std::max(Lmax, (size_t)gg->K[1] * (gg->K[2] / 2 + 1));
            }
        ALLOC(g, Hmax * 2 * ctx->F);
        ctx->d_hgrid = g;
        ctx->part_blocks = std::max<int>(ctx->part_blocks, (int)Lmax);
        dfree(ctx, ctx->d_part);
        ALLOC(ctx->d_part, (size_t)ctx->part_blocks * ctx->F * 7);
        ALLOC(ctx->d_gccount, Gmax * ctx->F);
        ALLOC(ctx->d_gcfill, Gmax * ctx->F);
        ALLOC(ctx->d_gcstart, (Gmax + 1) * ctx->F);
        size_t nbrmax = 0;
        for (const Grid* gg : {&ctx->eg, &ctx->dg})
            if (gg->G)
                nbrmax = std::max(nbrmax, (size_t)((gg->K[0] + BRX - 1) / BRX) * ((gg->K[1] + BRY - 1) / BRY) *
                                              ((gg->K[2] + BRZ - 1) / BRZ));
        ALLOC(ctx->d_bcount, nbrmax * ctx->F);
        ALLOC(ctx->d_bstart, (nbrmax + 1) * ctx->F);
        if (ctx->eg.G)
            ALLOC(ctx->d_infl, (size_t)ctx->eg.K[0] * ctx->eg.K[1] * (ctx->eg.K[2] / 2 + 1) * ctx->F);
        ctx->blist_cap = (long)BRICK_MAXCAND * ctx->N;  // per frame
        ALLOC(ctx->d_blist, (size_t)ctx->blist_cap * ctx->F);
    }
    ALLOC(ctx->d_wd, NF * 3);
    ALLOC(ctx->d_gcell, NF);
    ALLOC(ctx->d_gcatoms, NF);
    ALLOC(ctx->d_e_atom, NF); ALLOC(ctx->d_gpos, NF * 3); ALLOC(ctx->d_gq, NF * 10); ALLOC(ctx->d_vir_atom, NF * 9);
    ALLOC(ctx->d_gpol, NF); ALLOC(ctx->d_gth, NF); ALLOC(ctx->d_e_self, NF); ALLOC(ctx->d_e_pol, NF);
    ALLOC(ctx->d_gq_rec, NF * 10); ALLOC(ctx->d_gpos_rec, NF * 3); ALLOC(ctx->d_vir_rec, NF * 9);
    ALLOC(ctx->d_contrib, NF * 12); ALLOC(ctx->d_vir_fr, NF * 9); ALLOC(ctx->d_gQl, NF * 9);
    ALLOC(ctx->d_red, ctx->F * 16);
    ALLOC(ctx->d_red9a, ctx->F * 9); ALLOC(ctx->d_red9b, ctx->F * 9); ALLOC(ctx->d_red9c, ctx->F * 9);
    ALLOC(ctx->d_b, NF * 3); ALLOC(ctx->d_r, NF * 6); ALLOC(ctx->d_z, NF * 3); ALLOC(ctx->d_p, NF * 3);
    ALLOC(ctx->d_Ap, NF * 3); ALLOC(ctx->d_t1, NF * 3); ALLOC(ctx->d_t2, NF * 3); ALLOC(ctx->d_U, NF * 3);
    ALLOC(ctx->d_dpart, (size_t)ctx->dots_blocks * ctx->F * 3);
    ALLOC(ctx->d_dots, ctx->F * 3);
    ctx->pcg_nb = (int)nblk(ctx->N, PCG_BLOCK);
    ctx->cg_nw = std::max(ctx->pcg_nb, gather_blocks(ctx->N));
    ALLOC(ctx->d_pparts, (size_t)ctx->F * (ctx->cg_nw + 4 * ctx->pcg_nb));
    ALLOC(ctx->d_s, NF * 6);
    ALLOC(ctx->d_pnear, NF); ALLOC(ctx->d_pcj, NF * 3); ALLOC(ctx->d_pcd, NF * 3 * PC_ROW);
    ALLOC(ctx->d_pcn, NF); ALLOC(ctx->d_pcc, NF * PC_MAXCH);
    ALLOC(ctx->d_swfrac, NF * 3);
    ALLOC(ctx->d_sfq, NF * 10);
    ALLOC(ctx->d_res, NF * 3);
    ALLOC(ctx->d_dU, NF * 3);
    ALLOC(ctx->d_pcg, ctx->F);
    ALLOC(ctx->d_pol, ctx->N);
    ALLOC(ctx->d_thole, ctx->N);
    ALLOC(ctx->d_Qloc, (size_t)ctx->N * ctx->nh);
    ALLOC(ctx->d_eout, ctx->F);
    ALLOC(ctx->d_bout, (size_t)ctx->F * 9);
    ALLOC(ctx->d_coef, ctx->F * 4);
    ALLOC(ctx->d_active, ctx->F);
    ALLOC(ctx->d_maxcnt, 1);
    ALLOC(ctx->d_geom, 1);
    ALLOC(ctx->d_cgflag, ctx->F);
    ALLOC(ctx->d_ncnt, NF);
    ALLOC(ctx->d_nhalf, NF);
    ALLOC(ctx->d_cell, NF);
    ALLOC(ctx->d_catoms, NF);
    ALLOC(ctx->d_spos, NF);
#undef ALLOC
    // the zero fills of creation are the context's first work
    if (hipEventRecord(ctx->ev_out, ctx->own) != hipSuccess || hipEventRecord(ctx->ev_done, ctx->own) != hipSuccess) {
        ctx->err = "event record failed";
        return fail(ADMP_ERR_HIP);
    }
    *out = ctx;
    return ADMP_OK;
}

extern "C" admp_status admp_set_topology(admp_ctx* ctx, const int32_t* axis_type, const int32_t* axis_idx,
                                         int64_t n_excl, const int32_t* excl_i, const int32_t* excl_j,
                                         const int32_t* excl_nbond) {
    if (!ctx) return ADMP_ERR_ARG;
    { const admp_status ns = nbr_settle(ctx); if (ns != ADMP_OK) return ns; }
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
    HIPCHK(hipStreamSynchronize(ctx->stream));  // no queued work may still read the old tables
    dfree(ctx, ctx->d_axis_type); dfree(ctx, ctx->d_axis_idx); dfree(ctx, ctx->d_xoff); dfree(ctx, ctx->d_xent);
    dfree(ctx, ctx->d_rev_off); dfree(ctx, ctx->d_rev_ent);
    HIPCHK(dalloc(ctx, &ctx->d_axis_type, N));
    HIPCHK(dalloc(ctx, &ctx->d_axis_idx, 3 * N));
    HIPCHK(dalloc(ctx, &ctx->d_xoff, N + 1));
    HIPCHK(dalloc(ctx, &ctx->d_xent, xent.size()));
    HIPCHK(dalloc(ctx, &ctx->d_rev_off, N + 1));
    HIPCHK(dalloc(ctx, &ctx->d_rev_ent, rent.size()));
    HIPCHK(hipMemcpyAsync(ctx->d_axis_type, at.data(), N * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_axis_idx, ai.data(), 3 * N * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_xoff, xoff.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    if (!xent.empty()) HIPCHK(hipMemcpyAsync(ctx->d_xent, xent.data(), xent.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_rev_off, roff.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    if (!rent.empty()) HIPCHK(hipMemcpyAsync(ctx->d_rev_ent, rent.data(), rent.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));  // host vectors are released on return
    ctx->has_topo = true;
    return ADMP_OK;
}

// exclusive scan of F rows of n ints into out (n + 1 per row) on stream st
static admp_status scan_rows(admp_ctx* ctx, const int* in, int* out, int n, int F, hipStream_t st) {
    if (n <= 4 * 4096) {
        hipLaunchKernelGGL(k_scan_small, dim3(F), dim3(1024), 0, st, n, in, out);
        HIPCHK(hipGetLastError());
        return ADMP_OK;
    }
    const int ntile = (n + SCAN_TILE - 1) / SCAN_TILE;
    const size_t need = (size_t)ntile * F;
    if (need > ctx->scan_cap) {
        dfree(ctx, ctx->d_scan_tmp);
        HIPCHK(dalloc(ctx, &ctx->d_scan_tmp, need));
        ctx->scan_cap = need;
    }
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntile, F), dim3(256), 0, st, n, in, out, ctx->d_scan_tmp, ntile);
    hipLaunchKernelGGL(k_scan_top, dim3(F), dim3(256), 0, st, n, ntile, ctx->d_scan_tmp, out);
    if (ntile > 1) hipLaunchKernelGGL(k_scan_add, dim3(ntile, F), dim3(256), 0, st, n, ctx->d_scan_tmp, out, ntile);
    HIPCHK(hipGetLastError());
    return ADMP_OK;
}

static admp_status ensure_cap(admp_ctx* ctx, int need) {
    if (need <= ctx->cap && ctx->d_nbr) return ADMP_OK;
    clear_graphs(ctx);  // captured launches hold the old buffers and capacity
    int cap = std::max(64, ((int)(need * 1.25) + 63) / 64 * 64);
    const size_t NF = (size_t)ctx->N * ctx->F;
    int* old = ctx->d_nbr;
    const int old_cap = ctx->cap;
    ctx->d_nbr = nullptr;
    HIPCHK(dalloc(ctx, &ctx->d_nbr, NF * cap));
    if (old) {
        // carry every row over to the new stride (lists of other frames stay valid)
        HIPCHK(hipMemcpy2DAsync(ctx->d_nbr, (size_t)cap * sizeof(int), old, (size_t)old_cap * sizeof(int),
                                (size_t)old_cap * sizeof(int), NF, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        dfree(ctx, old);
    }
    dfree(ctx, ctx->d_pcoef);
    ctx->d_pcoef = nullptr;
    double* pc;
    HIPCHK(dalloc(ctx, &pc, NF * cap * 2));
    ctx->d_pcoef = pc;
    ctx->cap = cap;
    return ADMP_OK;
}

// Positions and boxes of every frame into the context.  Non-finite or
// out-of-range positions and unusable boxes are replaced on the device (so no
// kernel derives an index from them) and flagged in d_geom; the entry point
// queues geom_fetch() before its final synchronisation and returns
// ADMP_ERR_ARG from geom_check() after it.
static admp_status upload_geometry(admp_ctx* ctx, const double* pos, const double* box) {
    const size_t NF = (size_t)ctx->N * ctx->F;
    ctx->prepped = nullptr;  // site grid bases / sort follow the positions
    HIPCHK(hipMemsetAsync(ctx->d_geom, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_pack_pos, dim3(nblk(NF, 256)), dim3(256), 0, ctx->stream, NF, pos, ctx->d_pos, ctx->d_geom);
    const admp_config& c = ctx->cfg;
    hipLaunchKernelGGL(k_box_setup, dim3(nblk(ctx->F, 64)), dim3(64), 0, ctx->stream, ctx->F, box, c.lpme ? c.K1 : 0,
                       c.lpme ? c.K2 : 0, c.lpme ? c.K3 : 0, c.dK1, c.dK2, c.dK3, ctx->d_box, ctx->d_geom);
    HIPCHK(hipGetLastError());
    return ADMP_OK;
}

static admp_status geom_fetch(admp_ctx* ctx) {
    HIPCHK(hipMemcpyAsync(ctx->h_geom, ctx->d_geom, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    return ADMP_OK;
}

// after the stream has passed geom_fetch()
static admp_status geom_check(admp_ctx* ctx) {
    const int g = *ctx->h_geom;
    if (!g) return ADMP_OK;
    ctx->err = (g & GEOM_BAD_BOX) ? "box not finite or degenerate (each height must be >= 0.5 A)"
                                  : "positions not finite or beyond 1e5 A";
    return ADMP_ERR_ARG;
}

static admp_status admp_set_pairs_impl(admp_ctx* ctx, int32_t frame, const void* pairs, int64_t n_pairs,
                                      int32_t pairs_is_int64) {
    if (!ctx || frame < 0 || frame >= ctx->F || (n_pairs > 0 && !pairs)) return ADMP_ERR_ARG;
    { const admp_status ns = nbr_settle(ctx); if (ns != ADMP_OK) return ns; }
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
    if ((int)ctx->frame_nbr.size() != ctx->F) ctx->frame_nbr.assign(ctx->F, 0);
    ctx->frame_nbr[frame] = 1;
    ctx->nbr_built = false;
    return ADMP_OK;
}

// The list build writes rows up to the capacity and the longest row it found;
// that count reaches the host through a pinned copy behind an event, so the
// caller's next work is queued without waiting for the build.
static admp_status launch_nbr_build(admp_ctx* ctx) {
    const int N = ctx->N, F = ctx->F;
    const CellGrid cg = ctx->nbr_cg;
    const double rc = ctx->cfg.rc;
    HIPCHK(hipMemsetAsync(ctx->d_maxcnt, 0, sizeof(int), ctx->stream));
    { ProfScope ps_(ctx, 8);
    hipLaunchKernelGGL(k_nbr_build, dim3(nblk(N, 4), F), dim3(256), 0, ctx->stream, N, ctx->cap, ctx->d_pos, ctx->d_box,
                       cg, rc * rc, ctx->d_cell, ctx->d_cstart, ctx->d_spos, ctx->d_xoff, ctx->d_xent, ctx->d_nbr,
                       ctx->d_ncnt, ctx->d_nhalf, ctx->d_maxcnt); }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ctx->h_maxcnt, ctx->d_maxcnt, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev_nbr, ctx->stream));
    ctx->nbr_pending = true;
    return ADMP_OK;
}

// Resolve a pending list build before anything reads the lists, the geometry or
// the cell tables it used: when a row overflowed the capacity, grow it and build
// again from the same geometry (this one waits).  Every entry point that reads
// or replaces them calls this first.
static CellGrid host_cell_grid(const admp_ctx* ctx, const std::vector<double>& hb);
static admp_status launch_cell_sort(admp_ctx* ctx, const CellGrid& cg);

static admp_status nbr_settle(admp_ctx* ctx) {
    if (!ctx->nbr_pending) return ADMP_OK;
    ctx->nbr_pending = false;
    HIPCHK(hipEventSynchronize(ctx->ev_nbr));
    { const admp_status g = geom_check(ctx); if (g != ADMP_OK) return g; }
    if (ctx->cg_spec) {
        ctx->cg_spec = false;
        bool regrid = false;
        for (int f = 0; f < ctx->F; ++f) regrid = regrid || ctx->h_cgflag[f] != 0;
        if (regrid) {
            // a box outgrew the cell grid taken over from the earlier build: size
            // it from the uploaded boxes and build again (this one waits)
            std::vector<BoxInfo> bi(ctx->F);
            HIPCHK(hipMemcpyAsync(bi.data(), ctx->d_box, ctx->F * sizeof(BoxInfo), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            std::vector<double> hb(9 * (size_t)ctx->F);
            for (int f = 0; f < ctx->F; ++f)
                for (int k = 0; k < 9; ++k) hb[9 * (size_t)f + k] = bi[f].B[k];
            admp_status s = launch_cell_sort(ctx, host_cell_grid(ctx, hb));
            if (s != ADMP_OK) return s;
            if ((s = launch_nbr_build(ctx)) != ADMP_OK) return s;
            ctx->nbr_pending = false;
            HIPCHK(hipEventSynchronize(ctx->ev_nbr));
        }
    }
    const int mx = *ctx->h_maxcnt;
    if (mx <= ctx->cap) return ADMP_OK;
    admp_status s = ensure_cap(ctx, mx);
    if (s != ADMP_OK) return s;
    if ((s = launch_nbr_build(ctx)) != ADMP_OK) return s;
    ctx->nbr_pending = false;
    HIPCHK(hipEventSynchronize(ctx->ev_nbr));
    if (*ctx->h_maxcnt > ctx->cap) { ctx->err = "neighbour list capacity"; return ADMP_ERR_HIP; }
    return ADMP_OK;
}

// cell grid for the frames' boxes hb (9 per frame, host): the most cells per
// axis whose slabs stay at least rc wide in every frame
static CellGrid host_cell_grid(const admp_ctx* ctx, const std::vector<double>& hb) {
    const double rc = ctx->cfg.rc;
    CellGrid cg;
    cg.nc[0] = cg.nc[1] = cg.nc[2] = 1 << 20;
    for (int f = 0; f < ctx->F; ++f) {
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
    return cg;
}

// cell sort of the uploaded positions on grid cg (stored as nbr_cg)
static admp_status launch_cell_sort(admp_ctx* ctx, const CellGrid& cg) {
    const int N = ctx->N, F = ctx->F;
    if (cg.ncell > ctx->ncell_alloc) {
        // the tables may still be read by queued work of an earlier build
        HIPCHK(hipStreamSynchronize(ctx->stream));
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
    admp_status s = scan_rows(ctx, ctx->d_ccount, ctx->d_cstart, cg.ncell, F, ctx->stream);
    if (s != ADMP_OK) return s;
    hipLaunchKernelGGL(k_cell_fill, dim3(nblk(N, 256), F), dim3(256), 0, ctx->stream, N, cg.ncell, ctx->d_cell,
                       ctx->d_cstart, ctx->d_cfill, ctx->d_catoms, 1);
    hipLaunchKernelGGL(k_cell_sort_block, dim3(cg.ncell, F), dim3(256), 0, ctx->stream, cg.ncell, N, ctx->d_cstart,
                       ctx->d_catoms);
    hipLaunchKernelGGL(k_cell_pos, dim3(nblk(N, 256), F), dim3(256), 0, ctx->stream, N, ctx->d_pos, ctx->d_catoms,
                       ctx->d_spos);
    HIPCHK(hipGetLastError());
    ctx->nbr_cg = cg;
    return ADMP_OK;
}

// The cell grid depends on the boxes, which live on the device.  The first
// build (and one after a grid change) reads them back and waits; a later build
// takes the previous grid over without waiting, and k_cell_grid_check flags a
// box that needs another one: nbr_settle then sizes the grid from the boxes and
// builds again before anything reads the lists.
static admp_status admp_build_neighbors_impl(admp_ctx* ctx, const double* pos, const double* box) {
    if (!ctx || !pos || !box) return ADMP_ERR_ARG;
    if (!ctx->has_topo) { ctx->err = "admp_set_topology first"; return ADMP_ERR_STATE; }
    const int N = ctx->N, F = ctx->F;
    admp_status s = nbr_settle(ctx);
    if (s != ADMP_OK) return s;
    s = upload_geometry(ctx, pos, box);
    if (s != ADMP_OK) return s;
    CellGrid cg;
    std::vector<double> hb;
    const bool spec = ctx->cg_host && ctx->d_nbr;
    if (spec) {
        cg = ctx->nbr_cg;
        hipLaunchKernelGGL(k_cell_grid_check, dim3(nblk(F, 64)), dim3(64), 0, ctx->stream, F, ctx->d_box,
                           ctx->cfg.rc, cg, ctx->d_cgflag);
        HIPCHK(hipMemcpyAsync(ctx->h_cgflag, ctx->d_cgflag, F * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    } else {
        hb.resize(9 * (size_t)F);
        HIPCHK(hipMemcpyAsync(hb.data(), box, 9 * F * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        for (int f = 0; f < F; ++f)
            if (!box_usable(hb.data() + 9 * f)) {
                ctx->err = "box not finite or degenerate (each height must be >= 0.5 A)";
                return ADMP_ERR_ARG;
            }
        cg = host_cell_grid(ctx, hb);
    }
    if ((s = launch_cell_sort(ctx, cg)) != ADMP_OK) return s;
    if (!ctx->d_nbr) {
        // first estimate from density: 1.5 x expected full-list count + margin
        const double rc = ctx->cfg.rc;
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
    ctx->cg_host = true;
    ctx->cg_spec = spec;
    if ((s = geom_fetch(ctx)) != ADMP_OK) return s;  // read back with the build's longest row
    if ((s = launch_nbr_build(ctx)) != ADMP_OK) return s;
    ctx->frame_nbr.assign(ctx->F, 1);
    ctx->nbr_built = true;
    return ADMP_OK;
}

static admp_status admp_get_pairs_impl(admp_ctx* ctx, int32_t frame, int32_t* out, int64_t cap, int64_t* n_out) {
    if (!ctx || frame < 0 || frame >= ctx->F || !n_out) return ADMP_ERR_ARG;
    if (!ctx->nbr_built) { ctx->err = "no neighbour list built by admp_build_neighbors"; return ADMP_ERR_STATE; }
    { const admp_status ns = nbr_settle(ctx); if (ns != ADMP_OK) return ns; }
    const int N = ctx->N;
    // nhalf holds the j > i counts for lists built on the GPU
    int* off;
    HIPCHK(dalloc(ctx, &off, N + 1));
    {
        admp_status s2 = scan_rows(ctx, ctx->d_nhalf + (size_t)frame * N, off, N, 1, ctx->stream);
        if (s2 != ADMP_OK) { dfree(ctx, off); return s2; }
    }
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
// Reciprocal space.  recip_positions() computes every site's grid base and the
// counting sort of sites by base grid point; it depends only on positions and
// the box, so it runs once per evaluation and serves every spread of the solve.
// ---------------------------------------------------------------------------
// The brick spread keeps one modular representative of each site's support start
// relative to the brick, which covers every point of the brick only when each
// axis has at least brick + 5 points; smaller grids take the scatter path.
static bool small_grid(const Grid& g) { return g.K[0] < BRX + 5 || g.K[1] < BRY + 5 || g.K[2] < BRZ + 5; }

static admp_status recip_positions(admp_ctx* ctx, const Grid& g, int use_dA) {
    const int N = ctx->N, F = ctx->F;
    hipStream_t st = ctx->stream;
    { ProfScope ps_(ctx, 3);
    // the sort's counters are zeroed by k_grid_base (no separate fills)
    const bool sort = !small_grid(g);
    hipLaunchKernelGGL(k_grid_base, dim3(nblk(N, 128), F), dim3(128), 0, st, N, ctx->d_pos, ctx->d_box, use_dA,
                       ctx->d_mbase, ctx->d_wd, sort ? ctx->d_gccount : nullptr, sort ? ctx->d_gcfill : nullptr,
                       sort ? g.G * F : (size_t)0);
    if (sort) {
        const size_t G = g.G;
        hipLaunchKernelGGL(k_grid_cell, dim3(nblk(N, 256), F), dim3(256), 0, st, N, g.K[0], g.K[1], g.K[2],
                           ctx->d_mbase, ctx->d_gcell, ctx->d_gccount);
        admp_status s2 = scan_rows(ctx, ctx->d_gccount, ctx->d_gcstart, (int)G, F, st);
        if (s2 != ADMP_OK) return s2;
        hipLaunchKernelGGL(k_cell_fill, dim3(nblk(N, 256), F), dim3(256), 0, st, N, (int)G, ctx->d_gcell,
                           ctx->d_gcstart, ctx->d_gcfill, ctx->d_gcatoms, 0);
        hipLaunchKernelGGL(k_cell_sort, dim3(nblk(G, 256), F), dim3(256), 0, st, (int)G, N, ctx->d_gcstart,
                           ctx->d_gcatoms);
        // per-brick candidate lists for the spreads of this evaluation
        const int nbr = ((g.K[0] + BRX - 1) / BRX) * ((g.K[1] + BRY - 1) / BRY) * ((g.K[2] + BRZ - 1) / BRZ);
        hipLaunchKernelGGL(k_brick_cands<false>, dim3(nbr, F), dim3(64), 0, st, N, g.K[0], g.K[1], g.K[2],
                           ctx->d_mbase, ctx->d_gcstart, ctx->d_gcatoms, ctx->d_bcount, (const int*)nullptr,
                           (int2*)nullptr, 0L);
        s2 = scan_rows(ctx, ctx->d_bcount, ctx->d_bstart, nbr, F, st);
        if (s2 != ADMP_OK) return s2;
        hipLaunchKernelGGL(k_brick_cands<true>, dim3(nbr, F), dim3(64), 0, st, N, g.K[0], g.K[1], g.K[2],
                           ctx->d_mbase, ctx->d_gcstart, ctx->d_gcatoms, ctx->d_bcount, ctx->d_bstart, ctx->d_blist,
                           ctx->blist_cap);
        if (&g == &ctx->eg && ctx->d_infl && !use_dA) {
            const size_t nt = (size_t)g.K[0] * g.K[1] * (g.K[2] / 2 + 1);
            // laid out for the FP64 solver products' FFT path
            ctx->infl_pm = g.plane[1] ? 1 : 0;
            hipLaunchKernelGGL(k_infl_table, dim3(nblk(nt, 256), F), dim3(256), 0, st, g.K[0], g.K[1], g.K[2],
                               ctx->d_box, g.bmod[0], g.bmod[1], g.bmod[2], ctx->cfg.kappa, DIELECTRIC, ctx->infl_pm,
                               ctx->d_infl);
        }
    }
    }
    HIPCHK(hipGetLastError());
    ctx->prepped = &g;
    return ADMP_OK;
}

template <class R> static const Cplx<R>* twid(const Grid& g, int d) {
    return reinterpret_cast<const Cplx<R>*>(sizeof(R) == 8 ? g.tw64[d] : g.tw[d]);
}

// The real-space product v -> out (FP64 stored coefficients) launched with the
// plane pass of a solver product (k_plane_matvec)
struct MatvecJob {
    const double* v;
    double* out;
};

// Spread the sources (Qc, U, qmode), transform, apply the influence function,
// transform back and gather.  gather_mode 1: dE/dmu only into out3; 0: full
// gradients into gq, gpos, vir (dE/dA per site).  Energy/box sums land in d_rec.
template <class R>
static admp_status recip_run(admp_ctx* ctx, const Grid& g, int use_dA, const double* Qc, const double* U, int qmode,
                             int lmax_eff, int kind, double pref, double kappa, int gather_mode, double* out3,
                             double* gq, double* gpos, double* vir, bool want_energy, void* wf_in = nullptr,
                             void* fq_in = nullptr, const PcgState* skip = nullptr, const CgEpi* epi = nullptr,
                             hipEvent_t before_gather = nullptr, const MatvecJob* mv = nullptr) {
    const int N = ctx->N, F = ctx->F;
    hipStream_t st = ctx->stream;
    if (ctx->prepped != &g) {
        admp_status s = recip_positions(ctx, g, use_dA);
        if (s != ADMP_OK) return s;
    }
    R* wfrac = reinterpret_cast<R*>(fq_in ? wf_in : ctx->d_wfrac);
    R* fq = reinterpret_cast<R*>(fq_in ? fq_in : ctx->d_fq);
    if (!fq_in) {
        ProfScope ps_(ctx, 3);
        hipLaunchKernelGGL(k_frac_mult<R>, dim3(nblk(N, 128), F), dim3(128), 0, st, N, ctx->d_box, use_dA, Qc, U, qmode,
                           ctx->d_wd, wfrac, fq);
    }
    const int K0 = g.K[0], K1 = g.K[1], K2 = g.K[2];
    if (small_grid(g) && gather_mode == 2) {
        ctx->err = "fused CG product epilogue needs the brick spread path";
        return ADMP_ERR_ARG;
    }
    if (small_grid(g)) {
        // grids below 6 points per axis: scatter with atomics on the complex C2C path
        Cplx<R>* grid = reinterpret_cast<Cplx<R>*>(ctx->d_grid);
        const size_t tot = g.G * F;
        hipLaunchKernelGGL(k_zero_cplx<R>, dim3(nblk(tot, 256)), dim3(256), 0, st, tot, grid);
        { ProfScope ps_(ctx, 4);
        hipLaunchKernelGGL(k_spread_atomic<R>, dim3(nblk((size_t)N * 36, 256), F), dim3(256), 0, st, N, K0, K1, K2,
                           ctx->d_mbase, wfrac, fq, lmax_eff, grid); }
        const int maxpts = sizeof(R) == 8 ? 2048 : 4096;
        auto fft = [&](int inverse) {
            for (int dd = 0; dd < 3; ++dd) {
                const int d = inverse ? dd : 2 - dd;
                const int K = g.K[d];
                const int nlines = (int)(g.G / K);
                int LB = std::max(1, std::min(maxpts / K, nlines));
                const size_t sm = (size_t)2 * LB * K * sizeof(Cplx<R>);
                { ProfScope ps_(ctx, 5);
                hipLaunchKernelGGL(k_fft_pass<R>, dim3(nblk(nlines, LB), F), dim3(256), sm, st, grid, K0, K1, K2, d,
                                   g.plan[d], twid<R>(g, d), inverse, LB); }
            }
        };
        fft(0);
        const int nbI = (int)nblk(g.G, 256);
        { ProfScope ps_(ctx, 6);
        hipLaunchKernelGGL(k_influence<R>, dim3(nbI, F), dim3(256), 0, st, K0, K1, K2, ctx->d_box, use_dA, g.bmod[0],
                           g.bmod[1], g.bmod[2], kappa, kind, pref, grid, ctx->d_part); }
        if (want_energy)
            hipLaunchKernelGGL(k_reduce_cols, dim3(7, F), dim3(RED_BLOCK), 0, st, nbI, 7, ctx->d_part, ctx->d_rec, 0);
        fft(1);
        { ProfScope ps_(ctx, 7);
        hipLaunchKernelGGL(k_gather<R>, dim3(nblk(N, 64), F), dim3(64), 0, st, N, K0, K1, K2, ctx->d_pos, ctx->d_box,
                           use_dA, ctx->d_mbase, wfrac, fq, Qc, U, qmode, grid, gather_mode, out3, gq, gpos, vir); }
        HIPCHK(hipGetLastError());
        return ADMP_OK;
    }
    R* rg = reinterpret_cast<R*>(ctx->d_rgrid);
    Cplx<R>* hg = reinterpret_cast<Cplx<R>*>(ctx->d_hgrid);
    const int H = K2 / 2 + 1;
    const int nbricks = ((K0 + BRX - 1) / BRX) * ((K1 + BRY - 1) / BRY) * ((K2 + BRZ - 1) / BRZ);
    { ProfScope ps_(ctx, 4);
    // ADMP_SPREAD=list|col|cw forces one variant (measurements)
    static const int spread_force = [] {
        const char* e = std::getenv("ADMP_SPREAD");
        if (!e) return 0;
        if (std::strcmp(e, "list") == 0) return 1;
        if (std::strcmp(e, "col") == 0) return 2;
        if (std::strcmp(e, "cw") == 0) return 3;
        return 0;
    }();
    // default: one brick per wavefront with column accumulation (k_spread_cw); at
    // 12,288 sites it takes 19 us per solver spread where one grid point per lane
    // (list) takes 27 and four bricks per wavefront (col) 45, and it is also the
    // fastest for the 64-frame batch
    const int variant = spread_force ? spread_force : 3;
    const unsigned nb1 = xcd_blocks(nbricks), nb4 = xcd_blocks((nbricks + 3) / 4);
#define ADMP_SPREAD(LMv)                                                                                              \
    if (variant == 2)                                                                                                \
        hipLaunchKernelGGL((k_spread_col<R, LMv>), dim3(nb4, F), dim3(64), 0, st, N, K0, K1, K2, wfrac, fq,           \
                           ctx->d_bstart, ctx->d_blist, ctx->blist_cap, rg, skip);                                   \
    else if (variant == 3 && F == 1 && (LMv) <= 1)                                                                  \
        hipLaunchKernelGGL((k_spread_cw<R, LMv, SP_CHUNK_ONE_FRAME>), dim3(nb1, F), dim3(64), 0, st, N, K0, K1, K2,    \
                           wfrac, fq, ctx->d_bstart, ctx->d_blist, ctx->blist_cap, rg, skip, ctx->spread_chk);       \
    else if (variant == 3)                                                                                           \
        hipLaunchKernelGGL((k_spread_cw<R, LMv>), dim3(nb1, F), dim3(64), 0, st, N, K0, K1, K2, wfrac, fq,            \
                           ctx->d_bstart, ctx->d_blist, ctx->blist_cap, rg, skip, ctx->spread_chk);                  \
    else                                                                                                             \
        hipLaunchKernelGGL((k_spread_list<R, LMv>), dim3(nb1, F), dim3(64), 0, st, N, K0, K1, K2, wfrac, fq,          \
                           ctx->d_bstart, ctx->d_blist, ctx->blist_cap, rg, skip);
    if (lmax_eff >= 2) { ADMP_SPREAD(2) }
    else if (lmax_eff == 1) { ADMP_SPREAD(1) }
    else { ADMP_SPREAD(0) }
#undef ADMP_SPREAD
    }
    const int npairs = (K0 * K1 + 1) / 2;
    const int nly = K0 * H, nlx = K1 * H;
    const size_t Cs = sizeof(Cplx<R>);
    const int pm = g.plane[sizeof(R) == 8 ? 1 : 0] ? 1 : 0;
    if (mv && (!pm || want_energy)) {
        ctx->err = "real-space product with the plane pass needs a solver product on the plane path";
        return ADMP_ERR_STATE;
    }
    // z passes: 8 complex lines (16 real rows) per block where LDS allows, else 4
    const bool z8 = (size_t)(2 * 8 + 1) * K2 * Cs <= 48 * 1024;
    auto zpass = [&](int inverse) {
        ProfScope ps_(ctx, 5);
        if (!inverse) {
            if (z8)
                hipLaunchKernelGGL((k_fz_r2c<R, 8>), dim3(nblk(npairs, 8), F), dim3(128), (2 * 8 + 1) * K2 * Cs, st, K0,
                                   K1, K2, g.plan[2], twid<R>(g, 2), rg, hg, pm, skip);
            else
                hipLaunchKernelGGL((k_fz_r2c<R, 4>), dim3(nblk(npairs, 4), F), dim3(64), (2 * 4 + 1) * K2 * Cs, st, K0,
                                   K1, K2, g.plan[2], twid<R>(g, 2), rg, hg, pm, skip);
        } else {
            if (z8)
                hipLaunchKernelGGL((k_fz_c2r<R, 8>), dim3(nblk(npairs, 8), F), dim3(128), (2 * 8 + 1) * K2 * Cs, st, K0,
                                   K1, K2, g.plan[2], twid<R>(g, 2), hg, rg, pm, skip);
            else
                hipLaunchKernelGGL((k_fz_c2r<R, 4>), dim3(nblk(npairs, 4), F), dim3(64), (2 * 4 + 1) * K2 * Cs, st, K0,
                                   K1, K2, g.plan[2], twid<R>(g, 2), hg, rg, pm, skip);
        }
    };
    const bool use_tab = &g == &ctx->eg && kind == 0 && !use_dA && pref == DIELECTRIC && kappa == ctx->cfg.kappa &&
                         ctx->d_infl && pm == ctx->infl_pm;
    const double* tab = use_tab ? ctx->d_infl : nullptr;
    zpass(0);
    int npart;
    if (pm) {
        // fused y / x / influence / inverse x / inverse y, one block per (kz, frame)
        npart = H;
        const size_t sm = plane_smem(K0, K1, Cs);
        ProfScope ps_(ctx, mv ? 13 : 6);
        if (want_energy)
            hipLaunchKernelGGL((k_fxy_plane<R, true>), dim3(H, F), dim3(512), sm, st, K0, K1, K2, g.plan[0], g.plan[1],
                               twid<R>(g, 0), twid<R>(g, 1), hg, ctx->d_box, g.bmod[0], g.bmod[1], g.bmod[2], kappa,
                               kind, pref, ctx->d_part, (const double*)nullptr, skip);
        else if (mv) {
            const int npb = xcd_blocks(H * F);
            const int nbf = (int)nblk((size_t)N * TPA, PLANE_THREADS);
            hipLaunchKernelGGL((k_plane_matvec<R>), dim3(npb + xcd_blocks(nbf * F)), dim3(PLANE_THREADS), sm, st, K0, K1,
                               K2, g.plan[0], g.plan[1], twid<R>(g, 0), twid<R>(g, 1), hg, ctx->d_box, g.bmod[0], g.bmod[1],
                               g.bmod[2], kappa, kind, pref, tab, F, npb, N, nbf, ctx->cap, ctx->d_pos, ctx->d_nbr,
                               ctx->d_ncnt, reinterpret_cast<const double*>(ctx->d_pcoef), mv->v, mv->out, skip);
        } else
            hipLaunchKernelGGL((k_fxy_plane<R, false>), dim3(H, F), dim3(PLANE_THREADS), sm, st, K0, K1, K2, g.plan[0], g.plan[1],
                               twid<R>(g, 0), twid<R>(g, 1), hg, ctx->d_box, g.bmod[0], g.bmod[1], g.bmod[2], kappa,
                               kind, pref, ctx->d_part, tab, skip);
    } else {
        { ProfScope ps_(ctx, 5);
        hipLaunchKernelGGL(k_fy<R>, dim3(nblk(nly, FL), F), dim3(FTH), (2 * FL + 1) * K1 * Cs, st, K0, K1, H, g.plan[1],
                           twid<R>(g, 1), hg, 0, skip); }
        npart = (int)nblk(nlx, FL);
        { ProfScope ps_(ctx, 6);
        if (want_energy)
            hipLaunchKernelGGL((k_fx_infl<R, true>), dim3(npart, F), dim3(FTH), (2 * FL + 1) * K0 * Cs, st, K0, K1, K2,
                               g.plan[0], twid<R>(g, 0), hg, ctx->d_box, g.bmod[0], g.bmod[1], g.bmod[2], kappa, kind,
                               pref, ctx->d_part, (const double*)nullptr, skip);
        else
            hipLaunchKernelGGL((k_fx_infl<R, false>), dim3(npart, F), dim3(FTH), (2 * FL + 1) * K0 * Cs, st, K0, K1, K2,
                               g.plan[0], twid<R>(g, 0), hg, ctx->d_box, g.bmod[0], g.bmod[1], g.bmod[2], kappa, kind,
                               pref, ctx->d_part, tab, skip); }
        { ProfScope ps_(ctx, 5);
        hipLaunchKernelGGL(k_fy<R>, dim3(nblk(nly, FL), F), dim3(FTH), (2 * FL + 1) * K1 * Cs, st, K0, K1, H, g.plan[1],
                           twid<R>(g, 1), hg, 1, skip); }
    }
    if (want_energy)
        hipLaunchKernelGGL(k_reduce_cols, dim3(7, F), dim3(RED_BLOCK), 0, st, npart, 7, ctx->d_part, ctx->d_rec, 0);
    zpass(1);
    if (before_gather) HIPCHK(hipStreamWaitEvent(st, before_gather, 0));
    { ProfScope ps_(ctx, 7);
    const CgEpi none{};
    if (gather_mode == 2)
        hipLaunchKernelGGL((k_gather_ml<R, 2>), dim3(xcd_blocks(gather_blocks(N)), F), dim3(8 * GATHER_SITES_PER_BLOCK), 0, st, N, K0, K1, K2,
                           ctx->d_pos, ctx->d_box, use_dA, ctx->d_mbase, wfrac, fq, Qc, U, qmode, rg, out3, gq, gpos,
                           vir, ctx->d_gcatoms, skip, *epi);
    else if (gather_mode == 1)
        hipLaunchKernelGGL((k_gather_ml<R, 1>), dim3(xcd_blocks(gather_blocks(N)), F), dim3(8 * GATHER_SITES_PER_BLOCK), 0, st, N, K0, K1, K2,
                           ctx->d_pos, ctx->d_box, use_dA, ctx->d_mbase, wfrac, fq, Qc, U, qmode, rg, out3, gq, gpos,
                           vir, ctx->d_gcatoms, skip, none);
    else
        hipLaunchKernelGGL((k_gather_ml<R, 0>), dim3(xcd_blocks(gather_blocks(N)), F), dim3(8 * GATHER_SITES_PER_BLOCK), 0, st, N, K0, K1, K2,
                           ctx->d_pos, ctx->d_box, use_dA, ctx->d_mbase, wfrac, fq, Qc, U, qmode, rg, out3, gq, gpos,
                           vir, ctx->d_gcatoms, skip, none); }
    HIPCHK(hipGetLastError());
    return ADMP_OK;
}

// A v for the induced-dipole system (everything but the right-hand side).
// RC: type of the stored pair coefficients; R: arithmetic of the real-space and
// reciprocal products (the mixed solver's FP64 residual uses RC=float, R=double).
template <class RC, class R>
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
    { ProfScope ps_(ctx, 2);
    const int nbf = (int)nblk((size_t)N * TPA, PAIR_BLOCK);
    hipLaunchKernelGGL((k_real_matvec<RC, R>), dim3(xcd_blocks(nbf * F)), dim3(PAIR_BLOCK), 0, st, N, nbf, F,
                       ctx->cap, ctx->d_pos, ctx->d_box, ctx->d_nbr, ctx->d_ncnt,
                       reinterpret_cast<const RC*>(ctx->d_pcoef), vR, ctx->d_t1, (const PcgState*)nullptr); }
    double* rec = nullptr;
    if (ctx->cfg.lpme) {
        admp_status s = recip_run<R>(ctx, ctx->eg, 0, nullptr, v, 2, 1, 0, DIELECTRIC, ctx->cfg.kappa, 1, ctx->d_t2,
                                     nullptr, nullptr, nullptr, false);
        if (s != ADMP_OK) return s;
        rec = ctx->d_t2;
    }
    { ProfScope ps_(ctx, 11);
    hipLaunchKernelGGL(k_apply_diag, dim3(nblk(N, 256), F), dim3(256), 0, st, N, f1, pol, v, ctx->d_t1, rec, out); }
    HIPCHK(hipGetLastError());
    return ADMP_OK;
}

// per-frame (x.y, x.z, max |x| over polarisable sites) into ctx->d_dots (device only)
static void dots_dev(admp_ctx* ctx, const double* x, const double* y, const double* z, const double* pol) {
    const int F = ctx->F;
    ProfScope ps_(ctx, 11);
    hipLaunchKernelGGL(k_dots, dim3(ctx->dots_blocks, F), dim3(256), 0, ctx->stream, ctx->N, x, y, z, pol, ctx->d_dpart);
    hipLaunchKernelGGL(k_dots_final, dim3(nblk(F, 64)), dim3(64), 0, ctx->stream, F, ctx->dots_blocks, ctx->d_dpart,
                       ctx->d_dots);
}

static admp_status dots(admp_ctx* ctx, const double* x, const double* y, const double* z, const double* pol,
                        std::vector<double>& h) {
    const int F = ctx->F;
    dots_dev(ctx, x, y, z, pol);
    h.resize(3 * F);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->d_dots, 3 * F * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ADMP_OK;
}

struct SolveInfo {
    std::vector<int> n_iter, conv;
    std::vector<double> res;
};

static CgParts cg_parts(const admp_ctx* ctx) { return CgParts{ctx->d_pparts, ctx->pcg_nb, ctx->cg_nw}; }

// One product w = A z of the CG iteration.  z (and its R copy, fractional
// dipoles and spline coordinates) were written by k_cg_start / k_cg_update.
// w and its z.w partials come from the gather's CG epilogue (brick path, the
// real-space product joined before the gather) or from k_cg_w.
template <class RC, class R>
static admp_status pcg_apply(admp_ctx* ctx, const double* pol, double f1, int cur, double tol) {
    const int N = ctx->N, F = ctx->F;
    hipStream_t st = ctx->stream;
    // real-space products in FP64 whenever the stored coefficients are FP64 (the
    // near-field coupling sets the conditioning); R only for the reciprocal chain
    using RR = typename std::conditional<sizeof(RC) == 8, double, R>::type;
    const RR* vR = sizeof(RR) == 8 ? reinterpret_cast<const RR*>(ctx->d_z) : reinterpret_cast<const RR*>(ctx->d_vR);
    const bool fused = ctx->cfg.lpme && !small_grid(ctx->eg);
    // With few frames the plane pass fills only K2/2+1 workgroups per frame: the
    // real-space product then runs inside that launch (k_plane_matvec) instead of
    // on the side stream, which saves the fork and the join of every iteration.
    const int H = ctx->eg.K[2] / 2 + 1;
    const bool with_plane = fused && sizeof(RC) == 8 && ctx->eg.plane[sizeof(R) == 8 ? 1 : 0] && H * F <= 128;
    if (!with_plane) {
        // the real-space product runs on the side stream, beside the reciprocal chain
        HIPCHK(hipEventRecord(ctx->ev_fork, st));
        HIPCHK(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
        { ProfScope ps_(ctx, 2, ctx->side);
        const int nbf = (int)nblk((size_t)N * TPA, PAIR_BLOCK);
        hipLaunchKernelGGL((k_real_matvec<RC, RR>), dim3(xcd_blocks(nbf * F)), dim3(PAIR_BLOCK), 0, ctx->side, N, nbf,
                           F, ctx->cap, ctx->d_pos, ctx->d_box, ctx->d_nbr, ctx->d_ncnt,
                           reinterpret_cast<const RC*>(ctx->d_pcoef), vR, ctx->d_t1, ctx->d_pcg); }
        HIPCHK(hipEventRecord(ctx->ev_join, ctx->side));
    }
    // the product's first kernel tests the residual the next update would test
    ctx->spread_chk = CgCheck{cg_parts(ctx), ctx->d_pcg, tol, cur, 1};
    admp_status s = ADMP_OK;
    if (with_plane) {
        const CgEpi epi{ctx->d_t1, pol, f1, cg_parts(ctx)};
        const MatvecJob mv{ctx->d_z, ctx->d_t1};
        s = recip_run<R>(ctx, ctx->eg, 0, nullptr, ctx->d_z, 2, 1, 0, DIELECTRIC, ctx->cfg.kappa, 2, ctx->d_Ap,
                         nullptr, nullptr, nullptr, false, ctx->d_swfrac, ctx->d_sfq, ctx->d_pcg, &epi, nullptr, &mv);
        ctx->cg_nwu = gather_blocks(N);
    } else if (fused) {
        const CgEpi epi{ctx->d_t1, pol, f1, cg_parts(ctx)};
        s = recip_run<R>(ctx, ctx->eg, 0, nullptr, ctx->d_z, 2, 1, 0, DIELECTRIC, ctx->cfg.kappa, 2, ctx->d_Ap,
                         nullptr, nullptr, nullptr, false, ctx->d_swfrac, ctx->d_sfq, ctx->d_pcg, &epi, ctx->ev_join);
        ctx->cg_nwu = gather_blocks(N);
    } else if (ctx->cfg.lpme) {
        s = recip_run<R>(ctx, ctx->eg, 0, nullptr, ctx->d_z, 2, 1, 0, DIELECTRIC, ctx->cfg.kappa, 1, ctx->d_t2,
                         nullptr, nullptr, nullptr, false, ctx->d_swfrac, ctx->d_sfq, ctx->d_pcg);
    }
    // join before any return so that a capture never ends with an open fork
    ctx->spread_chk.on = 0;
    if (!with_plane) HIPCHK(hipStreamWaitEvent(st, ctx->ev_join, 0));
    if (s != ADMP_OK) return s;
    if (!fused) {
        ProfScope ps_(ctx, 11);
        hipLaunchKernelGGL(k_cg_w, dim3(ctx->pcg_nb, F), dim3(PCG_BLOCK), 0, st, N, f1, pol, ctx->d_z, ctx->d_t1,
                           ctx->cfg.lpme ? ctx->d_t2 : nullptr, ctx->d_Ap, cg_parts(ctx), ctx->d_pcg);
        ctx->cg_nwu = ctx->pcg_nb;
    }
    HIPCHK(hipGetLastError());
    return ADMP_OK;
}

// partners of the block preconditioner, or null for the diagonal one
static const int* pc_j(const admp_ctx* ctx) { return ctx->pc_maxm > 1 ? ctx->d_pcj : nullptr; }

// k_cg_update of parity cur with products of type R (z copy for the pair kernel
// and fractional dipoles typed R)
template <class R>
static void launch_update(admp_ctx* ctx, int cur, double* x, double tol, const double* pol) {
    hipLaunchKernelGGL(k_cg_update<R>, dim3(ctx->pcg_nb, ctx->F), dim3(PCG_BLOCK), 0, ctx->stream, ctx->N,
                       ctx->cg_nwu, cur, tol, ctx->d_box, pol, x, ctx->d_r, ctx->d_z, ctx->d_p, ctx->d_s, ctx->d_Ap,
                       sizeof(R) == 8 ? nullptr : reinterpret_cast<R*>(ctx->d_vR),
                       ctx->cfg.lpme ? reinterpret_cast<R*>(ctx->d_sfq) : nullptr, pc_j(ctx), ctx->d_pcd,
                       cg_parts(ctx), ctx->d_pcg);
}

// true residual r = rhs - A x with FP64 products (RC-typed stored pair
// coefficients), z = M^-1 r and the product inputs of z; mode 1 starts a solve,
// mode 2 only replaces the residual (direction kept).  dbl: the following
// iterations use FP64 products.  The next update has parity 0.
template <class RC>
static admp_status pcg_restart(admp_ctx* ctx, const double* rhs, const double* x, const double* pol, double f1,
                               double tol, int mode, bool dbl, bool x_zero = false, double* set_x_zero = nullptr) {
    const int N = ctx->N, F = ctx->F;
    hipStream_t st = ctx->stream;
    admp_status s;
    double* xz = set_x_zero;  // the start kernel zeroes x (a cold start whose x was not cleared)
    // from x = 0 the residual is the right-hand side: no product
    const double* Ax = x_zero ? nullptr : ctx->d_res;
    if (!x_zero && (s = matvec<RC, double>(ctx, x, ctx->d_res, pol, f1)) != ADMP_OK) return s;
    ProfScope ps_(ctx, 11);
    const double* wd = ctx->cfg.lpme ? ctx->d_wd : nullptr;
    if (dbl)
        hipLaunchKernelGGL(k_cg_start<double>, dim3(ctx->pcg_nb, F), dim3(PCG_BLOCK), 0, st, N, ctx->d_box, wd, pol,
                           rhs, Ax, ctx->d_r, ctx->d_z, (double*)nullptr,
                           reinterpret_cast<double*>(ctx->d_swfrac),
                           ctx->cfg.lpme ? reinterpret_cast<double*>(ctx->d_sfq) : nullptr, pc_j(ctx), ctx->d_pcd,
                           cg_parts(ctx), xz);
    else
        hipLaunchKernelGGL(k_cg_start<float>, dim3(ctx->pcg_nb, F), dim3(PCG_BLOCK), 0, st, N, ctx->d_box, wd, pol,
                           rhs, Ax, ctx->d_r, ctx->d_z, reinterpret_cast<float*>(ctx->d_vR),
                           reinterpret_cast<float*>(ctx->d_swfrac),
                           ctx->cfg.lpme ? reinterpret_cast<float*>(ctx->d_sfq) : nullptr, pc_j(ctx), ctx->d_pcd,
                           cg_parts(ctx), xz);
    hipLaunchKernelGGL(k_cg_publish, dim3(F), dim3(PCG_BLOCK), 0, st, mode, tol, cg_parts(ctx), ctx->d_pcg);
    HIPCHK(hipGetLastError());
    return ADMP_OK;
}

// Preconditioned CG on A x = rhs (Jacobi preconditioner pol/DIELECTRIC) with
// all scalars on the device, in the Chronopoulos-Gear form: each iteration is
// pcg_apply (w = A z) + k_cg_update; chunks of iterations are replayed from
// captured HIP graphs and the
// host reads the per-frame state once per chunk.  Real-space products use the
// stored coefficients (FP64 on both paths).  With fp32 (mixed contexts only)
// the reciprocal products start in FP32: the recursively updated residual is
// replaced by the true FP64-product residual whenever it has dropped by 1e-2
// ("reliable updates", Sleijpen & van der Vorst 1996), and once the true
// residual fell by 1e3 (or after 32 FP32 iterations) the solve restarts its
// direction and continues with FP64 products.  x holds the initial guess.
template <class RC, class R>
static admp_status pcg(admp_ctx* ctx, const double* rhs, double* x, const double* pol, double f1, double tol,
                       int maxiter, std::vector<int>& it_out, std::vector<double>& res_out, bool fp32 = false,
                       bool cold = false, bool started = false) {
    const int N = ctx->N, F = ctx->F;
    hipStream_t st = ctx->stream;
    // optional first phase with FP32 reciprocal products (fp32 on a mixed context)
    bool dbl = sizeof(R) == 8 || !fp32;
    int fp32_iters = 0;
    std::vector<int> it_base(F, 0);
    admp_status s;
    // (started: the caller already ran this first restart, in its set-up graph)
    if (!started && (s = pcg_restart<RC>(ctx, rhs, x, pol, f1, tol, 1, dbl, cold)) != ADMP_OK) return s;
    const int chunk = 4;  // FP32-phase chunk (host reads the state after each)
    std::vector<PcgState> hs(F);
    int done_iters = 0;
    auto read_state = [&](double& worst, bool& all) -> admp_status {
        HIPCHK(hipMemcpyAsync(hs.data(), ctx->d_pcg, F * sizeof(PcgState), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        all = true;
        worst = 0;
        for (int f = 0; f < F; ++f) {
            all = all && hs[f].done;
            worst = std::max(worst, hs[f].res);
        }
        return ADMP_OK;
    };
    // launch nrun iterations (FP64 or FP32 reciprocal products), eagerly or as a
    // captured HIP graph replayed on later calls
    // (par: buffer parity of the chunk's first iteration; r and s alternate
    // between two buffers, and a restart leaves its state in parity 0)
    auto launch_chunk = [&](int nrun, bool use_dbl, int par) -> admp_status {
        auto iterate = [&]() -> admp_status {
            for (int c = 0; c < nrun; ++c) {
                const int cur = (par + c) & 1;
                admp_status s2;
                if (use_dbl) s2 = pcg_apply<RC, double>(ctx, pol, f1, cur, tol);
                else s2 = pcg_apply<RC, float>(ctx, pol, f1, cur, tol);
                if (s2 != ADMP_OK) return s2;
                ProfScope ps_(ctx, 11);
                if (use_dbl) launch_update<double>(ctx, cur, x, tol, pol);
                else launch_update<float>(ctx, cur, x, tol, pol);
            }
            HIPCHK(hipGetLastError());
            return ADMP_OK;
        };
        if (!ctx->use_graphs) return iterate();
        const int type = (sizeof(RC) == 8 ? 1 : 0) + (use_dbl ? 2 : 0);
        admp_ctx::ChunkGraph* cg = nullptr;
        auto same = [&](const admp_ctx::ChunkGraph& g) {
            return g.type == type && g.nrun == nrun && g.par == par && g.rhs == rhs && g.x == x && g.tol == tol &&
                   g.f1 == f1 && g.pc == (ctx->pc_maxm > 1);
        };
        for (auto& g : ctx->graphs)
            if (g.exec && same(g)) cg = &g;
        if (!cg) {
            // captured on the second use of a variant: a chunk length a context meets
            // once (a server's many short-lived contexts) runs eagerly instead of
            // paying for a capture and an instantiation
            bool seen = false;
            for (auto& g : ctx->chunk_seen)
                if (g.nrun > 0 && same(g)) seen = true;
            if (!seen) {
                admp_ctx::ChunkGraph& k = ctx->chunk_seen[ctx->chunk_seen_next++ % 16];
                k = admp_ctx::ChunkGraph();
                k.type = type; k.nrun = nrun; k.par = par; k.rhs = rhs; k.x = x; k.tol = tol; k.f1 = f1;
                k.pc = ctx->pc_maxm > 1;
                return iterate();
            }
        }
        if (!cg) {
            admp_ctx::ChunkGraph* slot = nullptr;
            for (auto& g : ctx->graphs)
                if (!g.exec) { slot = &g; break; }
            if (!slot) { clear_chunk_graphs(ctx); slot = &ctx->graphs[0]; }
            const bool prof = ctx->prof;
            ctx->prof = false;
            hipGraph_t graph = nullptr;
            hipError_t be = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
            if (be != hipSuccess) {
                ctx->prof = prof;
                (void)hipGetLastError();
                return iterate();  // capture unavailable: launch eagerly
            }
            admp_status cs = iterate();
            hipError_t ce = hipStreamEndCapture(st, &graph);
            ctx->prof = prof;
            hipGraphExec_t exec = nullptr;
            hipError_t ie = hipErrorUnknown;
            if (cs == ADMP_OK && ce == hipSuccess && graph)
                ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (graph) (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) {
                // the capture was invalidated (or failed): nothing of it ran, so the
                // chunk is launched eagerly; errors of the eager launch are reported
                (void)hipGetLastError();
                ctx->err.clear();
                return iterate();
            }
            slot->exec = exec; slot->type = type; slot->nrun = nrun; slot->par = par; slot->rhs = rhs; slot->x = x;
            slot->tol = tol; slot->f1 = f1; slot->pc = ctx->pc_maxm > 1;
            cg = slot;
        }
        ProfScope ps_(ctx, 12);
        HIPCHK(hipGraphLaunch(cg->exec, st));
        return ADMP_OK;
    };
    double worst = 0;
    bool all = false;
    // From x = 0 the start state cannot be converged: with FP64 products the first
    // chunk is queued at once, without waiting for that state (its residual is
    // only needed to steer the FP32 phase).
    const bool read_start = !(dbl && cold && maxiter > 0);
    if (read_start && (s = read_state(worst, all)) != ADMP_OK) return s;
    const double r0 = worst;
    double last_true = worst;
    for (;;) {
        const bool sw = fp32_iters >= 32;
        if (!dbl && done_iters > 0 && (all || sw || worst < 1e-2 * last_true)) {
            // replace the recursively updated residual by the true one (FP64 products);
            // switch to FP64 products after three orders of reduction or 32 FP32 iterations
            if ((s = pcg_restart<RC>(ctx, rhs, x, pol, f1, tol, 2, false)) != ADMP_OK) return s;
            if ((s = read_state(worst, all)) != ADMP_OK) return s;
            last_true = worst;
            if (!all && (sw || worst < 1e-3 * r0)) {
                // restart the direction: an FP32-built p is not conjugate under FP64 products
                dbl = true;
                for (int f = 0; f < F; ++f) it_base[f] += hs[f].iter;
                if ((s = pcg_restart<RC>(ctx, rhs, x, pol, f1, tol, 1, true)) != ADMP_OK) return s;
                if ((s = read_state(worst, all)) != ADMP_OK) return s;
            }
        }
        if (all || done_iters >= maxiter) break;
        if (dbl) break;  // FP64 products from here on: pipelined loop below
        const int nrun = std::min(chunk, maxiter - done_iters);
        fp32_iters += nrun;
        // (FP32 chunks have even length but a last one cut by maxiter: parity 0)
        if ((s = launch_chunk(nrun, false, 0)) != ADMP_OK) return s;
        done_iters += nrun;
        if ((s = read_state(worst, all)) != ADMP_OK) return s;
    }
    if (dbl && !all && done_iters < maxiter) {
        // The host stays one chunk ahead: chunk k+1 is launched before the state
        // after chunk k is read (pinned copy + event), so the GPU does not idle at
        // chunk boundaries.  A chunk launched after convergence costs little: the
        // product kernels return at once for converged frames.
        const int ch = 4;
        int q[2], qn[2], nq = 0, last_slot = 1, par = 0;  // (qn: iterations of each chunk in flight)
        auto post = [&](int sl) -> admp_status {
            HIPCHK(hipMemcpyAsync(ctx->h_pcg + (size_t)sl * F, ctx->d_pcg, F * sizeof(PcgState),
                                  hipMemcpyDeviceToHost, st));
            HIPCHK(hipEventRecord(ctx->ev_state[sl], st));
            return ADMP_OK;
        };
        // nrun: iterations of the next chunk (at most ch)
        auto launch_next = [&](int want = 1 << 30) -> admp_status {
            const int nrun = std::min(std::min(ch, want), maxiter - done_iters);
            admp_status s2;
            if ((s2 = launch_chunk(nrun, true, par)) != ADMP_OK) return s2;
            done_iters += nrun;
            par = (par + nrun) & 1;
            last_slot ^= 1;
            if ((s2 = post(last_slot)) != ADMP_OK) return s2;
            qn[nq] = nrun;
            q[nq++] = last_slot;
            return ADMP_OK;
        };
        // A chunk launched after convergence still costs every graph node's launch
        // (~10 us each), so the next chunk is not launched ahead when the chunk in
        // flight is predicted to finish the solve: each frame's residual is
        // extrapolated geometrically from its last two states.  A wrong
        // prediction costs one host round trip before the next chunk starts.
        std::vector<PcgState> older = hs;
        bool have_older = false, have_hs = read_start;
        // the most iterations any frame is predicted to need from the last state
        // read (geometric extrapolation from its last two states); -1: no prediction
        auto predicted_need = [&]() -> double {
            if (!have_older) return -1.0;
            double worst_need = 0.0;
            for (int f = 0; f < F; ++f) {
                if (hs[f].done) continue;
                const int di = hs[f].iter - older[f].iter;
                if (di <= 0 || !(hs[f].res > 0.0) || !(hs[f].res < older[f].res)) return -1.0;
                const double rate = std::log(hs[f].res / older[f].res) / di;   // per iteration, < 0
                worst_need = std::max(worst_need, std::log(tol / hs[f].res) / rate);
            }
            return worst_need;
        };
        // (both with a margin of 1.5 iterations; called with one chunk in flight)
        auto finishes_in_flight = [&]() -> bool {
            const double need = predicted_need();
            return need >= 0.0 && need <= qn[0] - 1.5;
        };
        // a chunk launched ahead of the one in flight is cut to the predicted
        // remainder, so that the solve does not end in iterations that only
        // return (each still costs its launches)
        auto ahead_size = [&]() -> int {
            const double need = predicted_need();
            if (need < 0.0) return ch;
            return std::max(1, std::min(ch, (int)std::ceil(need - qn[0] + 1.5)));
        };
        if ((s = launch_next()) != ADMP_OK) return s;
        while (nq > 0) {
            if (done_iters < maxiter && nq < 2 && !finishes_in_flight())
                if ((s = launch_next(ahead_size())) != ADMP_OK) return s;
            const int sl = q[0];
            q[0] = q[1];
            qn[0] = qn[1];
            --nq;
            HIPCHK(hipEventSynchronize(ctx->ev_state[sl]));
            older = hs;
            have_older = have_hs;
            have_hs = true;
            all = true;
            for (int f = 0; f < F; ++f) {
                hs[f] = ctx->h_pcg[(size_t)sl * F + f];
                all = all && hs[f].done;
            }
            if (all) break;
            if (nq == 0 && done_iters < maxiter)
                if ((s = launch_next()) != ADMP_OK) return s;   // the prediction missed
        }
    }
    it_out.resize(F);
    res_out.resize(F);
    for (int f = 0; f < F; ++f) { it_out[f] = it_base[f] + hs[f].iter; res_out[f] = hs[f].res; }
    HIPCHK(hipGetLastError());
    return ADMP_OK;
}

// right-hand side b = -dE/dU at U = 0: real permanent field (+ stored pair
// coefficients), minus the reciprocal dE/dmu of the permanent multipoles, plus
// the self term 2 D f1 mu.  RB: precision of the reciprocal part.
template <class R, class RB>
static admp_status solver_rhs(admp_ctx* ctx, const double* pol, const double* thole, const PairParams& pp) {
    const admp_config& c = ctx->cfg;
    const int N = ctx->N, F = ctx->F;
    const size_t NF = (size_t)N * F;
    hipStream_t st = ctx->stream;
    const double kappa = c.lpme ? c.kappa : 0.0;
    const double f1 = c.lpme ? kappa / SQRT_PI * 2.0 * kappa * kappa / 3.0 : 0.0;
    const int lmax_eff = c.lmax == 0 ? 1 : c.lmax;
    R* site = reinterpret_cast<R*>(ctx->d_site);
    hipLaunchKernelGGL(k_pack_site<R>, dim3(nblk(N, 128), F), dim3(128), 0, st, N, ctx->d_Qc, pol, thole,
                       (const double*)nullptr, site);
    // the real-space permanent field (and the stored coefficients and preconditioner
    // blocks) runs on the side stream beside the reciprocal part: they share only
    // inputs; k_rhs_finish joins them
    const hipStream_t sr = c.lpme ? ctx->side : st;
    if (c.lpme) {
        HIPCHK(hipEventRecord(ctx->ev_fork, st));
        HIPCHK(hipStreamWaitEvent(sr, ctx->ev_fork, 0));
    }
    { ProfScope ps_(ctx, 1, sr);
    hipLaunchKernelGGL((k_real_permfield<R>), dim3(nblk((size_t)N * TPA, PAIR_BLOCK), F), dim3(PAIR_BLOCK), 0, sr, N,
                       ctx->cap, ctx->d_pos, ctx->d_box, site, ctx->d_nbr, ctx->d_ncnt, pp, ctx->d_b,
                       reinterpret_cast<R*>(ctx->d_pcoef), ctx->d_pnear, ctx->d_pcn);
    if (ctx->pc_maxm > 1) {
        hipLaunchKernelGGL(k_pc_children, dim3(nblk(N, 256), F), dim3(256), 0, sr, N, ctx->cap, ctx->d_nbr,
                           ctx->d_pnear, ctx->d_pcn, ctx->d_pcc);
        hipLaunchKernelGGL((k_pc_clusters<R>), dim3(nblk((size_t)N * PC_LANES, 256), F), dim3(256), 0, sr, N, ctx->cap,
                           ctx->pc_maxm, ctx->d_pos, ctx->d_box, site, pol, ctx->d_nbr, ctx->d_ncnt, ctx->d_pnear, ctx->d_pcn,
                           ctx->d_pcc, pp, ctx->d_pcj, ctx->d_pcd);
    } }
    HIPCHK(hipGetLastError());
    if (c.lpme) {
        HIPCHK(hipEventRecord(ctx->ev_join, sr));
        admp_status s = recip_run<RB>(ctx, ctx->eg, 0, ctx->d_Qc, nullptr, 1, lmax_eff, 0, DIELECTRIC, kappa, 1,
                                      ctx->d_t2, nullptr, nullptr, nullptr, false);
        // join before any return, so that no error leaves the side stream unjoined
        HIPCHK(hipStreamWaitEvent(st, ctx->ev_join, 0));
        if (s != ADMP_OK) return s;
        hipLaunchKernelGGL(k_rhs_finish, dim3(nblk(N, 256), F), dim3(256), 0, st, N, f1, ctx->d_Qc, ctx->d_t2,
                           ctx->d_b);
    }
    HIPCHK(hipGetLastError());
    return ADMP_OK;
}

// local frames and global multipoles of the context's copy of Q_local
static admp_status launch_frames_fwd(admp_ctx* ctx) {
    const int N = ctx->N, F = ctx->F;
    { ProfScope ps_(ctx, 9); hipLaunchKernelGGL(k_frames_fwd, dim3(nblk(N, 128), F), dim3(128), 0, ctx->stream, N, ctx->nh,
                       ctx->use_frames, ctx->d_pos, ctx->d_box, ctx->d_axis_type, ctx->d_axis_idx, ctx->d_Qloc,
                       ctx->d_frames, ctx->d_Qc); }
    HIPCHK(hipGetLastError());
    return ADMP_OK;
}

// Induced-dipole solve: U (ctx->d_U) from A U = b.
template <class R>
static admp_status solve(admp_ctx* ctx, const double* pol, const double* thole, const PairParams& pp,
                         const admp_solver_opts* o, const double* U_in, SolveInfo& info) {
    const admp_config& c = ctx->cfg;
    const int N = ctx->N, F = ctx->F;
    const size_t NF = (size_t)N * F;
    hipStream_t st = ctx->stream;
    const double kappa = c.lpme ? c.kappa : 0.0;
    const double f1 = c.lpme ? kappa / SQRT_PI * 2.0 * kappa * kappa / 3.0 : 0.0;
    double* U = ctx->d_U;
    const int kind = o ? o->kind : ADMP_SOLVER_PCG;
    const double tol = o ? o->tol : 1e-6;
    const int maxiter = o ? (kind == ADMP_SOLVER_STEPS ? o->steps : o->maxiter) : 500;
    const bool mixed = sizeof(R) == 4;
    {
        const int pk = o ? o->precond : ADMP_PRECOND_CLUSTERS;
        ctx->pc_maxm = pk == ADMP_PRECOND_DIAGONAL ? 1 : pk == ADMP_PRECOND_PAIRS ? 2 : PC_MAXM;
    }
    const bool warm = o && o->warm_start && U_in;
    const bool fp32 = mixed && o && o->fp32_products;
    const bool started = kind == ADMP_SOLVER_PCG && !fp32 && !warm;   // see below
    if (warm)
        HIPCHK(hipMemcpyAsync(U, U_in, NF * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    else if (!started)
        HIPCHK(hipMemsetAsync(U, 0, NF * 3 * sizeof(double), st));
    admp_status s;
    // right-hand side and the stored real-space dipole coupling in FP64 on both
    // paths: the near-field coupling is what makes close contacts ill-conditioned,
    // and FP32 rounding of it changes the solution beyond the residual tolerance.
    // A cold FP64-product CG also starts here (residual, preconditioner, first
    // direction): the set-up is one captured graph from its second use on.
    if (started) {
        std::string key;
        key_add(key, pp);
        key_add(key, tol);
        key_add(key, ctx->pc_maxm);
        s = run_phase(ctx, ctx->g_setup, key, [&]() -> admp_status {
            admp_status s2 = launch_frames_fwd(ctx);
            if (s2 != ADMP_OK) return s2;
            if ((s2 = solver_rhs<double, double>(ctx, pol, thole, pp)) != ADMP_OK) return s2;
            return pcg_restart<double>(ctx, ctx->d_b, U, pol, f1, tol, 1, true, true, U);
        });
        if (s != ADMP_OK) return s;
        // a replayed set-up graph sorted the sites for the spread as well (the host
        // state its capture left): later launches and captures must not sort again
        if (c.lpme) ctx->prepped = &ctx->eg;
    } else {
        if ((s = launch_frames_fwd(ctx)) != ADMP_OK) return s;
        if ((s = solver_rhs<double, double>(ctx, pol, thole, pp)) != ADMP_OK) return s;
    }
    info.n_iter.assign(F, 0);
    info.conv.assign(F, 0);
    info.res.assign(F, 0.0);
    std::vector<int> act(F, 1);
    std::vector<double> h;
    // With FP64 products throughout, the CG's recursively updated residual (which
    // decides convergence) is what is reported: its drift from the true residual
    // is rounding, far below any tolerance, and the check would cost a full
    // product and a host synchronisation.  A solve that ran FP32 products ends
    // with the true residual below.
    if (kind == ADMP_SOLVER_PCG && !mixed) {
        std::vector<int> it;
        std::vector<double> rr;
        if ((s = pcg<double, double>(ctx, ctx->d_b, U, pol, f1, tol, maxiter, it, rr, false, !warm, started)) !=
            ADMP_OK)
            return s;
        for (int f = 0; f < F; ++f) {
            info.n_iter[f] = it[f];
            info.res[f] = rr[f];
            info.conv[f] = rr[f] < tol ? 1 : 0;
        }
        HIPCHK(hipGetLastError());
        return ADMP_OK;
    } else if (kind == ADMP_SOLVER_PCG) {
        // FP64 products by default; with fp32_products the first phase runs the
        // reciprocal products in FP32 with FP64 true residuals (reliable updates)
        std::vector<int> it;
        std::vector<double> rr;
        if ((s = pcg<double, R>(ctx, ctx->d_b, U, pol, f1, tol, maxiter, it, rr, fp32, !warm, started)) != ADMP_OK)
            return s;
        for (int f = 0; f < F; ++f) info.n_iter[f] = it[f];
        if (!fp32) {
            for (int f = 0; f < F; ++f) {
                info.res[f] = rr[f];
                info.conv[f] = rr[f] < tol ? 1 : 0;
            }
            HIPCHK(hipGetLastError());
            return ADMP_OK;
        }
    } else {
        // Jacobi sweeps (pme.py:380-422)
        HIPCHK(hipMemcpyAsync(ctx->d_active, act.data(), F * sizeof(int), hipMemcpyHostToDevice, st));
        int i = 0;
        for (i = 0; i < maxiter; ++i) {
            if ((s = matvec<double, R>(ctx, U, ctx->d_Ap, pol, f1)) != ADMP_OK) return s;
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
        if (kind == ADMP_SOLVER_REFERENCE) {
            HIPCHK(hipGetLastError());
            return ADMP_OK;
        }
    }
    // true residual of the final iterate (FP64 products)
    if ((s = matvec<double, double>(ctx, U, ctx->d_res, pol, f1)) != ADMP_OK) return s;
    hipLaunchKernelGGL(k_lincomb, dim3(nblk(NF * 3, 256)), dim3(256), 0, st, NF * 3, 1.0, ctx->d_res, -1.0, ctx->d_b,
                       ctx->d_res);
    if ((s = dots(ctx, ctx->d_res, nullptr, nullptr, pol, h)) != ADMP_OK) return s;
    for (int f = 0; f < F; ++f) {
        info.res[f] = h[3 * f + 2];
        if (kind != ADMP_SOLVER_STEPS) info.conv[f] = h[3 * f + 2] < tol ? 1 : 0;
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
    // local frames and global multipoles (inside the solver's set-up when polarisable)
    if (!lpol) { const admp_status sf = launch_frames_fwd(ctx); if (sf != ADMP_OK) return sf; }
    SolveInfo info;
    admp_status s;
    const double* U = nullptr;
    if (lpol) {
        if ((s = solve<R>(ctx, pol, thole, pp, opts, U_inout, info)) != ADMP_OK) return s;
        U = ctx->d_U;
        if (U_inout) HIPCHK(hipMemcpyAsync(U_inout, U, NF * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    // ---------------- final pass: energy and gradients with U fixed ----------------
    // (energy and dE/dbox land in context buffers; one launch copies every output)
    auto final_body = [&]() -> admp_status {
    hipLaunchKernelGGL(k_pack_site<R>, dim3(nblk(N, 128), F), dim3(128), 0, st, N, ctx->d_Qc, pol, thole, U, site);
    // the real-space pass runs on the side stream beside the reciprocal one (their
    // outputs are separate buffers; k_self, which adds to both, comes after the join)
    const hipStream_t sr = lpme ? ctx->side : st;
    if (lpme) {
        HIPCHK(hipEventRecord(ctx->ev_fork, st));
        HIPCHK(hipStreamWaitEvent(sr, ctx->ev_fork, 0));
    }
    if (lpol)
        { ProfScope ps_(ctx, 0, sr); hipLaunchKernelGGL((k_real_final<R, true>), pgrid, dim3(PAIR_BLOCK), 0, sr, N, ctx->cap, ctx->d_pos, ctx->d_box,
                           site, ctx->d_nbr, ctx->d_ncnt, pp, ctx->d_e_atom, ctx->d_gpos, ctx->d_gq, ctx->d_vir_atom,
                           ctx->d_gpol, ctx->d_gth); }
    else
        { ProfScope ps_(ctx, 0, sr); hipLaunchKernelGGL((k_real_final<R, false>), pgrid, dim3(PAIR_BLOCK), 0, sr, N, ctx->cap, ctx->d_pos, ctx->d_box,
                           site, ctx->d_nbr, ctx->d_ncnt, pp, ctx->d_e_atom, ctx->d_gpos, ctx->d_gq, ctx->d_vir_atom,
                           ctx->d_gpol, ctx->d_gth); }
    HIPCHK(hipGetLastError());
    if (!lpol) {
        HIPCHK(hipMemsetAsync(ctx->d_gpol, 0, NF * sizeof(double), sr));
        HIPCHK(hipMemsetAsync(ctx->d_gth, 0, NF * sizeof(double), sr));
    }
    admp_status s2 = ADMP_OK;
    if (lpme) {
        HIPCHK(hipEventRecord(ctx->ev_join, sr));
        s2 = recip_run<R>(ctx, ctx->eg, 0, ctx->d_Qc, U, lpol ? 3 : 1, lmax_eff, 0, DIELECTRIC, kappa, 0, nullptr,
                          ctx->d_gq_rec, ctx->d_gpos_rec, ctx->d_vir_rec, true);
        HIPCHK(hipStreamWaitEvent(st, ctx->ev_join, 0));
        if (s2 != ADMP_OK) return s2;
    } else {
        HIPCHK(hipMemsetAsync(ctx->d_gq_rec, 0, NF * 10 * sizeof(double), st));
        HIPCHK(hipMemsetAsync(ctx->d_gpos_rec, 0, NF * 3 * sizeof(double), st));
        HIPCHK(hipMemsetAsync(ctx->d_vir_rec, 0, NF * 9 * sizeof(double), st));
        HIPCHK(hipMemsetAsync(ctx->d_rec, 0, F * 7 * sizeof(double), st));
    }
    // the self terms, and (with gradients) the reciprocal dE/dQ_global added to the real one
    hipLaunchKernelGGL(k_self, dim3(nblk(N, 256), F), dim3(256), 0, st, N, kappa, lpme, lpol, ctx->d_Qc, U, pol,
                       ctx->d_e_self, ctx->d_e_pol, ctx->d_gq, lpol ? ctx->d_gpol : nullptr,
                       want_grad ? ctx->d_gq_rec : (const double*)nullptr);
    // per-frame energy components: real, self, polarisation penalty -> d_red [0..F), [F..2F), [2F..3F)
    hipLaunchKernelGGL(k_reduce3, dim3(3, F), dim3(RED_BLOCK), 0, st, N,
                       Reduce3{{ctx->d_e_atom, ctx->d_e_self, ctx->d_e_pol}}, ctx->d_red);
    // total energy per frame on the device (real + self + polarisation + reciprocal)
    hipLaunchKernelGGL(k_energy_total, dim3(nblk(F, 64)), dim3(64), 0, st, F, ctx->d_red, ctx->d_rec, ctx->d_eout);
    if (want_grad) {
        // dE/dQ_global total -> frames backward
        { ProfScope ps_(ctx, 9); hipLaunchKernelGGL(k_frames_bwd, dim3(nblk(N, 128), F), dim3(128), 0, st, N, ctx->nh, ctx->use_frames, ctx->d_pos,
                           ctx->d_box, ctx->d_axis_type, ctx->d_axis_idx, Q_local, ctx->d_gq, ctx->d_gQl,
                           ctx->d_contrib, ctx->d_vir_fr); }
        // forces: real + reciprocal + the frames' contributions
        hipLaunchKernelGGL(k_frames_gather, dim3(nblk(N, 256), F), dim3(256), 0, st, N, ctx->d_rev_off, ctx->d_rev_ent,
                           ctx->d_contrib, ctx->d_gpos_rec, ctx->d_gpos);
        // box: direct terms (real pairs + frames), reciprocal A-terms and C_k terms
        // (real pairs and frames are summed separately and added in k_box_grad, in
        // that order)
        hipLaunchKernelGGL(k_reduce_vir, dim3(9, F, 3), dim3(RED_BLOCK), 0, st, N,
                           Vir3{{ctx->d_vir_atom, ctx->d_vir_fr, ctx->d_vir_rec},
                                {ctx->d_red9a, ctx->d_red9c, ctx->d_red9b}});
        hipLaunchKernelGGL(k_box_grad, dim3(nblk(F, 64)), dim3(64), 0, st, F, ctx->d_box, ctx->d_red9a,
                           ctx->d_red9c, lpme ? ctx->d_red9b : nullptr, (const double*)nullptr, lpme ? ctx->d_rec : nullptr,
                           c.K1, c.K2, c.K3, 0, ctx->d_bout, 0);
    }
    HIPCHK(hipGetLastError());
    return ADMP_OK;
    };
    if (lpol && lpme) {
        // graph key: every value the captured launches take by value
        std::string key;
        key_add(key, want_grad);
        key_add(key, pp);
        if ((s = run_phase(ctx, ctx->g_final, key, final_body)) != ADMP_OK) return s;
    } else if ((s = final_body()) != ADMP_OK) {
        return s;
    }
    {
        // every requested output in one launch
        CopyJobs cj{};
        cj.src[0] = ctx->d_eout; cj.dst[0] = energy; cj.n[0] = energy ? F : 0;
        if (want_grad) {
            cj.src[1] = ctx->d_gpos; cj.dst[1] = dpos; cj.n[1] = dpos ? NF * 3 : 0;
            cj.src[2] = ctx->d_gQl; cj.dst[2] = dQ_local; cj.n[2] = dQ_local ? NF * ctx->nh : 0;
            cj.src[3] = ctx->d_gpol; cj.dst[3] = dpol; cj.n[3] = dpol ? NF : 0;
            cj.src[4] = ctx->d_gth; cj.dst[4] = dthole; cj.n[4] = dthole ? NF : 0;
            cj.src[5] = ctx->d_bout; cj.dst[5] = dbox; cj.n[5] = dbox ? (size_t)F * 9 : 0;
        }
        size_t nmax = 0;
        for (int k = 0; k < 6; ++k) nmax = std::max(nmax, cj.n[k]);
        if (nmax) hipLaunchKernelGGL(k_copy4, dim3(nblk(nmax, 256), 6), dim3(256), 0, st, cj);
        HIPCHK(hipGetLastError());
    }
    // Without stats the call returns once everything is queued, provided the
    // geometry flags have already reached the host (the solve's convergence
    // reads waited on the device after they were copied); otherwise, and with
    // stats (energy components: one pinned copy after everything is queued),
    // it waits for the stream.
    bool geom_known = false;
    if (!stats) {
        const hipError_t q = hipEventQuery(ctx->ev_geom);
        if (q == hipSuccess) geom_known = true;
        else if (q != hipErrorNotReady) HIPCHK(q);
    }
    if (!geom_known) {
        if (stats) {
            HIPCHK(hipMemcpyAsync(ctx->h_comp, ctx->d_red, 3 * F * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(ctx->h_comp + 3 * F, ctx->d_rec, 7 * F * sizeof(double), hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipStreamSynchronize(st));
    }
    if ((s = geom_check(ctx)) != ADMP_OK) return s;
    const double* comp = ctx->h_comp;
    const double* rec = ctx->h_comp + 3 * F;
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

static admp_status admp_compute_impl(admp_ctx* ctx, const double* pos, const double* box, const double* Q_local,
                                    const double* pol, const double* thole, const double* mScales,
                                    const double* pScales, const double* dScales, int32_t n_scales, double* U_inout,
                                    const admp_solver_opts* opts, int32_t want_grad, double* energy, double* dpos,
                                    double* dbox, double* dQ_local, double* dpol, double* dthole, admp_stats* stats) {
    if (!ctx || !pos || !box || !Q_local) return ADMP_ERR_ARG;
    if (!ctx->has_topo) { ctx->err = "admp_set_topology first"; return ADMP_ERR_STATE; }
    { const admp_status ns = nbr_settle(ctx); if (ns != ADMP_OK) return ns; }
    if (!all_frames_have_pairs(ctx)) {
        ctx->err = "no pair list for every frame: admp_set_pairs per frame or admp_build_neighbors first";
        return ADMP_ERR_STATE;
    }
    if (ctx->cfg.lpol && (!pol || !thole)) { ctx->err = "pol and thole required"; return ADMP_ERR_ARG; }
    {
        // the caller's per-atom inputs into context buffers (stable arguments of
        // the captured phases), in one launch
        CopyJobs cj{};
        cj.src[0] = Q_local; cj.dst[0] = ctx->d_Qloc; cj.n[0] = (size_t)ctx->N * ctx->nh;
        if (ctx->cfg.lpol) {
            cj.src[1] = pol; cj.dst[1] = ctx->d_pol; cj.n[1] = ctx->N;
            cj.src[2] = thole; cj.dst[2] = ctx->d_thole; cj.n[2] = ctx->N;
            pol = ctx->d_pol;
            thole = ctx->d_thole;
        }
        hipLaunchKernelGGL(k_copy4, dim3(nblk(cj.n[0], 256), ctx->cfg.lpol ? 3 : 1), dim3(256), 0, ctx->stream, cj);
        HIPCHK(hipGetLastError());
        Q_local = ctx->d_Qloc;
    }
    if (n_scales < 1 || n_scales > 8) { ctx->err = "n_scales must be 1..8"; return ADMP_ERR_ARG; }
    admp_status s = upload_geometry(ctx, pos, box);
    if (s != ADMP_OK) return s;
    // the geometry flags go to the host early: a call whose solve already waited
    // on the device checks them without a final synchronisation
    if ((s = geom_fetch(ctx)) != ADMP_OK) return s;
    HIPCHK(hipEventRecord(ctx->ev_geom, ctx->stream));
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
                       ctx->d_gpos, ctx->d_vir_atom, ctx->d_gpos_rec /* (F,N,3) dE/dc */, 3); }
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
            hipLaunchKernelGGL(k_reduce_cols, dim3(9, F), dim3(RED_BLOCK), 0, st, N, 9, ctx->d_vir_rec, ctx->d_red9c, 1);
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
    hipLaunchKernelGGL(k_reduce_cols, dim3(1, F), dim3(RED_BLOCK), 0, st, N, 1, ctx->d_e_atom, ctx->d_red9a, 0);
    hipLaunchKernelGGL(k_reduce_cols, dim3(1, F), dim3(RED_BLOCK), 0, st, N, 1, ctx->d_e_self, ctx->d_red9a + F, 0);
    std::vector<double> comp(2 * F);
    HIPCHK(hipMemcpyAsync(comp.data(), ctx->d_red9a, 2 * F * sizeof(double), hipMemcpyDeviceToHost, st));
    if (want_grad) {
        hipLaunchKernelGGL(k_reduce_cols, dim3(9, F), dim3(RED_BLOCK), 0, st, N, 9, ctx->d_vir_atom, ctx->d_red9b, 0);
        if (dbox)
            hipLaunchKernelGGL(k_box_grad, dim3(nblk(F, 64)), dim3(64), 0, st, F, ctx->d_box, ctx->d_red9b,
                               (const double*)nullptr,
                               lpme ? ctx->d_red9c : nullptr, (const double*)nullptr, lpme ? ctx->d_rec : nullptr, c.dK1,
                               c.dK2, c.dK3, 1, dbox, 0);
        if (dpos) HIPCHK(hipMemcpyAsync(dpos, ctx->d_gpos, NF * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (dc) HIPCHK(hipMemcpyAsync(dc, gc, NF * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    if (geom_fetch(ctx) != ADMP_OK) return ADMP_ERR_HIP;
    HIPCHK(hipStreamSynchronize(st));
    { const admp_status g = geom_check(ctx); if (g != ADMP_OK) return g; }
    std::vector<double> e(F);
    for (int f = 0; f < F; ++f) e[f] = comp[f] + comp[F + f] + erec[f];
    if (energy) {
        HIPCHK(hipMemcpyAsync(energy, e.data(), F * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return ADMP_OK;
}

static admp_status admp_disp_compute_impl(admp_ctx* ctx, const double* pos, const double* box, const double* c_list,
                                         const double* mScales, int32_t n_scales, int32_t want_grad, double* energy,
                                         double* dpos, double* dbox, double* dc) {
    if (!ctx || !pos || !box || !c_list) return ADMP_ERR_ARG;
    { const admp_status ns = nbr_settle(ctx); if (ns != ADMP_OK) return ns; }
    if (!all_frames_have_pairs(ctx)) { ctx->err = "no pair list for every frame"; return ADMP_ERR_STATE; }
    if (n_scales < 1 || n_scales > 8) { ctx->err = "n_scales must be 1..8"; return ADMP_ERR_ARG; }
    admp_status s = upload_geometry(ctx, pos, box);
    if (s != ADMP_OK) return s;
    PairParams pp = make_pp(mScales, nullptr, nullptr, n_scales, 0.0);
    if (ctx->cfg.precision == ADMP_MIXED) return disp_impl<float>(ctx, c_list, pp, want_grad, energy, dpos, dbox, dc);
    return disp_impl<double>(ctx, c_list, pp, want_grad, energy, dpos, dbox, dc);
}

template <class R, int KIND>
static void launch_pair_short(admp_ctx* ctx, const double* prm, const PairParams& pp) {
    const int N = ctx->N, F = ctx->F;
    ProfScope ps_(ctx, 10);
    hipLaunchKernelGGL((k_pair_short<R, KIND>), dim3(nblk((size_t)N * TPA, PAIR_BLOCK), F), dim3(PAIR_BLOCK), 0,
                       ctx->stream, N, ctx->cap, ctx->d_pos, ctx->d_box, prm, ctx->d_nbr, ctx->d_ncnt, pp, 0,
                       ctx->d_e_atom, ctx->d_gpos, ctx->d_vir_atom, ctx->d_gq, sr_nprm(KIND));
}

// generate_pairwise_interaction(kernel) (pairwise.py:47-95) on the context's
// lists: per-site energies, forces, virial and parameter gradients from one
// k_pair_short launch, then fixed-order reductions.
static admp_status pairwise_impl(admp_ctx* ctx, int kind, const double* prm, const PairParams& pp, int want_grad,
                                 double* energy, double* dpos, double* dbox, double* dprm) {
    const int N = ctx->N, F = ctx->F;
    const size_t NF = (size_t)N * F;
    hipStream_t st = ctx->stream;
    const bool f32 = ctx->cfg.precision == ADMP_MIXED;
    switch (kind) {
        case SR_SLATER: f32 ? launch_pair_short<float, SR_SLATER>(ctx, prm, pp) : launch_pair_short<double, SR_SLATER>(ctx, prm, pp); break;
        case SR_TT_QQ_C6: f32 ? launch_pair_short<float, SR_TT_QQ_C6>(ctx, prm, pp) : launch_pair_short<double, SR_TT_QQ_C6>(ctx, prm, pp); break;
        case SR_TT_QQ: f32 ? launch_pair_short<float, SR_TT_QQ>(ctx, prm, pp) : launch_pair_short<double, SR_TT_QQ>(ctx, prm, pp); break;
        case SR_SLATER_DAMP: f32 ? launch_pair_short<float, SR_SLATER_DAMP>(ctx, prm, pp) : launch_pair_short<double, SR_SLATER_DAMP>(ctx, prm, pp); break;
        default: ctx->err = "unknown short-range kind"; return ADMP_ERR_ARG;
    }
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_cols, dim3(1, F), dim3(RED_BLOCK), 0, st, N, 1, ctx->d_e_atom, ctx->d_red9a, 0);
    if (energy) HIPCHK(hipMemcpyAsync(energy, ctx->d_red9a, F * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (want_grad) {
        hipLaunchKernelGGL(k_reduce_cols, dim3(9, F), dim3(RED_BLOCK), 0, st, N, 9, ctx->d_vir_atom, ctx->d_red9b, 0);
        if (dbox)
            hipLaunchKernelGGL(k_box_grad, dim3(nblk(F, 64)), dim3(64), 0, st, F, ctx->d_box, ctx->d_red9b,
                               (const double*)nullptr,
                               (const double*)nullptr, (const double*)nullptr, (const double*)nullptr, 1, 1, 1, 0, dbox, 0);
        if (dpos) HIPCHK(hipMemcpyAsync(dpos, ctx->d_gpos, NF * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (dprm)
            HIPCHK(hipMemcpyAsync(dprm, ctx->d_gq, NF * sr_nprm(kind) * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    HIPCHK(hipGetLastError());
    admp_status s;
    if ((s = geom_fetch(ctx)) != ADMP_OK) return s;
    HIPCHK(hipStreamSynchronize(st));
    return geom_check(ctx);
}

static admp_status admp_pairwise_compute_impl(admp_ctx* ctx, int32_t kind, const double* pos, const double* box,
                                             const double* prm, int32_t n_prm, const double* mScales,
                                             int32_t n_scales, int32_t want_grad, double* energy, double* dpos,
                                             double* dbox, double* dprm) {
    if (!ctx || !pos || !box || !prm) return ADMP_ERR_ARG;
    if (kind < 1 || kind > SR_KIND_MAX) { ctx->err = "unknown short-range kind"; return ADMP_ERR_ARG; }
    if (n_prm != sr_nprm(kind)) { ctx->err = "n_prm does not match the kind's parameter count"; return ADMP_ERR_ARG; }
    { const admp_status ns = nbr_settle(ctx); if (ns != ADMP_OK) return ns; }
    if (!all_frames_have_pairs(ctx)) { ctx->err = "no pair list for every frame"; return ADMP_ERR_STATE; }
    if (n_scales < 1 || n_scales > 8) { ctx->err = "n_scales must be 1..8"; return ADMP_ERR_ARG; }
    admp_status s = upload_geometry(ctx, pos, box);
    if (s != ADMP_OK) return s;
    PairParams pp = make_pp(mScales, nullptr, nullptr, n_scales, 0.0);
    return pairwise_impl(ctx, kind, prm, pp, want_grad, energy, dpos, dbox, dprm);
}

static admp_status admp_slater_compute_impl(admp_ctx* ctx, const double* pos, const double* box, const double* a,
                                           const double* b, const double* mScales, int32_t n_scales, int32_t want_grad,
                                           double* energy, double* dpos, double* dbox, double* da, double* db) {
    if (!ctx || !pos || !box || !a || !b) return ADMP_ERR_ARG;
    { const admp_status ns = nbr_settle(ctx); if (ns != ADMP_OK) return ns; }
    if (!all_frames_have_pairs(ctx)) { ctx->err = "no pair list for every frame"; return ADMP_ERR_STATE; }
    if (n_scales < 1 || n_scales > 8) { ctx->err = "n_scales must be 1..8"; return ADMP_ERR_ARG; }
    admp_status s = upload_geometry(ctx, pos, box);
    if (s != ADMP_OK) return s;
    const int N = ctx->N, F = ctx->F;
    const size_t NF = (size_t)N * F;
    hipStream_t st = ctx->stream;
    PairParams pp = make_pp(mScales, nullptr, nullptr, n_scales, 0.0);
    // (A, B) per site interleaved into t1 (N x 2 <= NF x 3)
    hipLaunchKernelGGL(k_interleave2, dim3(nblk(N, 256)), dim3(256), 0, st, N, a, b, ctx->d_t1);
    double* g2 = want_grad && (da || db) ? ctx->d_t2 : nullptr;  // (F,N,2) dE/d(A,B)
    if ((s = pairwise_impl(ctx, SR_SLATER, ctx->d_t1, pp, want_grad, energy, dpos, dbox, g2)) != ADMP_OK) return s;
    if (g2) {
        if (da) hipLaunchKernelGGL(k_copy_cols, dim3(nblk(NF, 256)), dim3(256), 0, st, NF, 1, 2, 0, g2, da);
        if (db) hipLaunchKernelGGL(k_copy_cols, dim3(nblk(NF, 256)), dim3(256), 0, st, NF, 1, 2, 1, g2, db);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    }
    return ADMP_OK;
}

// ---------------------------------------------------------------------------
// Public entry points: order the context stream after the caller's stream,
// run, then order the caller's stream after the context's work.
// ---------------------------------------------------------------------------
#define ADMP_WRAP(call)                                   \
    do {                                                  \
        if (!ctx) return ADMP_ERR_ARG;                    \
        admp_status s_ = enter(ctx);                      \
        if (s_ != ADMP_OK) return s_;                     \
        admp_status r_ = (call);                          \
        admp_status l_ = leave(ctx);                      \
        return r_ != ADMP_OK ? r_ : l_;                   \
    } while (0)

extern "C" admp_status admp_set_pairs(admp_ctx* ctx, int32_t frame, const void* pairs, int64_t n_pairs,
                                      int32_t pairs_is_int64) {
    ADMP_WRAP(admp_set_pairs_impl(ctx, frame, pairs, n_pairs, pairs_is_int64));
}

extern "C" admp_status admp_build_neighbors(admp_ctx* ctx, const double* pos, const double* box) {
    ADMP_WRAP(admp_build_neighbors_impl(ctx, pos, box));
}

extern "C" admp_status admp_get_pairs(admp_ctx* ctx, int32_t frame, int32_t* out, int64_t cap, int64_t* n_out) {
    ADMP_WRAP(admp_get_pairs_impl(ctx, frame, out, cap, n_out));
}

extern "C" admp_status admp_compute(admp_ctx* ctx, const double* pos, const double* box, const double* Q_local,
                                    const double* pol, const double* thole, const double* mScales,
                                    const double* pScales, const double* dScales, int32_t n_scales, double* U_inout,
                                    const admp_solver_opts* opts, int32_t want_grad, double* energy, double* dpos,
                                    double* dbox, double* dQ_local, double* dpol, double* dthole, admp_stats* stats) {
    ADMP_WRAP(admp_compute_impl(ctx, pos, box, Q_local, pol, thole, mScales, pScales, dScales, n_scales, U_inout, opts,
                                want_grad, energy, dpos, dbox, dQ_local, dpol, dthole, stats));
}

extern "C" admp_status admp_disp_compute(admp_ctx* ctx, const double* pos, const double* box, const double* c_list,
                                         const double* mScales, int32_t n_scales, int32_t want_grad, double* energy,
                                         double* dpos, double* dbox, double* dc) {
    ADMP_WRAP(admp_disp_compute_impl(ctx, pos, box, c_list, mScales, n_scales, want_grad, energy, dpos, dbox, dc));
}

extern "C" admp_status admp_slater_compute(admp_ctx* ctx, const double* pos, const double* box, const double* a,
                                           const double* b, const double* mScales, int32_t n_scales, int32_t want_grad,
                                           double* energy, double* dpos, double* dbox, double* da, double* db) {
    ADMP_WRAP(admp_slater_compute_impl(ctx, pos, box, a, b, mScales, n_scales, want_grad, energy, dpos, dbox, da, db));
}

extern "C" admp_status admp_pairwise_compute(admp_ctx* ctx, int32_t kind, const double* pos, const double* box,
                                             const double* prm, int32_t n_prm, const double* mScales,
                                             int32_t n_scales, int32_t want_grad, double* energy, double* dpos,
                                             double* dbox, double* dprm) {
    ADMP_WRAP(admp_pairwise_compute_impl(ctx, kind, pos, box, prm, n_prm, mScales, n_scales, want_grad, energy, dpos,
                                         dbox, dprm));
}

// Q_local_to_global (dmff/admp/pme.py:298-311): the frames' rotation of every
// frame's per-site local harmonic multipoles to global Cartesian
// [q, mu_x, mu_y, mu_z, Theta_xx, Theta_yy, Theta_zz, Theta_xy, Theta_xz, Theta_yz]
static admp_status admp_local_to_global_impl(admp_ctx* ctx, const double* pos, const double* box,
                                             const double* Q_local, double* Q_global) {
    if (!ctx || !pos || !box || !Q_local || !Q_global) return ADMP_ERR_ARG;
    if (!ctx->has_topo) { ctx->err = "admp_set_topology first"; return ADMP_ERR_STATE; }
    { const admp_status ns = nbr_settle(ctx); if (ns != ADMP_OK) return ns; }
    admp_status s = upload_geometry(ctx, pos, box);
    if (s != ADMP_OK) return s;
    const int N = ctx->N, F = ctx->F;
    hipLaunchKernelGGL(k_frames_fwd, dim3(nblk(N, 128), F), dim3(128), 0, ctx->stream, N, ctx->nh, ctx->use_frames,
                       ctx->d_pos, ctx->d_box, ctx->d_axis_type, ctx->d_axis_idx, Q_local, ctx->d_frames, ctx->d_Qc);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(Q_global, ctx->d_Qc, (size_t)N * F * 10 * sizeof(double), hipMemcpyDeviceToDevice,
                          ctx->stream));
    if ((s = geom_fetch(ctx)) != ADMP_OK) return s;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return geom_check(ctx);
}

extern "C" admp_status admp_local_to_global(admp_ctx* ctx, const double* pos, const double* box,
                                            const double* Q_local, double* Q_global) {
    ADMP_WRAP(admp_local_to_global_impl(ctx, pos, box, Q_local, Q_global));
}

extern "C" admp_status admp_esp(int32_t n_sites, const double* pos, const double* Q_global, const double* U,
                                int32_t n_grid, const double* grids, double* esp, void* stream) {
    if (n_sites < 1 || n_grid < 0 || !pos || !Q_global || (n_grid > 0 && (!grids || !esp))) return ADMP_ERR_ARG;
    if (n_grid == 0) return ADMP_OK;
    hipLaunchKernelGGL(k_esp, dim3(n_grid), dim3(256), 0, (hipStream_t)stream, n_sites, n_grid, pos, Q_global, U,
                       grids, esp);
    return hipGetLastError() == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
}

extern "C" admp_status admp_add_forces_fixed(const double* dE_dpos, double scale, const int32_t* atom_index,
                                             int64_t* force_buffers, int32_t n_atoms, int32_t padded_n_atoms,
                                             void* stream) {
    if (n_atoms < 0 || padded_n_atoms < n_atoms || (n_atoms > 0 && (!dE_dpos || !force_buffers)))
        return ADMP_ERR_ARG;
    if (n_atoms == 0) return ADMP_OK;
    hipLaunchKernelGGL(k_add_forces_fixed, dim3(std::min<unsigned>(nblk(n_atoms, 256), 4096)), dim3(256), 0,
                       (hipStream_t)stream, n_atoms, padded_n_atoms, scale, dE_dpos, atom_index,
                       reinterpret_cast<long long*>(force_buffers));
    return hipGetLastError() == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
}

extern "C" admp_status admp_type_reduce(const double* per_site, int32_t n_frames, int32_t n_sites, int32_t d,
                                        const int32_t* type_map, int32_t n_types, double* out, void* stream) {
    if (n_frames < 1 || n_sites < 0 || d < 1 || n_types < 1 || !type_map || !out || (n_sites > 0 && !per_site))
        return ADMP_ERR_ARG;
    if ((long)n_types * d > 65535 * 64 || n_frames > 65535) return ADMP_ERR_ARG;
    hipLaunchKernelGGL(k_type_reduce, dim3(n_types * d, n_frames), dim3(TR_BLOCK), 0, (hipStream_t)stream, per_site,
                       n_sites, d, type_map, out, n_types);
    return hipGetLastError() == hipSuccess ? ADMP_OK : ADMP_ERR_HIP;
}

extern "C" admp_status admp_memory_bytes(const admp_ctx* ctx, int64_t* bytes) {
    if (!ctx || !bytes) return ADMP_ERR_ARG;
    int64_t t = 0;
    for (size_t b : ctx->alloc_bytes) t += (int64_t)b;
    *bytes = t;
    return ADMP_OK;
}

// ---------------------------------------------------------------------------
// admp_pairwise_compute_noalloc: the pair terms of admp_pairwise_compute for
// callers that cannot allocate (code under HIP graph capture, serving loops).
// It takes the reference-format pair list directly (pairwise.py:47-95), issues
// four launches on the caller's stream and nothing else: no allocation, no
// memset, no host synchronisation.  Its only device memory besides the caller's
// inputs and outputs is `scratch` below: the box record and the per-block
// partial sums of the energy and the virial.  Its size does not depend on the
// input (the pair kernel is a grid-stride loop over at most NA_MAXBLK blocks).
// Every call on a device shares that device's copy, so calls that overlap on
// one device race; each call rewrites what it reads.
// ---------------------------------------------------------------------------
namespace {

constexpr int NA_BLOCK = 256;
constexpr int NA_MAXBLK = 1024;
constexpr int NA_BOX_DOUBLES = 64;                         // BoxInfo + the flag word
constexpr int NA_SCRATCH_DOUBLES = NA_BOX_DOUBLES + NA_MAXBLK * 10;  // 82,432 bytes
static_assert(sizeof(BoxInfo) + sizeof(int) <= NA_BOX_DOUBLES * sizeof(double), "box record");
static_assert(PC_ROW == 3 * PC_MAXM, "preconditioner rows hold one 3x3 block per cluster site");
constexpr int NA_MAX_DEVICES = 64;

}  // namespace

static __device__ __align__(256) double scratch[NA_SCRATCH_DOUBLES];

namespace {

thread_local std::string g_err;          // last error of a context-free entry point
std::mutex g_scratch_mu;
double* g_scratch_ptr[NA_MAX_DEVICES] = {nullptr};

// The device address of the current device's `scratch` (looked up once per
// device; the array's name in host code is not a device address).
admp_status scratch_ptr(double** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { g_err = "hipGetDevice failed"; return ADMP_ERR_HIP; }
    if (dev < 0 || dev >= NA_MAX_DEVICES) { g_err = "device ordinal beyond 64"; return ADMP_ERR_ARG; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            g_err = "hipGetSymbolAddress(scratch) failed";
            return ADMP_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return ADMP_OK;
}

}  // namespace

static const char* context_free_error() { return g_err.empty() ? "null context" : g_err.c_str(); }

namespace admp {

// zero the gradient outputs and the flag word, then the box record (one thread)
__global__ void k_na_prep(int N, int nprm, const double* __restrict__ box, double* __restrict__ sbox,
                          double* __restrict__ dpos, double* __restrict__ dprm) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nt = (size_t)gridDim.x * blockDim.x;
    if (dpos)
        for (size_t k = tid; k < (size_t)N * 3; k += nt) dpos[k] = 0.0;
    if (dprm)
        for (size_t k = tid; k < (size_t)N * nprm; k += nt) dprm[k] = 0.0;
    if (tid == 0) {
        int* flags = reinterpret_cast<int*>(sbox + NA_BOX_DOUBLES - 1);
        *flags = 0;
        box_setup_one(box, 0, 0, 0, 0, 0, 0, reinterpret_cast<BoxInfo*>(sbox), flags);
    }
}

// One thread per pair row (grid-stride).  Rows that are not live pairs
// (i >= j, the (N, N) padding rows, indices outside [0, N)) contribute nothing,
// as regularize_pairs / pair_buffer_scales make them (dmff/utils.py:65-84).
// Forces and parameter gradients go to the outputs with atomic adds; the energy
// and the virial to per-block partials in scratch.
template <class R, int KIND, class I>
__global__ void __launch_bounds__(NA_BLOCK)
k_na_pairs(int64_t np, int N, const I* __restrict__ pairs, const double* __restrict__ pos,
           const double* __restrict__ prm, const double* __restrict__ sbox, PairParams pp, int want_grad,
           double* __restrict__ part, double* __restrict__ dpos, double* __restrict__ dprm) {
    __shared__ double sh[16];
    const BoxInfo& bx = *reinterpret_cast<const BoxInfo*>(sbox);
    constexpr int npr = sr_nprm(KIND);
    double E = 0, v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t r = (int64_t)blockIdx.x * NA_BLOCK + threadIdx.x; r < np; r += (int64_t)gridDim.x * NA_BLOCK) {
        const int64_t i = (int64_t)pairs[r * 3], j = (int64_t)pairs[r * 3 + 1];
        if (!(i >= 0 && i < j && j < N)) continue;
        const int64_t nb64 = (int64_t)pairs[r * 3 + 2];
        const int nb = (int)(nb64 < -64 ? -64 : (nb64 > 64 ? 64 : nb64));
        double dx = pos[i * 3] - pos[j * 3], dy = pos[i * 3 + 1] - pos[j * 3 + 1], dz = pos[i * 3 + 2] - pos[j * 3 + 2];
        double n0, n1, n2;
        pbc_shift(bx, dx, dy, dz, n0, n1, n2);
        const double rd = sqrt(dx * dx + dy * dy + dz * dz);
        const R rr = (R)rd, rinv = (R)(1.0 / rd);
        const R m = (R)pp.ms[scale_index(nb, pp.nscale)];
        R pI[SR_NPRM_MAX] = {0, 0, 0, 0}, pJ[SR_NPRM_MAX] = {0, 0, 0, 0}, dI[SR_NPRM_MAX], dJ[SR_NPRM_MAX];
        for (int q = 0; q < npr; ++q) { pI[q] = (R)prm[i * npr + q]; pJ[q] = (R)prm[j * npr + q]; }
        R e, dedr;
        sr_pair<R, KIND>(rr, m, pI, pJ, e, dedr, dI, dJ);
        E += (double)e;
        if (want_grad) {
            const double f[3] = {(double)(dedr * (R)dx * rinv), (double)(dedr * (R)dy * rinv),
                                 (double)(dedr * (R)dz * rinv)};
            const double nn[3] = {n0, n1, n2};
            for (int c = 0; c < 3; ++c) {
                atomicAdd(&dpos[i * 3 + c], f[c]);
                atomicAdd(&dpos[j * 3 + c], -f[c]);
            }
            for (int q = 0; q < npr; ++q) {
                atomicAdd(&dprm[i * npr + q], (double)dI[q]);
                atomicAdd(&dprm[j * npr + q], (double)dJ[q]);
            }
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) v[p * 3 + q] -= nn[p] * f[q];
        }
    }
    const double Eb = block_sum(E, sh);
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * 10] = Eb;
    for (int k = 0; k < 9; ++k) {
        const double vb = block_sum(v[k], sh);
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * 10 + 1 + k] = vb;
    }
}

// energy and dE/dbox from the block partials, summed in block order; NaN when
// the box was unusable (box_usable), since nothing here can return an error
__global__ void k_na_finish(int nblk, const double* __restrict__ sbox, const double* __restrict__ part,
                            double* __restrict__ energy, double* __restrict__ dbox) {
    const int k = threadIdx.x;
    if (k >= 10) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(size_t)b * 10 + k];
    const int flags = *reinterpret_cast<const int*>(sbox + NA_BOX_DOUBLES - 1);
    if (flags) s = __builtin_nan("");
    if (k == 0) {
        if (energy) energy[0] = s;
    } else if (dbox) {
        dbox[k - 1] = s;
    }
}

}  // namespace admp

template <class R, int KIND>
static void na_launch(int nb, hipStream_t st, int64_t np, int N, const void* pairs, int is64, const double* pos,
                      const double* prm, const double* sbox, const PairParams& pp, int want_grad, double* part,
                      double* dpos, double* dprm) {
    if (is64)
        hipLaunchKernelGGL((k_na_pairs<R, KIND, int64_t>), dim3(nb), dim3(NA_BLOCK), 0, st, np, N,
                           (const int64_t*)pairs, pos, prm, sbox, pp, want_grad, part, dpos, dprm);
    else
        hipLaunchKernelGGL((k_na_pairs<R, KIND, int32_t>), dim3(nb), dim3(NA_BLOCK), 0, st, np, N,
                           (const int32_t*)pairs, pos, prm, sbox, pp, want_grad, part, dpos, dprm);
}

template <class R>
static void na_dispatch(int kind, int nb, hipStream_t st, int64_t np, int N, const void* pairs, int is64,
                        const double* pos, const double* prm, const double* sbox, const PairParams& pp,
                        int want_grad, double* part, double* dpos, double* dprm) {
    switch (kind) {
        case SR_SLATER: na_launch<R, SR_SLATER>(nb, st, np, N, pairs, is64, pos, prm, sbox, pp, want_grad, part, dpos, dprm); break;
        case SR_TT_QQ_C6: na_launch<R, SR_TT_QQ_C6>(nb, st, np, N, pairs, is64, pos, prm, sbox, pp, want_grad, part, dpos, dprm); break;
        case SR_TT_QQ: na_launch<R, SR_TT_QQ>(nb, st, np, N, pairs, is64, pos, prm, sbox, pp, want_grad, part, dpos, dprm); break;
        default: na_launch<R, SR_SLATER_DAMP>(nb, st, np, N, pairs, is64, pos, prm, sbox, pp, want_grad, part, dpos, dprm); break;
    }
}

extern "C" admp_status admp_pairwise_compute_noalloc(int32_t kind, int32_t precision, int32_t n_atoms,
                                                     const double* pos, const double* box, const void* pairs,
                                                     int64_t n_pairs, int32_t pairs_is_int64, const double* prm,
                                                     int32_t n_prm, const double* mScales, int32_t n_scales,
                                                     int32_t want_grad, double* energy, double* dpos, double* dbox,
                                                     double* dprm, void* stream) {
    if (kind < 1 || kind > SR_KIND_MAX) { g_err = "unknown short-range kind"; return ADMP_ERR_ARG; }
    if (n_prm != sr_nprm(kind)) { g_err = "n_prm does not match the kind's parameter count"; return ADMP_ERR_ARG; }
    if (precision != ADMP_FP64 && precision != ADMP_MIXED) { g_err = "precision must be ADMP_FP64 or ADMP_MIXED"; return ADMP_ERR_ARG; }
    if (n_atoms < 1 || n_atoms >= (1 << NB_SHIFT)) { g_err = "n_atoms must be in 1 .. 2^28 - 1"; return ADMP_ERR_ARG; }
    if (n_pairs < 0 || n_pairs > ((int64_t)1 << 40)) { g_err = "n_pairs must be in 0 .. 2^40"; return ADMP_ERR_ARG; }
    if (n_scales < 1 || n_scales > 8 || !mScales) { g_err = "n_scales must be 1..8"; return ADMP_ERR_ARG; }
    if (!pos || !box || !prm || (n_pairs > 0 && !pairs)) { g_err = "null input"; return ADMP_ERR_ARG; }
    if (want_grad && (!dpos || !dprm)) { g_err = "want_grad needs dpos and dprm"; return ADMP_ERR_ARG; }
    double* sc = nullptr;
    admp_status s = scratch_ptr(&sc);
    if (s != ADMP_OK) return s;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    double* sbox = sc;
    double* part = sc + NA_BOX_DOUBLES;
    const PairParams pp = make_pp(mScales, nullptr, nullptr, n_scales, 0.0);
    double* gp = want_grad ? dpos : nullptr;
    double* gq = want_grad ? dprm : nullptr;
    const int64_t want_blocks = (n_pairs + NA_BLOCK - 1) / NA_BLOCK;
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(want_blocks, NA_MAXBLK));
    const size_t nz = (size_t)n_atoms * std::max(3, (int)n_prm);
    hipLaunchKernelGGL(k_na_prep, dim3((unsigned)std::min<size_t>(nblk(nz, 256), 1024)), dim3(256), 0, st, n_atoms,
                       n_prm, box, sbox, gp, gq);
    if (precision == ADMP_MIXED)
        na_dispatch<float>(kind, nb, st, n_pairs, n_atoms, pairs, pairs_is_int64, pos, prm, sbox, pp, want_grad, part,
                           gp, gq);
    else
        na_dispatch<double>(kind, nb, st, n_pairs, n_atoms, pairs, pairs_is_int64, pos, prm, sbox, pp, want_grad, part,
                            gp, gq);
    hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(64), 0, st, nb, sbox, part, energy, want_grad ? dbox : nullptr);
    if (hipGetLastError() != hipSuccess) { g_err = "kernel launch failed"; return ADMP_ERR_HIP; }
    return ADMP_OK;
}
