// Host runtime of the gfx950 QSM library: planning, device workspaces, the
// iteration loop with the reference's progress semantics, and the extern "C"
// entry points declared in include/qsm_gpu.h.
#include "qsm_internal.hpp"
#include "fft_col2.hpp"
#include "fft_chirp.hpp"
#include "../../include/qsm_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace qsm {

void launch_tv_output(hipStream_t st, const VolDesc* d, const LaunchTable& lt, const VolStatus* status);
void launch_fill_zero(hipStream_t st, float* p, size_t n);
void launch_status_reset(hipStream_t st, VolStatus* s, int nvol);
void launch_vs_init(hipStream_t st, const VolDesc* d, const LaunchTable& lt);
void launch_tv_lbar(hipStream_t st, const VolDesc* d, int nvol, int nrows);
void launch_nltv_weights(hipStream_t st, const VolDesc* d, const LaunchTable& lt, int parity, int sel);
void launch_nltv_wmax(hipStream_t st, const VolDesc* d, int nvol, int nblk, int sel);
void launch_nltv_init(hipStream_t st, const VolDesc* d, const LaunchTable& lt);
void launch_status_clear_done(hipStream_t st, VolStatus* s, int nvol);
void launch_copy_probe(hipStream_t st, const float2* a, float2* b, size_t n);
int launch_stream_probe(hipStream_t st, int reads, int writes, const float* in, float* out, size_t n);
bool fixed_length(int n);

// ------------------------------------------------------------ error state --
thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define QSM_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(e_ == hipErrorOutOfMemory ? QSM_ENOMEM : QSM_EHIP, "%s: %s (%s:%d)", \
                        #call, hipGetErrorString(e_), __FILE__, __LINE__);            \
    } while (0)

#define QSM_CHECK_LAUNCH()                                                             \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess)                                                          \
            return fail(QSM_EHIP, "kernel launch: %s (%s:%d)", hipGetErrorString(e_),  \
                        __FILE__, __LINE__);                                           \
    } while (0)

// -------------------------------------------------------------- planning --
constexpr int kMaxLen = qsmfft::kMaxDirectLen;
using qsmfft::direct_length;
using qsmfft::chirp_length;
using qsmfft::chirp_tables;
using qsmfft::kMaxChirpLen;
bool length_ok(size_t n) {
    if (n < 1 || n > (size_t)kMaxLen) return false;
    return direct_length((int)n) || ((int)n <= kMaxChirpLen && chirp_length((int)n) > 0);
}

// Length the FFT stages of an axis of length n run at (n, or the chirp-z
// convolution length), which sizes the LDS tiles.
int work_length(int n) { return direct_length(n) ? n : chirp_length(n); }

// Columns per tile of a column pass of length n: kTile for the compiled
// lengths and for runtime plans of work length <= 256; kTileLong above (half
// the LDS per block of the two-buffer transform, so twice the blocks per
// CU).  Measured, one volume: 164x205x205 (chirp-z on every axis) 4.29 ->
// 5.14 solves/s and 232x232x160 4.30 -> 4.69 with kTileLong, 130^3 (work
// length 130) 27.8 -> 26.7, so kTile stays there.
int col_tile(int n) {
    if (has_fixed_kernels(n) || work_length(n) <= 256) return kTile;
    return kTileLong;
}

// Runtime-plan column kernel of length n: its elements per thread (EG),
// plus kColLongTile when it runs in tiles of kTileLong columns.
int emax_for_col(int n) {
    if (col_tile(n) == kTileLong) return kColLongTile + (work_length(n) <= 512 ? 16 : 32);
    return 16;
}

// Elements of the tile-major TV / RTS z-pass operands (f^, q) of a volume:
// ny * nz rows of whole z-pass tiles (at least the half spectrum).
size_t zop_elems(const VolDesc& V) {
    const int tw = col_tile(V.nz);
    const size_t cols = (size_t)(tw == kTile ? V.ntx : V.ntx_long) * tw;
    return cols * V.ny * V.nz;
}

// exp(-2 pi i q/n) from an exactly reduced angle, rounded once
inline float2 twiddle_entry(int q, int n) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * q / n;
    return make_float2((float)cosl(a), (float)sinl(a));
}

struct TwiddleBank {
    std::vector<float2> host;
    std::map<int, int> off, chirp_off, filt_off;
    int add(int n) {
        auto it = off.find(n);
        if (it != off.end()) return it->second;
        int o = (int)host.size();
        for (int q = 0; q < n; ++q) host.push_back(twiddle_entry(q, n));
        off[n] = o;
        return o;
    }
    // chirp (n entries) and filter (M entries) offsets of length n
    void add_chirp(int n, int M, int& oc, int& of) {
        auto it = chirp_off.find(n);
        if (it != chirp_off.end()) {
            oc = it->second;
            of = filt_off[n];
            return;
        }
        std::vector<float2> c, f;
        chirp_tables(n, M, c, f);
        oc = (int)host.size();
        host.insert(host.end(), c.begin(), c.end());
        of = (int)host.size();
        host.insert(host.end(), f.begin(), f.end());
        chirp_off[n] = oc;
        filt_off[n] = of;
    }
};

// The same tables for one volume (three axes, each at most a length-1024
// table, or a length-M table plus a chirp of n <= 512 and a filter of M)
// in fixed storage, for the entry point that allocates nothing.
struct FixedTwiddleBank {
    static constexpr int kCap = 3 * (2 * kMaxLen + kMaxChirpLen);
    float2 host[kCap];
    int len[6] = {0, 0, 0, 0, 0, 0}, off[6] = {0, 0, 0, 0, 0, 0};
    int nlen = 0, count = 0;
    int add(int n) {
        for (int i = 0; i < nlen; ++i)
            if (len[i] == n) return off[i];
        const int o = count;
        for (int q = 0; q < n; ++q) host[count++] = twiddle_entry(q, n);
        len[nlen] = n;
        off[nlen++] = o;
        return o;
    }
    void add_chirp(int n, int M, int& oc, int& of) {
        for (int i = 0; i < nlen; ++i)
            if (len[i] == -n) {  // chirp tables are keyed by -n
                oc = off[i];
                of = oc + n;
                return;
            }
        std::vector<float2> c, f;
        chirp_tables(n, M, c, f);
        oc = count;
        for (const float2& v : c) host[count++] = v;
        of = count;
        for (const float2& v : f) host[count++] = v;
        len[nlen] = -n;
        off[nlen++] = oc;
    }
};

template <class TB>
qsmfft::Plan make_plan(int n, TB& tb) {
    qsmfft::Plan p{};
    if (direct_length(n)) {
        qsmfft::make_plan_stages(n, p);
        p.tw_off = tb.add(n);
        return p;
    }
    // chirp-z: the stages run at the convolution length
    const int M = chirp_length(n);
    qsmfft::make_plan_stages(M, p);
    p.tw_off = tb.add(M);
    p.bl_n = n;
    tb.add_chirp(n, M, p.bl_chirp, p.bl_filt);
    return p;
}

template <class T>
int dev_alloc(T** p, size_t count) {
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? QSM_ENOMEM : QSM_EHIP, "hipMalloc(%zu bytes): %s",
                    count * sizeof(T), hipGetErrorString(e));
    return QSM_OK;
}

int validate_grid(const qsm_grid* g) {
    if (!g) return fail(QSM_EPARAM, "grid is null");
    if (g->nx == 0 || g->ny == 0 || g->nz == 0) return fail(QSM_EPARAM, "grid has a zero dimension");
    if (!(g->vsx > 0) || !(g->vsy > 0) || !(g->vsz > 0))
        return fail(QSM_EPARAM, "voxel sizes must be positive");
    const size_t dims[3] = {g->nx, g->ny, g->nz};
    for (size_t d : dims)
        if (!length_ok(d))
            return fail(QSM_EUNSUPPORTED,
                        "axis length %zu not supported by the GPU FFT (2^a3^b5^c7^d11^e13^f <= %d, "
                        "or any length <= %d)",
                        d, kMaxLen, kMaxChirpLen);
    if (g->nx * g->ny * g->nz > (size_t)1 << 29)
        return fail(QSM_EUNSUPPORTED, "volume of %zu voxels exceeds 2^29", g->nx * g->ny * g->nz);
    return QSM_OK;
}

// Fill the geometry part of a descriptor and the block counts of each pass.
struct PassBlocks {
    int xpass, cpass, ypass, zpass, out;
};

// Pair-rows per block of the x passes: as many as 16 elements per thread
// hold (kRowEmaxRt for the runtime-plan kernels), at most 32.
static int default_rows_x(int nx) {
    const int tile = qsmfft::kThreads * (has_fixed_kernels(nx) ? 16 : kRowEmaxRt);
    return std::max(1, std::min(32, tile / work_length(nx)));
}

// xshift / cshift 2 (compiled lengths only): the R2C / TV C2R passes take a
// quarter of their default rows per block, as the kernels launched with
// PassArgs::xshift = 2 do (C2R: half the R2C pass's rows before its shift).
template <class TB>
PassBlocks fill_geometry(VolDesc& V, const qsm_grid& g, TB& tb, int xshift = 0, int cshift = 0) {
    V.nx = (int)g.nx; V.ny = (int)g.ny; V.nz = (int)g.nz;
    V.nxh = (V.nx + 1) / 2;  // even nx: DC and Nyquist share column 0
    V.rows_x = default_rows_x(V.nx) >> xshift;
    // compiled C2R kernels take half the R2C pass's pair-rows (the previous
    // iterate stays in registers across the transform); the runtime-plan one
    // its whole tile, whose 8 elements per thread cover it
    V.rows_c = has_fixed_kernels(V.nx) ? std::max(1, (default_rows_x(V.nx) >> cshift) / 2)
                                       : default_rows_x(V.nx);
    V.ntx = (V.nxh + kTile - 1) / kTile;
    V.ntx_long = (V.nxh + kTileLong - 1) / kTileLong;
    V.m_nx = qsmfft::magic_for((uint32_t)V.nx);
    V.m_nxh = qsmfft::magic_for((uint32_t)V.nxh);
    V.m_ny = qsmfft::magic_for((uint32_t)V.ny);
    V.px = make_plan(V.nx, tb);
    V.py = make_plan(V.ny, tb);
    V.pz = make_plan(V.nz, tb);
    V.vsx = g.vsx; V.vsy = g.vsy; V.vsz = g.vsz;
    V.hx = (float)(1.0 / g.vsx); V.hy = (float)(1.0 / g.vsy); V.hz = (float)(1.0 / g.vsz);
    PassBlocks pb;
    const int npairs = (V.ny * V.nz + 1) / 2;
    pb.xpass = (npairs + V.rows_x - 1) / V.rows_x;
    pb.cpass = (npairs + V.rows_c - 1) / V.rows_c;
    pb.ypass = (col_tile(V.ny) == kTile ? V.ntx : V.ntx_long) * V.nz;
    pb.zpass = (col_tile(V.nz) == kTile ? V.ntx : V.ntx_long) * V.ny;
    const size_t N = (size_t)V.nx * V.ny * V.nz;
    pb.out = (int)((N + 1023) / 1024);
    return pb;
}

// Volumes of a batch grouped by the length of one pass axis: one launch per
// group, so each group runs the kernel compiled for its length.
struct PassGroup {
    int n;
    int emax;
    LaunchTable lt;
    int xshift = 0;  // x passes: row shift of every volume in the group
};
using Groups = std::vector<PassGroup>;

void launch_tv_reduce(hipStream_t st, const VolDesc* d, int nvol, VolStatus* status, VolStatus* mirror);

// shift (x passes): each volume's row shift; volumes of one length with
// different shifts run in separate groups.
Groups make_groups(const std::vector<int>& len, const std::vector<int>& blocks, bool column,
                   const std::vector<int>* shift = nullptr) {
    Groups gs;
    for (size_t v = 0; v < len.size(); ++v) {
        const int sh = shift ? (*shift)[v] : 0;
        PassGroup* g = nullptr;
        for (auto& x : gs)
            if (x.n == len[v] && x.xshift == sh) g = &x;
        if (!g) {
            gs.push_back(PassGroup{});
            g = &gs.back();
            g->n = len[v];
            g->emax = column ? emax_for_col(len[v]) : 16;
            g->xshift = sh;
            g->lt.nvol = 0;
            g->lt.start[0] = 0;
        }
        LaunchTable& lt = g->lt;
        lt.vol[lt.nvol] = (int)v;
        lt.start[lt.nvol + 1] = lt.start[lt.nvol] + blocks[v];
        ++lt.nvol;
    }
    return gs;
}

// ------------------------------------------------- column-pass schedules --
// Per (pass, axis length): 0 two-stage, 1 LDS-staged.  Pass 0: the y passes;
// pass 1: the per-iteration z passes.  Mode 0 (default): a fixed table, so
// every process computes the same bits; 1 always two-stage; 2 always staged;
// 3 tuned: the first batch using a length times both forms and the faster
// one is kept for the process (-1 in the table until then).
namespace {
constexpr int kMaxAxis = 512;
std::mutex g_sched_mu;
int8_t g_sched_tab[2][kMaxAxis + 1];
std::atomic<int> g_sched_mode{0};
struct SchedInit {
    SchedInit() { std::memset(g_sched_tab, -1, sizeof g_sched_tab); }
} g_sched_init;

int sched_pass(int kind) {
    if (kind == P_Y_FWD || kind == P_Y_INV) return 0;
    if (kind == P_Z_TV_SOLVE || kind == P_Z_TV_SOLVE0 || kind == P_Z_TV_ADJ || kind == P_Z_RTS_SOLVE) return 1;
    return -1;
}

// The fixed table: the forms the tuner chose on MI355X over the TV workloads
// (256x256x160, the uneven 32-volume batch): the staged form for y passes of
// length 144 and 176, where radices 9 and 11 in registers cost more than
// LDS stages, and the two-stage form everywhere else.
int fixed_schedule(int pass, int n) { return (pass == 0 && (n == 144 || n == 176)) ? 1 : 0; }
}  // namespace

int column_schedule(int kind, int n) {
    const int pass = sched_pass(kind);
    if (pass < 0 || n < 0 || n > kMaxAxis || !qsmfft::col2_ok(n)) return 1;
    const int mode = g_sched_mode.load();
    if (mode == 1) return 0;
    if (mode == 2) return 1;
    if (mode == 0) return fixed_schedule(pass, n);
    std::lock_guard<std::mutex> lk(g_sched_mu);
    const int t = g_sched_tab[pass][n];
    return t < 0 ? fixed_schedule(pass, n) : t;
}

void run_groups(int kind, const Groups& gs, const PassArgs& a) {
    PassArgs b = a;
    for (const auto& g : gs) {
        b.sched = column_schedule(kind, g.n);
        b.xshift = g.xshift;
        launch_pass(kind, g.n, g.emax, g.lt, b);
    }
}

LaunchTable identity_table(const std::vector<int>& blocks) {
    LaunchTable lt{};
    lt.nvol = (int)blocks.size();
    lt.start[0] = 0;
    for (size_t v = 0; v < blocks.size(); ++v) {
        lt.vol[v] = (int)v;
        lt.start[v + 1] = lt.start[v] + blocks[v];
    }
    return lt;
}

}  // namespace qsm

using namespace qsm;

// ============================================================ TV batch ====
struct qsm_tv_batch {
    int nvol = 0;
    int method = 0;  // 0 TV-ADMM (src/inversion/tv.rs), 1 RTS (src/inversion/rts.rs)
    qsm_tv_params params{};
    std::vector<VolDesc> h;
    VolDesc* d_vols = nullptr;
    VolStatus* d_status = nullptr;
    VolStatus* h_status = nullptr;   // pinned mirror written by the reduce kernel
    VolStatus* h_status_dev = nullptr;
    float2* d_tw = nullptr;
    Groups gx, gxc, gy, gz;         // x / TV C2R / y / z passes grouped by axis length
    LaunchTable lt_out{};
    std::vector<void*> allocs;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    static constexpr int kWindow = 4;
    PassArgs args(const VolStatus* status) const {
        PassArgs a{};
        a.stream = stream;
        a.vols = d_vols;
        a.tw = d_tw;
        a.status = status;
        a.status_rw = nullptr;
        a.mirror = nullptr;
        a.sel = SPEC_WORK;
        a.sel2 = -1;
        return a;
    }
    hipEvent_t ev[kWindow] = {};
    // Iterations it >= 2 replay captured graphs: graph p is one iteration with
    // (it - 1) & 1 == p (its kernel arguments depend only on the parity).
    // graph_state: 0 not built yet, 1 built, -1 eager launches (disabled, or
    // the stream cannot be captured).
    hipGraphExec_t gexec[2] = {};
    int graph_state = 0;
    // Passes with several launch groups (volumes of several axis lengths)
    // fork them over the batch's stream and up to kSide side streams and
    // join before the next pass, so the groups' last waves overlap.  Created
    // only for batches with more than one group in some pass.
    static constexpr int kSide = 3;
    hipStream_t side[kSide] = {};
    hipEvent_t ev_fork[5] = {};
    hipEvent_t ev_join[5][kSide] = {};
};

static void batch_free(qsm_tv_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto& g : b->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    for (auto& st : b->side)
        if (st) (void)hipStreamSynchronize(st);
    for (void* p : b->allocs) (void)hipFree(p);
    for (auto& e : b->ev_fork)
        if (e) (void)hipEventDestroy(e);
    for (auto& row : b->ev_join)
        for (auto& e : row)
            if (e) (void)hipEventDestroy(e);
    for (auto& st : b->side)
        if (st) (void)hipStreamDestroy(st);
    if (b->h_status) (void)hipHostFree(b->h_status);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// The autotuner (schedule mode 3): for every y and z length of a new batch
// that has no entry yet, time the forward y pass and the z solve pass under
// both schedules on the batch's own launch tables and buffers (before any
// run, so only scratch is overwritten, after clearing it so that no
// uninitialised values enter the timed launches; convergence flags are not
// consulted) and keep the faster one for the rest of the process.  One
// warm-up and the best of five timed launches each.  The two forms round
// differently, so tuned results can differ between processes in the last
// bits; the default mode uses the fixed table instead.
static int tune_column_schedules(qsm_tv_batch* B) {
    if (g_sched_mode.load() != 3) return QSM_OK;
    {
        // a stream under capture cannot be timed: keep the defaults
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(B->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
            (void)hipGetLastError();
            return QSM_OK;
        }
    }
    for (const auto& V : B->h) {
        launch_fill_zero(B->stream, reinterpret_cast<float*>(V.spec[SPEC_WORK]), 2 * (size_t)V.nxh * V.ny * V.nz);
        launch_fill_zero(B->stream, reinterpret_cast<float*>(V.spec[SPEC_A]), 2 * zop_elems(V));
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto best_ms = [&](int kind, const PassGroup& g, int sched, float& best) {
        PassArgs a = B->args(nullptr);
        a.sched = sched;
        launch_pass(kind, g.n, g.emax, g.lt, a);
        best = 1e30f;
        for (int r = 0; r < 5 && e == hipSuccess; ++r) {
            float ms = 0.f;
            if ((e = hipEventRecord(e0, B->stream)) != hipSuccess) break;
            launch_pass(kind, g.n, g.emax, g.lt, a);
            if ((e = hipGetLastError()) != hipSuccess) break;
            if ((e = hipEventRecord(e1, B->stream)) != hipSuccess) break;
            if ((e = hipEventSynchronize(e1)) != hipSuccess) break;
            if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
            best = std::min(best, ms);
        }
    };
    const struct { const Groups* gs; int kind; int pass; } work[2] = {
        {&B->gy, P_Y_FWD, 0}, {&B->gz, P_Z_TV_SOLVE, 1}};
    for (const auto& w : work) {
        for (const auto& g : *w.gs) {
            if (e != hipSuccess) break;
            if (g.n > kMaxAxis || !qsmfft::col2_ok(g.n)) continue;
            {
                std::lock_guard<std::mutex> lk(g_sched_mu);
                if (g_sched_tab[w.pass][g.n] >= 0) continue;
            }
            float t0 = 0.f, t1 = 0.f;
            best_ms(w.kind, g, 0, t0);
            best_ms(w.kind, g, 1, t1);
            if (e != hipSuccess) break;
            // the staged form only when clearly faster: measured at creation,
            // with the batch's data in cache, the two come closer than in the
            // iteration (where the two-stage y pass leads by ~9 % at 256)
            std::lock_guard<std::mutex> lk(g_sched_mu);
            g_sched_tab[w.pass][g.n] = (int8_t)(t1 < 0.97f * t0 ? 1 : 0);
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(QSM_EHIP, "schedule tuning: %s", hipGetErrorString(e));
    return QSM_OK;
}

// rts != null selects RTS: its rho, tol, max_iter replace params' (params
// may then be null).
static int tv_batch_create_impl(const qsm_tv_volume* vols, size_t n_vols, const qsm_tv_params* params_in,
                                void* stream, qsm_tv_batch** out, const qsm_rts_params* rts = nullptr) {
    qsm_tv_params rts_as_tv{};
    if (rts) {
        if (!(rts->rho > 0) || std::isnan(rts->delta) || std::isnan(rts->mu))
            return fail(QSM_EPARAM, "RTS needs rho > 0 and finite delta, mu");
        rts_as_tv.lambda = 1.0;  // threshold 1/rho is set per volume below
        rts_as_tv.rho = rts->rho;
        rts_as_tv.tol = rts->tol;
        rts_as_tv.max_iter = rts->max_iter;
        params_in = &rts_as_tv;
    }
    const qsm_tv_params* params = params_in;
    if (!vols || !params || !out) return fail(QSM_EPARAM, "null argument");
    if (n_vols == 0 || n_vols > (size_t)kMaxVolumes)
        return fail(QSM_EPARAM, "batch size %zu outside 1..%d", n_vols, kMaxVolumes);
    if (!(params->rho != 0.0) || std::isnan(params->rho) || std::isnan(params->lambda))
        return fail(QSM_EPARAM, "rho must be non-zero and lambda/rho not NaN");
    for (size_t v = 0; v < n_vols; ++v) {
        int rc = validate_grid(&vols[v].grid);
        if (rc) return rc;
        if (!vols[v].d_field || !vols[v].d_mask || !vols[v].d_chi)
            return fail(QSM_EPARAM, "volume %zu has a null device pointer", v);
        const double* b = vols[v].bdir;
        double bn = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        if (!(bn > 0)) return fail(QSM_EPARAM, "volume %zu: B0 direction has zero length", v);
    }
    qsm_tv_batch* B = new qsm_tv_batch();
    B->nvol = (int)n_vols;
    B->method = rts ? 1 : 0;
    B->params = *params;
    B->h.resize(n_vols);
    TwiddleBank tb;
    std::vector<int> lx, ly, lz, bx, bc, by, bz, bo;
    auto cleanup = [&](int rc) { batch_free(B); return rc; };
    // A small volume of a compiled x length, whose R2C pass would launch
    // fewer than 1024 blocks (128 for 128x128x64 against ~768 resident), runs
    // that pass with a quarter of the rows per block, and its TV C2R pass
    // likewise, where either keeps at least 4 pair-rows per block (fewer made
    // the C2R pass slower, measured).  The choice depends on the volume's own
    // shape only: the two forms round differently in the last bits, and a
    // volume must give the same bits alone and in any batch.
    std::vector<int> xshift(n_vols), cshift(n_vols);
    for (size_t v = 0; v < n_vols; ++v) {
        const int nx = (int)vols[v].grid.nx, r0 = default_rows_x(nx), rc = std::max(1, r0 / 2);
        const long npairs = ((long)vols[v].grid.ny * (long)vols[v].grid.nz + 1) / 2;
        const bool fixed = has_fixed_kernels(nx);
        xshift[v] = (fixed && (npairs + r0 - 1) / r0 < 1024 && (r0 >> 2) >= 4) ? 2 : 0;
        cshift[v] = (fixed && (npairs + rc - 1) / rc < 1024 && (r0 >> 2) / 2 >= 4) ? 2 : 0;
    }
    for (size_t v = 0; v < n_vols; ++v) {
        VolDesc& V = B->h[v];
        std::memset(&V, 0, sizeof V);
        PassBlocks pb = fill_geometry(V, vols[v].grid, tb, xshift[v], cshift[v]);
        const double* b = vols[v].bdir;
        double bn = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        V.bx = b[0] / bn; V.by = b[1] / bn; V.bz = b[2] / bn;  // src/kernels/dipole.rs:39-43
        V.rho = params->rho;
        V.tol = params->tol;
        V.thr = (float)(params->lambda / params->rho);          // src/inversion/tv.rs:78
        if (rts) {
            V.thr = (float)(1.0 / rts->rho);                     // src/inversion/rts.rs:191-201
            V.filt = 5;                                          // FILT_RTS_X0
            V.filt_a = rts->delta;
            V.filt_b = rts->mu;
            V.lsmr_iter = (int)rts->lsmr_iter;
        }
        V.field = vols[v].d_field;
        V.mask = vols[v].d_mask;
        V.out = vols[v].d_chi;
        lx.push_back(V.nx); bx.push_back(pb.xpass); bc.push_back(pb.cpass);
        ly.push_back(V.ny); by.push_back(pb.ypass);
        lz.push_back(V.nz); bz.push_back(pb.zpass);
        bo.push_back(pb.out);
        V.npartial = pb.cpass;
        const size_t N = (size_t)V.nx * V.ny * V.nz;
        const size_t Nh = (size_t)V.nxh * V.ny * V.nz;
        int rc;
        for (int p = 0; p < 2; ++p) {
            if ((rc = dev_alloc(&V.x[p], N))) return cleanup(rc);
            B->allocs.push_back(V.x[p]);
            if ((rc = dev_alloc(&V.u[p], 3 * N))) return cleanup(rc);
            B->allocs.push_back(V.u[p]);
        }
        if ((rc = dev_alloc(&V.spec[SPEC_WORK], Nh))) return cleanup(rc);
        B->allocs.push_back(V.spec[SPEC_WORK]);
        // f^ and q: tile-major z-pass operands (zop_base in the kernels)
        if ((rc = dev_alloc(&V.spec[SPEC_A], zop_elems(V)))) return cleanup(rc);
        B->allocs.push_back(V.spec[SPEC_A]);
        if ((rc = dev_alloc(&V.q, zop_elems(V)))) return cleanup(rc);
        B->allocs.push_back(V.q);
        if ((rc = dev_alloc(&V.partial, 2 * (size_t)pb.cpass))) return cleanup(rc);
        B->allocs.push_back(V.partial);
    }
    B->gx = make_groups(lx, bx, false, &xshift);
    B->gxc = make_groups(lx, bc, false, &cshift);
    B->gy = make_groups(ly, by, true);
    B->gz = make_groups(lz, bz, true);
    B->lt_out = identity_table(bo);
    int rc;
    if ((rc = dev_alloc(&B->d_tw, tb.host.size()))) return cleanup(rc);
    B->allocs.push_back(B->d_tw);
    if ((rc = dev_alloc(&B->d_vols, n_vols))) return cleanup(rc);
    B->allocs.push_back(B->d_vols);
    if ((rc = dev_alloc(&B->d_status, n_vols))) return cleanup(rc);
    B->allocs.push_back(B->d_status);
    {
        hipError_t e = hipHostMalloc((void**)&B->h_status, n_vols * sizeof(VolStatus),
                                     hipHostMallocMapped | hipHostMallocCoherent);
        if (e != hipSuccess) { B->h_status = nullptr; return cleanup(fail(QSM_EHIP, "hipHostMalloc: %s", hipGetErrorString(e))); }
        e = hipHostGetDevicePointer((void**)&B->h_status_dev, B->h_status, 0);
        if (e != hipSuccess) return cleanup(fail(QSM_EHIP, "hipHostGetDevicePointer: %s", hipGetErrorString(e)));
        std::memset(B->h_status, 0, n_vols * sizeof(VolStatus));
    }
    if (stream) {
        B->stream = (hipStream_t)stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking);
        if (e != hipSuccess) { B->stream = nullptr; return cleanup(fail(QSM_EHIP, "hipStreamCreate: %s", hipGetErrorString(e))); }
        B->own_stream = true;
    }
    for (auto& e : B->ev) {
        hipError_t he = hipEventCreateWithFlags(&e, hipEventDisableTiming);
        if (he != hipSuccess) { e = nullptr; return cleanup(fail(QSM_EHIP, "hipEventCreate: %s", hipGetErrorString(he))); }
    }
    hipError_t e1 = hipMemcpy(B->d_tw, tb.host.data(), tb.host.size() * sizeof(float2), hipMemcpyHostToDevice);
    hipError_t e2 = hipMemcpy(B->d_vols, B->h.data(), n_vols * sizeof(VolDesc), hipMemcpyHostToDevice);
    if (e1 != hipSuccess || e2 != hipSuccess)
        return cleanup(fail(QSM_EHIP, "hipMemcpy of plan tables: %s",
                            hipGetErrorString(e1 != hipSuccess ? e1 : e2)));
    {
        int rc = tune_column_schedules(B);
        if (rc) return cleanup(rc);
    }
    if (B->gx.size() > 1 || B->gy.size() > 1 || B->gz.size() > 1 || B->gxc.size() > 1) {
        for (auto& st : B->side) {
            hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
            if (e != hipSuccess) { st = nullptr; return cleanup(fail(QSM_EHIP, "hipStreamCreate: %s", hipGetErrorString(e))); }
        }
        auto mk = [&](hipEvent_t& ev) {
            hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            if (e != hipSuccess) ev = nullptr;
            return e;
        };
        for (auto& ev : B->ev_fork)
            if (mk(ev) != hipSuccess) return cleanup(fail(QSM_EHIP, "hipEventCreate failed"));
        for (auto& row : B->ev_join)
            for (auto& ev : row)
                if (mk(ev) != hipSuccess) return cleanup(fail(QSM_EHIP, "hipEventCreate failed"));
    }
    *out = B;
    return QSM_OK;
}

// One TV-ADMM iteration (src/inversion/admm.rs:76-145) for every volume.
// Its first pass also closes iteration it-1 (relative change, convergence);
// `closed` (if set) is recorded right after that pass.
// One pass of an iteration over all groups: sequential on the batch's
// stream, or forked over the side streams (group i on lane i mod (S+1), lane
// 0 the batch's stream) and joined back.  slot: the pass's event set.
static int run_pass(qsm_tv_batch* B, int slot, int kind, const Groups& gs, const PassArgs& a) {
    if (gs.size() < 2 || !B->side[0]) {
        run_groups(kind, gs, a);
        return QSM_OK;
    }
    const int S = std::min<int>(qsm_tv_batch::kSide, (int)gs.size() - 1);
    QSM_HIP(hipEventRecord(B->ev_fork[slot], B->stream));
    for (int s = 0; s < S; ++s) QSM_HIP(hipStreamWaitEvent(B->side[s], B->ev_fork[slot], 0));
    PassArgs b = a;
    for (size_t i = 0; i < gs.size(); ++i) {
        const int lane = (int)(i % (size_t)(S + 1));
        b.stream = lane == 0 ? B->stream : B->side[lane - 1];
        b.sched = column_schedule(kind, gs[i].n);
        b.xshift = gs[i].xshift;
        launch_pass(kind, gs[i].n, gs[i].emax, gs[i].lt, b);
    }
    for (int s = 0; s < S; ++s) {
        QSM_HIP(hipEventRecord(B->ev_join[slot][s], B->side[s]));
        QSM_HIP(hipStreamWaitEvent(B->stream, B->ev_join[slot][s], 0));
    }
    return QSM_OK;
}

// rec >= 0 also records the shrink pattern of iteration it-1 as checkpoint
// `rec` (unrolled adjoint).
static int enqueue_iteration(qsm_tv_batch* B, size_t it, hipEvent_t closed, int rec = -1) {
    PassArgs a = B->args(B->d_status);
    a.parity = (int)((it - 1) & 1);
    if (B->method == 1 && it == 1) {
        // RTS starts from x0 != 0 with g = u = 0: the first step is x^ = f^
        // (no z/u update and no divergence).
        run_groups(P_Z_TV_SOLVE0, B->gz, a);
        run_groups(P_Y_INV, B->gy, a);
        run_groups(P_C2R_TV, B->gxc, a);
        QSM_CHECK_LAUNCH();
        return QSM_OK;
    }
    if (it >= 2) {
        a.status_rw = B->d_status;
        a.mirror = B->h_status_dev;
    }
    a.ridx = rec;
    int rc;
    if ((rc = run_pass(B, 0, rec >= 0 ? P_R2C_TVDIV_REC : P_R2C_TVDIV, B->gx, a))) return rc;
    if (closed) QSM_HIP(hipEventRecord(closed, B->stream));
    a.status_rw = nullptr;
    a.mirror = nullptr;
    if ((rc = run_pass(B, 1, P_Y_FWD, B->gy, a))) return rc;
    if ((rc = run_pass(B, 2, B->method == 1 ? P_Z_RTS_SOLVE : P_Z_TV_SOLVE, B->gz, a))) return rc;
    if ((rc = run_pass(B, 3, P_Y_INV, B->gy, a))) return rc;
    if ((rc = run_pass(B, 4, P_C2R_TV, B->gxc, a))) return rc;
    QSM_CHECK_LAUNCH();
    return QSM_OK;
}

// x_0 = 0, u_{-1} = 0 (AdmmBuffers::new, src/inversion/admm.rs:39-52), then
// f^ and inv_a (prepare_admm_spectral, src/inversion/admm.rs:150-178).
static int enqueue_setup(qsm_tv_batch* B) {
    hipStream_t st = B->stream;
    if (B->method == 1) {
        // RTS (src/inversion/rts.rs:93-183): x0 = mask * Re IFFT(FFT(f) w(D)),
        // then f^ = FFT(x0) (M/rho) inv_a and inv_a.  u[0] is what the second
        // step reads as the previous dual variable.
        for (auto& V : B->h) launch_fill_zero(st, V.u[0], 3 * (size_t)V.nx * V.ny * V.nz);
        launch_status_reset(st, B->d_status, B->nvol);
        PassArgs a = B->args(nullptr);
        run_groups(P_R2C_F64, B->gx, a);
        run_groups(P_Y_FWD, B->gy, a);
        run_groups(P_Z_FILTER, B->gz, a);
        run_groups(P_Y_INV, B->gy, a);
        run_groups(P_C2R_X0, B->gxc, a);
        run_groups(P_R2C_X0, B->gx, a);
        run_groups(P_Y_FWD, B->gy, a);
        run_groups(P_Z_RTS_SETUP, B->gz, a);
        QSM_CHECK_LAUNCH();
        return QSM_OK;
    }
    for (auto& V : B->h) {
        const size_t N = (size_t)V.nx * V.ny * V.nz;
        launch_fill_zero(st, V.x[0], N);
        launch_fill_zero(st, V.u[1], 3 * N);
    }
    launch_status_reset(st, B->d_status, B->nvol);
    PassArgs a = B->args(nullptr);
    run_groups(P_R2C_F64, B->gx, a);
    run_groups(P_Y_FWD, B->gy, a);
    run_groups(P_Z_TV_SETUP, B->gz, a);
    QSM_CHECK_LAUNCH();
    return QSM_OK;
}

// Capture the two iteration graphs on the batch's stream.  A stream that
// cannot be captured (the legacy null stream, or one already capturing)
// leaves the batch on eager launches; nothing runs during capture.  The
// graphs hold no event: the host's "iteration closed" event is recorded
// after the replay instead (see launch_iteration).
static int build_iteration_graphs(qsm_tv_batch* B) {
    for (int s = 0; s < 2; ++s) {
        const size_t it = (size_t)s + 3;  // (it - 1) & 1 == s, it >= 2
        if (hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            B->graph_state = -1;
            return QSM_OK;
        }
        const int rc = enqueue_iteration(B, it, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(B->stream, &g);
        hipError_t ei = hipSuccess;
        if (!rc && e == hipSuccess && g) ei = hipGraphInstantiate(&B->gexec[s], g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (rc) return rc;
        if (e != hipSuccess || !g || ei != hipSuccess) {
            (void)hipGetLastError();
            for (auto& x : B->gexec)
                if (x) { (void)hipGraphExecDestroy(x); x = nullptr; }
            B->graph_state = -1;
            return QSM_OK;
        }
    }
    B->graph_state = 1;
    return QSM_OK;
}

// Iteration `it` of a run: a graph replay where available, else the eager
// launches of enqueue_iteration (the first iteration always: its kernels
// differ, and RTS starts differently).  `closed` marks the point after which
// iteration it-1's convergence state is in the host mirror: eagerly right
// after the first pass of iteration it, after a replay at its end.
static int launch_iteration(qsm_tv_batch* B, size_t it, hipEvent_t closed) {
    if (it >= 2 && B->graph_state == 1) {
        QSM_HIP(hipGraphLaunch(B->gexec[(it - 1) & 1], B->stream));
        if (closed) QSM_HIP(hipEventRecord(closed, B->stream));
        return QSM_OK;
    }
    return enqueue_iteration(B, it, closed);
}

static int tv_batch_run_impl(qsm_tv_batch* B, qsm_progress_fn progress, void* user) {
    if (!B) return fail(QSM_EPARAM, "batch is null");
    hipStream_t st = B->stream;
    if (B->graph_state == 0 && B->params.max_iter >= 2) {
        int rc = build_iteration_graphs(B);
        if (rc) return rc;
    }
    std::memset(B->h_status, 0, B->nvol * sizeof(VolStatus));
    {
        int rc = enqueue_setup(B);
        if (rc) return rc;
    }

    const size_t max_iter = B->params.max_iter;
    // relative_change >= 0 (or NaN), so with tol <= 0 no iteration can stop
    // early and the host never needs to look at the device state.
    const bool can_converge = B->params.tol > 0.0;
    constexpr size_t W = qsm_tv_batch::kWindow;
    if (!progress && !can_converge) {
        for (size_t it = 1; it <= max_iter; ++it) {
            int rc = launch_iteration(B, it, nullptr);
            if (rc) return rc;
        }
        if (max_iter > 0) launch_tv_reduce(st, B->d_vols, B->nvol, B->d_status, B->h_status_dev);
    } else {
        // Iteration j is closed by the first pass of iteration j+1 (or by
        // k_tv_reduce after the last one); ev[j % W] marks that point.  Keep
        // up to W iterations in flight; report progress for iteration it+1
        // only once iteration it is known not to have converged, so the
        // callback sequence equals the reference's.
        size_t next = 1;
        if (max_iter > 0 && progress) progress(1, max_iter, user);
        for (size_t it = 1; it <= max_iter; ++it) {
            while (next <= max_iter && next <= it + W) {
                int rc = launch_iteration(B, next, next >= 2 ? B->ev[(next - 1) % W] : nullptr);
                if (rc) return rc;
                ++next;
            }
            if (it == max_iter) {
                launch_tv_reduce(st, B->d_vols, B->nvol, B->d_status, B->h_status_dev);
                QSM_HIP(hipEventRecord(B->ev[it % W], st));
            }
            QSM_HIP(hipEventSynchronize(B->ev[it % W]));
            bool all_done = true;
            for (int v = 0; v < B->nvol; ++v) {
                const volatile VolStatus* hs = B->h_status + v;
                const int done = hs->done, iters = hs->iters;
                if (!(done && (size_t)iters <= it)) all_done = false;
            }
            if (all_done) {
                if (progress) progress(it, it, user);
                break;
            }
            if (it < max_iter && progress) progress(it + 1, max_iter, user);
        }
    }
    launch_tv_output(st, B->d_vols, B->lt_out, B->d_status);
    QSM_CHECK_LAUNCH();
    QSM_HIP(hipStreamSynchronize(st));
    return QSM_OK;
}



// Event-timed copy of the iteration loop: every kernel of `iters` iterations
// is bracketed by HIP events on the batch's stream; ms_out[k] receives the
// mean duration of kernel slot k (0 r2c+div, 1 y fwd, 2 z solve, 3 y inv,
// 4 c2r).  The r2c slot includes the reduction of the previous iteration's
// relative change.  Runs the spectral setup first so the data are realistic.
static int tv_batch_profile_impl(qsm_tv_batch* B, size_t iters, double* ms_out, size_t n_out) {
    if (!B || !ms_out || n_out < 5 || iters == 0) return fail(QSM_EPARAM, "bad profile arguments");
    hipStream_t st = B->stream;
    {
        int rc = enqueue_setup(B);
        if (rc) return rc;
    }
    hipEvent_t ev[6];
    for (auto& e : ev) QSM_HIP(hipEventCreate(&e));
    double acc[5] = {0, 0, 0, 0, 0};
    PassArgs a = B->args(nullptr);
    for (size_t it = 1; it <= iters; ++it) {
        a.parity = (int)((it - 1) & 1);
        a.status_rw = it >= 2 ? B->d_status : nullptr;
        QSM_HIP(hipEventRecord(ev[0], st));
        run_groups(P_R2C_TVDIV, B->gx, a);
        a.status_rw = nullptr;
        QSM_HIP(hipEventRecord(ev[1], st));
        run_groups(P_Y_FWD, B->gy, a);
        QSM_HIP(hipEventRecord(ev[2], st));
        run_groups(B->method == 1 ? P_Z_RTS_SOLVE : P_Z_TV_SOLVE, B->gz, a);
        QSM_HIP(hipEventRecord(ev[3], st));
        run_groups(P_Y_INV, B->gy, a);
        QSM_HIP(hipEventRecord(ev[4], st));
        run_groups(P_C2R_TV, B->gxc, a);
        QSM_HIP(hipEventRecord(ev[5], st));
        QSM_CHECK_LAUNCH();
        QSM_HIP(hipEventSynchronize(ev[5]));
        for (int k = 0; k < 5; ++k) {
            float ms = 0.f;
            QSM_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            acc[k] += ms;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k < 5; ++k) ms_out[k] = acc[k] / (double)iters;
    return QSM_OK;
}

// ============================================ unrolled TV-ADMM adjoint ====
// K iterations of tv_admm with tol = 0 (src/inversion/tv.rs:63-97) and the
// vector-Jacobian product of chi = mask * x_K with respect to the local
// field and lambda.  The forward records, for iterations 1..K-1, the shrink
// pattern q = sign(v) [|v| > lambda/rho] (one byte per component): with it
// the backward recursion is linear and needs no recomputation of the
// forward iterates.
struct qsm_tv_unrolled {
    qsm_tv_batch* B = nullptr;
    size_t K = 0;
    std::vector<int> nblk;  // x-pass blocks per volume
    std::vector<void*> allocs;
    double* d_lbar = nullptr;
    bool forward_done = false;
};

static void unrolled_free(qsm_tv_unrolled* U) {
    if (!U) return;
    if (U->B) {
        if (U->B->stream) (void)hipStreamSynchronize(U->B->stream);
        batch_free(U->B);
    }
    for (void* p : U->allocs) (void)hipFree(p);
    delete U;
}

static int unrolled_create_impl(const qsm_tv_volume* vols, size_t n_vols, const qsm_tv_params* params,
                                void* stream, qsm_tv_unrolled** out) {
    if (!params || !out) return fail(QSM_EPARAM, "null argument");
    qsm_tv_params p = *params;
    p.tol = 0.0;  // unrolled: exactly max_iter iterations
    qsm_tv_unrolled* U = new qsm_tv_unrolled();
    U->K = p.max_iter;
    auto cleanup = [&](int rc) { unrolled_free(U); return rc; };
    int rc = tv_batch_create_impl(vols, n_vols, &p, stream, &U->B);
    if (rc) return cleanup(rc);
    qsm_tv_batch* B = U->B;
    const size_t nrec = U->K > 1 ? U->K - 1 : 0;
    if ((rc = dev_alloc(&U->d_lbar, n_vols))) return cleanup(rc);
    U->allocs.push_back(U->d_lbar);
    for (int v = 0; v < B->nvol; ++v) {
        VolDesc& V = B->h[v];
        const size_t N = (size_t)V.nx * V.ny * V.nz;
        const int npairs = (V.ny * V.nz + 1) / 2;
        const int nb = (npairs + V.rows_x - 1) / V.rows_x;
        U->nblk.push_back(nb);
        V.adj_nblk = nb;
        if ((rc = dev_alloc(&V.ckpt, std::max<size_t>(1, nrec) * 3 * N))) return cleanup(rc);
        U->allocs.push_back(V.ckpt);
        if ((rc = dev_alloc(&V.adj_c, N))) return cleanup(rc);
        U->allocs.push_back(V.adj_c);
        for (int q = 0; q < 2; ++q) {
            if ((rc = dev_alloc(&V.adj_s[q], 3 * N))) return cleanup(rc);
            U->allocs.push_back(V.adj_s[q]);
        }
        if ((rc = dev_alloc(&V.adj_acc, N))) return cleanup(rc);
        U->allocs.push_back(V.adj_acc);
        if ((rc = dev_alloc(&V.adj_tpart, std::max<size_t>(1, nrec) * (size_t)nb))) return cleanup(rc);
        U->allocs.push_back(V.adj_tpart);
        V.adj_lbar = U->d_lbar + v;
    }
    hipError_t e = hipMemcpy(B->d_vols, B->h.data(), B->h.size() * sizeof(VolDesc), hipMemcpyHostToDevice);
    if (e != hipSuccess) return cleanup(fail(QSM_EHIP, "hipMemcpy descriptors: %s", hipGetErrorString(e)));
    *out = U;
    return QSM_OK;
}

static int unrolled_forward_impl(qsm_tv_unrolled* U) {
    if (!U) return fail(QSM_EPARAM, "handle is null");
    qsm_tv_batch* B = U->B;
    hipStream_t st = B->stream;
    std::memset(B->h_status, 0, B->nvol * sizeof(VolStatus));
    int rc = enqueue_setup(B);
    if (rc) return rc;
    for (size_t it = 1; it <= U->K; ++it) {
        if ((rc = enqueue_iteration(B, it, nullptr, it >= 2 ? (int)(it - 2) : -1))) return rc;
    }
    if (U->K > 0) launch_tv_reduce(st, B->d_vols, B->nvol, B->d_status, B->h_status_dev);
    launch_tv_output(st, B->d_vols, B->lt_out, B->d_status);
    QSM_CHECK_LAUNCH();
    QSM_HIP(hipStreamSynchronize(st));
    U->forward_done = true;
    return QSM_OK;
}

static int unrolled_backward_impl(qsm_tv_unrolled* U, const double* const* d_chi_bar,
                                  double* const* d_field_bar, double* lambda_bar) {
    if (!U || !d_chi_bar || !d_field_bar) return fail(QSM_EPARAM, "null argument");
    if (!U->forward_done) return fail(QSM_EPARAM, "backward needs a forward pass first");
    qsm_tv_batch* B = U->B;
    hipStream_t st = B->stream;
    for (int v = 0; v < B->nvol; ++v) {
        if (!d_chi_bar[v] || !d_field_bar[v]) return fail(QSM_EPARAM, "volume %d has a null pointer", v);
        B->h[v].adj_chibar = d_chi_bar[v];
        B->h[v].adj_out = d_field_bar[v];
    }
    QSM_HIP(hipMemcpyAsync(B->d_vols, B->h.data(), B->h.size() * sizeof(VolDesc), hipMemcpyHostToDevice, st));
    PassArgs a = B->args(nullptr);
    const size_t K = U->K;
    if (K == 0) {
        // chi = mask * x_0 = 0 does not depend on the field or lambda
        for (int v = 0; v < B->nvol; ++v) {
            const VolDesc& V = B->h[v];
            const size_t N = (size_t)V.nx * V.ny * V.nz;
            launch_fill_zero(st, reinterpret_cast<float*>(V.adj_out), 2 * N);
        }
        QSM_HIP(hipStreamSynchronize(st));
        if (lambda_bar)
            for (int v = 0; v < B->nvol; ++v) lambda_bar[v] = 0.0;
        return QSM_OK;
    }
    for (size_t j = K; j >= 1; --j) {
        a.parity = (int)(j & 1);
        a.ridx = (j == K) ? -1 : (int)(j - 1);
        run_groups(j == K ? P_R2C_TVADJ_FIRST : P_R2C_TVADJ, B->gx, a);
        if (j > 1) {  // c-bar_j feeds iteration j-1; c-bar_1 is not needed
            run_groups(P_Y_FWD, B->gy, a);
            run_groups(P_Z_TV_ADJ, B->gz, a);
            run_groups(P_Y_INV, B->gy, a);
            run_groups(P_C2R_ADJ, B->gxc, a);
        }
        QSM_CHECK_LAUNCH();
    }
    // field-bar = A (sum_j x-bar_j)
    run_groups(P_R2C_ACC, B->gx, a);
    run_groups(P_Y_FWD, B->gy, a);
    run_groups(P_Z_TV_APPLYA, B->gz, a);
    run_groups(P_Y_INV, B->gy, a);
    run_groups(P_C2R_F64OUT, B->gxc, a);
    launch_tv_lbar(st, B->d_vols, B->nvol, (int)(K - 1));
    QSM_CHECK_LAUNCH();
    if (lambda_bar)
        QSM_HIP(hipMemcpyAsync(lambda_bar, U->d_lbar, B->nvol * sizeof(double), hipMemcpyDeviceToHost, st));
    QSM_HIP(hipStreamSynchronize(st));
    return QSM_OK;
}

// =================================================== dipole operator ======
// field = Re IFFT(D FFT(chi)) (tests/chisep_integration.rs:18-62, kernel of
// src/kernels/dipole.rs:24-72), device pointers, f64 in and out, FP32
// transforms.  D is real and even after taking its Hermitian part, so the
// operator is self-adjoint and the adjoint entry point runs the same passes.
// zkind P_Z_DIPOLE (no mask, f64 out) or P_Z_FILTER (filter `filt` with
// parameter `fa`, output masked by d_mask).
// src_kind P_R2C_F64 (d_in as is) or P_R2C_WRAPLAP (masked wrapped
// Laplacian of d_in); bdir may be null for filters that do not use D.
// The five passes of one spectral round trip of a single volume: R2C of
// src_kind, y, z pass zkind, inverse y, C2R to f64 (masked for filters).
// Launch tables live on the stack.
static void enqueue_round_trip(int src_kind, int zkind, const VolDesc& V, const PassBlocks& pb,
                               const VolDesc* d_vols, const float2* d_tw, hipStream_t st) {
    auto table = [](int blocks) {
        LaunchTable lt{};
        lt.nvol = 1;
        lt.start[0] = 0;
        lt.start[1] = blocks;
        lt.vol[0] = 0;
        return lt;
    };
    const LaunchTable tx = table(pb.xpass), txc = table(pb.cpass), ty = table(pb.ypass),
                      tz = table(pb.zpass);
    PassArgs a{};
    a.stream = st;
    a.vols = d_vols;
    a.tw = d_tw;
    a.sel = SPEC_WORK;
    a.sel2 = -1;
    a.xshift = 0;
    auto run = [&](int kind, int n, bool column, const LaunchTable& lt) {
        a.sched = column_schedule(kind, n);
        launch_pass(kind, n, column ? emax_for_col(n) : 16, lt, a);
    };
    run(src_kind, V.nx, false, tx);
    run(P_Y_FWD, V.ny, true, ty);
    run(zkind, V.nz, true, tz);
    run(P_Y_INV, V.ny, true, ty);
    run(zkind == P_Z_FILTER ? P_C2R_F64_MASKED : P_C2R_F64OUT, V.nx, false, txc);
}

static int spectral_apply(int zkind, int filt, double fa, const double* d_in, const uint8_t* d_mask,
                          const qsm_grid* grid, const double bdir_in[3], double* d_out, void* stream,
                          int src_kind = P_R2C_F64) {
    static const double z_axis[3] = {0.0, 0.0, 1.0};
    const double* bdir = bdir_in ? bdir_in : z_axis;
    if (!d_in || !grid || !d_out || (zkind == P_Z_FILTER && !d_mask))
        return fail(QSM_EPARAM, "null argument");
    int rc = validate_grid(grid);
    if (rc) return rc;
    const double bn = std::sqrt(bdir[0] * bdir[0] + bdir[1] * bdir[1] + bdir[2] * bdir[2]);
    if (!(bn > 0)) return fail(QSM_EPARAM, "B0 direction has zero length");
    VolDesc V;
    std::memset(&V, 0, sizeof V);
    TwiddleBank tb;
    PassBlocks pb = fill_geometry(V, *grid, tb);
    V.bx = bdir[0] / bn; V.by = bdir[1] / bn; V.bz = bdir[2] / bn;
    V.field = d_in;
    V.adj_out = d_out;
    V.out = d_out;
    V.mask = d_mask;
    V.filt = filt;
    V.filt_a = fa;
    const size_t Nh = (size_t)V.nxh * V.ny * V.nz;
    std::vector<void*> allocs;
    float2* d_tw = nullptr;
    VolDesc* d_vols = nullptr;
    hipStream_t st = (hipStream_t)stream;
    bool own = false;
    auto done = [&](int code) {
        if (st) (void)hipStreamSynchronize(st);
        for (void* p : allocs) (void)hipFree(p);
        if (own && st) (void)hipStreamDestroy(st);
        return code;
    };
    if ((rc = dev_alloc(&V.spec[SPEC_WORK], Nh))) return done(rc);
    allocs.push_back(V.spec[SPEC_WORK]);
    if ((rc = dev_alloc(&d_tw, tb.host.size()))) return done(rc);
    allocs.push_back(d_tw);
    if ((rc = dev_alloc(&d_vols, 1))) return done(rc);
    allocs.push_back(d_vols);
    if (!st) {
        hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) { st = nullptr; return done(fail(QSM_EHIP, "hipStreamCreate: %s", hipGetErrorString(e))); }
        own = true;
    }
    hipError_t e1 = hipMemcpy(d_tw, tb.host.data(), tb.host.size() * sizeof(float2), hipMemcpyHostToDevice);
    hipError_t e2 = hipMemcpy(d_vols, &V, sizeof V, hipMemcpyHostToDevice);
    if (e1 != hipSuccess || e2 != hipSuccess)
        return done(fail(QSM_EHIP, "hipMemcpy of plan tables: %s", hipGetErrorString(e1 != hipSuccess ? e1 : e2)));
    enqueue_round_trip(src_kind, zkind, V, pb, d_vols, d_tw, st);
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return done(fail(QSM_EHIP, "kernel launch: %s", hipGetErrorString(le)));
    hipError_t se = hipStreamSynchronize(st);
    if (se != hipSuccess) return done(fail(QSM_EHIP, "hipStreamSynchronize: %s", hipGetErrorString(se)));
    return done(QSM_OK);
}

// ========================================================= V-SHARP batch ==
// Voxels of the wrapped sphere of src/kernels/smv.rs:19-67, counted over the
// sphere's bounding box only (same coordinates and the same f64 operation
// order as the kernel, so the count equals the reference's `count`).
static double sphere_count(const qsm_grid& g, double radius) {
#pragma clang fp contract(off)
    const double r2 = radius * radius;
    auto coords = [&](size_t n, double vs, std::vector<double>& out) {
        const double lim = std::fabs(radius) / vs + 1.0;
        for (size_t i = 0; i < n; ++i) {
            double c = (i <= n / 2) ? (double)i : (double)((long long)i - (long long)n);
            if (std::fabs(c) <= lim) out.push_back(c * vs);
        }
    };
    std::vector<double> xs, ys, zs;
    coords(g.nx, g.vsx, xs);
    coords(g.ny, g.vsy, ys);
    coords(g.nz, g.vsz, zs);
    double count = 0.0;
    for (double dz : zs)
        for (double dy : ys)
            for (double dx : xs) {
                volatile double d2 = dx * dx;
                d2 = d2 + dy * dy;
                d2 = d2 + dz * dz;
                if (d2 <= r2) count += 1.0;
            }
    return count;
}

struct qsm_vsharp_batch {
    int nvol = 0;
    std::vector<VolDesc> h;
    VolDesc* d_vols = nullptr;
    float2* d_tw = nullptr;
    double* d_radii = nullptr;
    std::vector<double> radii;  // processing order (descending)
    double threshold = 0.05;
    bool sharp = false;         // one radius: sharp_core semantics
    Groups gx, grows, gy, gy2, gz;   // passes grouped by axis length
    LaunchTable lt_init{};
    std::vector<void*> allocs;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    PassArgs args() const {
        PassArgs a{};
        a.stream = stream;
        a.vols = d_vols;
        a.tw = d_tw;
        a.status = nullptr;
        a.sel = SPEC_WORK;
        a.sel2 = -1;
        return a;
    }
};

static void vs_free(qsm_vsharp_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

static int vs_create_impl(const qsm_vsharp_volume* vols, size_t n_vols, const double* radii,
                          size_t n_radii, double threshold, void* stream, qsm_vsharp_batch** out) {
    if (!vols || !out || (n_radii && !radii)) return fail(QSM_EPARAM, "null argument");
    if (n_vols == 0 || n_vols > (size_t)kMaxVolumes)
        return fail(QSM_EPARAM, "batch size %zu outside 1..%d", n_vols, kMaxVolumes);
    if (n_radii == 0) return fail(QSM_EPARAM, "at least one radius is required for a batch");
    for (size_t r = 0; r < n_radii; ++r)
        if (std::isnan(radii[r])) return fail(QSM_EPARAM, "radius %zu is NaN", r);
    for (size_t v = 0; v < n_vols; ++v) {
        int rc = validate_grid(&vols[v].grid);
        if (rc) return rc;
        if (!vols[v].d_field || !vols[v].d_mask || !vols[v].d_local || !vols[v].d_mask_out)
            return fail(QSM_EPARAM, "volume %zu has a null device pointer", v);
    }
    qsm_vsharp_batch* B = new qsm_vsharp_batch();
    auto cleanup = [&](int rc) { vs_free(B); return rc; };
    B->nvol = (int)n_vols;
    B->threshold = threshold;
    // src/bgremove/vsharp.rs:114-116: sort descending (stable)
    B->radii.assign(radii, radii + n_radii);
    std::stable_sort(B->radii.begin(), B->radii.end(), [](double a, double b) { return b < a; });
    B->sharp = (n_radii == 1);
    const int R = (int)n_radii;
    B->h.resize(n_vols);
    TwiddleBank tb;
    std::vector<int> lx, ly, lz, bx, brows, by, by2, bz, bi;
    int rc;
    if ((rc = dev_alloc(&B->d_radii, (size_t)R * (1 + n_vols)))) return cleanup(rc);
    B->allocs.push_back(B->d_radii);
    std::vector<double> host_tab((size_t)R * (1 + n_vols));
    for (int r = 0; r < R; ++r) host_tab[r] = B->radii[r];
    for (size_t v = 0; v < n_vols; ++v) {
        VolDesc& V = B->h[v];
        std::memset(&V, 0, sizeof V);
        PassBlocks pb = fill_geometry(V, vols[v].grid, tb);
        V.field = vols[v].d_field;
        V.mask = vols[v].d_mask;
        V.out = vols[v].d_local;
        V.mask_out = vols[v].d_mask_out;
        V.vs_threshold = threshold;
        V.vs_radii = B->d_radii;
        V.vs_counts = B->d_radii + (size_t)R * (1 + v);
        for (int r = 0; r < R; ++r) host_tab[(size_t)R * (1 + v) + r] = sphere_count(vols[v].grid, B->radii[r]);
        const int nrows = V.ny * V.nz;
        lx.push_back(V.nx); bx.push_back(pb.xpass);
        brows.push_back((nrows + V.rows_x - 1) / V.rows_x);
        ly.push_back(V.ny); by.push_back(pb.ypass); by2.push_back(2 * pb.ypass);
        lz.push_back(V.nz); bz.push_back(pb.zpass);
        bi.push_back(pb.out);
        const size_t N = (size_t)V.nx * V.ny * V.nz;
        const size_t Nh = (size_t)V.nxh * V.ny * V.nz;
        for (int sidx = 0; sidx < SPEC_COUNT; ++sidx) {
            if ((rc = dev_alloc(&V.spec[sidx], Nh))) return cleanup(rc);
            B->allocs.push_back(V.spec[sidx]);
        }
        // S^ of the largest sphere; 2 nz more for the Nyquist plane's doubly packed columns
        if ((rc = dev_alloc(&V.vs_sfirst, Nh + 2 * (size_t)V.nz))) return cleanup(rc);
        B->allocs.push_back(V.vs_sfirst);
        if ((rc = dev_alloc(&V.vs_local, N))) return cleanup(rc);
        B->allocs.push_back(V.vs_local);
        if ((rc = dev_alloc(&V.vs_final, N))) return cleanup(rc);
        B->allocs.push_back(V.vs_final);
    }
    B->gx = make_groups(lx, bx, false);
    B->grows = make_groups(lx, brows, false);
    B->gy = make_groups(ly, by, true);
    B->gy2 = make_groups(ly, by2, true);
    B->gz = make_groups(lz, bz, true);
    B->lt_init = identity_table(bi);
    if ((rc = dev_alloc(&B->d_tw, tb.host.size()))) return cleanup(rc);
    B->allocs.push_back(B->d_tw);
    if ((rc = dev_alloc(&B->d_vols, n_vols))) return cleanup(rc);
    B->allocs.push_back(B->d_vols);
    if (stream) {
        B->stream = (hipStream_t)stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking);
        if (e != hipSuccess) { B->stream = nullptr; return cleanup(fail(QSM_EHIP, "hipStreamCreate: %s", hipGetErrorString(e))); }
        B->own_stream = true;
    }
    hipError_t e1 = hipMemcpy(B->d_tw, tb.host.data(), tb.host.size() * sizeof(float2), hipMemcpyHostToDevice);
    hipError_t e2 = hipMemcpy(B->d_vols, B->h.data(), n_vols * sizeof(VolDesc), hipMemcpyHostToDevice);
    hipError_t e3 = hipMemcpy(B->d_radii, host_tab.data(), host_tab.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
        return cleanup(fail(QSM_EHIP, "hipMemcpy of plan tables failed"));
    *out = B;
    return QSM_OK;
}

static int vs_run_impl(qsm_vsharp_batch* B, qsm_progress_fn progress, void* user) {
    if (!B) return fail(QSM_EPARAM, "batch is null");
    hipStream_t st = B->stream;
    const int R = (int)B->radii.size();
    const double delta = 1.0 - std::sqrt(1e-7);  // src/bgremove/vsharp.rs:129
    // F^ = FFT(field) (vsharp.rs:117-122), M^ = FFT(mask) once for all radii
    PassArgs a = B->args();
    a.sel = SPEC_A;
    run_groups(P_R2C_F64, B->gx, a);
    run_groups(P_Y_FWD, B->gy, a);
    run_groups(P_Z_VS_FWD, B->gz, a);
    a.sel = SPEC_B;
    run_groups(P_R2C_U8, B->gx, a);
    run_groups(P_Y_FWD, B->gy, a);
    run_groups(P_Z_VS_FWD, B->gz, a);
    launch_vs_init(st, B->d_vols, B->lt_init);
    QSM_CHECK_LAUNCH();
    a.sel = SPEC_WORK;
    a.delta = delta;
    for (int r = 0; r < R; ++r) {
        if (progress) {
            // report radius r+1 once radius r has finished on the device
            if (r > 0) QSM_HIP(hipStreamSynchronize(st));
            if (B->sharp) progress(1, 1, user);           // vsharp.rs:103-108
            else progress((size_t)r + 1, (size_t)R, user); // vsharp.rs:134
        }
        a.ridx = r;
        a.first = (r == 0) ? 1 : 0;
        a.sel2 = -1;
        run_groups(P_R2C_SPHERE, B->gx, a);
        run_groups(P_Y_FWD, B->gy, a);
        run_groups(P_Z_VS_RADIUS, B->gz, a);
        a.sel2 = SPEC_C;
        run_groups(P_Y_INV, B->gy2, a);
        a.sel2 = -1;
        run_groups(P_C2R_VS_SELECT, B->grows, a);
        QSM_CHECK_LAUNCH();
    }
    // deconvolution with the largest kernel (vsharp.rs:180-196 / sharp.rs:105-131)
    a.divide = B->sharp ? 1 : 0;
    run_groups(P_R2C_F32, B->gx, a);
    run_groups(P_Y_FWD, B->gy, a);
    run_groups(P_Z_VS_DECONV, B->gz, a);
    run_groups(P_Y_INV, B->gy, a);
    run_groups(P_C2R_VS_OUT, B->gx, a);
    QSM_CHECK_LAUNCH();
    QSM_HIP(hipStreamSynchronize(st));
    return QSM_OK;
}

// Host-pointer V-SHARP with explicit radii (src/bgremove/vsharp.rs:88-199).
static int vsharp_host(const double* field, const uint8_t* mask, const qsm_grid* grid,
                       const double* radii, size_t n_radii, double threshold,
                       qsm_progress_fn progress, void* user, double* local_out, uint8_t* mask_out) {
    if (!field || !mask || !grid || !local_out || !mask_out || (n_radii && !radii))
        return fail(QSM_EPARAM, "null argument");
    int rc = validate_grid(grid);
    if (rc) return rc;
    const size_t N = grid->nx * grid->ny * grid->nz;
    if (n_radii == 0) {  // vsharp.rs:98-100
        std::fill(local_out, local_out + N, 0.0);
        std::memcpy(mask_out, mask, N);
        return QSM_OK;
    }
    double *d_f = nullptr, *d_l = nullptr;
    uint8_t *d_m = nullptr, *d_mo = nullptr;
    qsm_vsharp_batch* B = nullptr;
    auto done = [&](int code) {
        if (B) vs_free(B);
        if (d_f) (void)hipFree(d_f);
        if (d_l) (void)hipFree(d_l);
        if (d_m) (void)hipFree(d_m);
        if (d_mo) (void)hipFree(d_mo);
        return code;
    };
    if ((rc = dev_alloc(&d_f, N)) || (rc = dev_alloc(&d_l, N)) || (rc = dev_alloc(&d_m, N)) ||
        (rc = dev_alloc(&d_mo, N)))
        return done(rc);
    hipError_t e = hipMemcpy(d_f, field, N * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_m, mask, N, hipMemcpyHostToDevice);
    if (e != hipSuccess) return done(fail(QSM_EHIP, "hipMemcpy H2D: %s", hipGetErrorString(e)));
    qsm_vsharp_volume vol;
    vol.grid = *grid;
    vol.d_field = d_f; vol.d_mask = d_m; vol.d_local = d_l; vol.d_mask_out = d_mo;
    if ((rc = vs_create_impl(&vol, 1, radii, n_radii, threshold, nullptr, &B))) return done(rc);
    if ((rc = vs_run_impl(B, progress, user))) return done(rc);
    e = hipMemcpy(local_out, d_l, N * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(mask_out, d_mo, N, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return done(fail(QSM_EHIP, "hipMemcpy D2H: %s", hipGetErrorString(e)));
    return done(QSM_OK);
}

// src/bgremove/vsharp.rs:70-82.  A non-positive step would never terminate
// in the reference; it is rejected instead.
static int vsharp_radii(const qsm_vsharp_params* p, std::vector<double>& radii) {
    if (!(p->min_radius > 0.0))
        return fail(QSM_EPARAM, "min_radius must be > 0 (it is the step between radii)");
    if (std::isnan(p->max_radius)) return fail(QSM_EPARAM, "max_radius is NaN");
    double r = p->max_radius;
    const double step = p->min_radius;
    while (r >= step) {
        radii.push_back(r);
        r -= step;
        if (radii.size() > 100000) return fail(QSM_EPARAM, "too many radii");
    }
    if (radii.empty()) radii.push_back(p->max_radius);
    return QSM_OK;
}

// ====================================================== extern "C" API ====
extern "C" {

void qsm_tv_params_default(qsm_tv_params* p) {
    if (!p) return;
    p->lambda = 2e-4;  // src/inversion/tv.rs:34-43
    p->rho = 2e-2;
    p->tol = 1e-3;
    p->max_iter = 250;
}

void qsm_vsharp_params_default(qsm_vsharp_params* p) {
    if (!p) return;
    p->threshold = 0.05;  // src/bgremove/vsharp.rs:31-39
    p->max_radius = 12.0;
    p->min_radius = 1.0;
}

const char* qsm_last_error_message(void) { return g_err.c_str(); }

int qsm_fft_length_supported(size_t n) { return length_ok(n) ? 1 : 0; }

int qsm_backend_info(char* buf, size_t len) {
    if (!buf || len == 0) return fail(QSM_EPARAM, "null buffer");
    int dev = 0;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) {
        snprintf(buf, len, "qsm_gpu (gfx950 build); no HIP device: %s", hipGetErrorString(e));
        return QSM_OK;
    }
    snprintf(buf, len, "qsm_gpu gfx950 build on %s (%s), %d CUs, %.1f GiB HBM", prop.name,
             prop.gcnArchName, prop.multiProcessorCount, prop.totalGlobalMem / 1073741824.0);
    return QSM_OK;
}

int qsm_tv_batch_create(const qsm_tv_volume* vols, size_t n_vols, const qsm_tv_params* params,
                        void* stream, qsm_tv_batch** out) {
    g_err.clear();
    return tv_batch_create_impl(vols, n_vols, params, stream, out);
}

int qsm_tv_batch_run(qsm_tv_batch* batch, qsm_progress_fn progress, void* user) {
    g_err.clear();
    return tv_batch_run_impl(batch, progress, user);
}

int qsm_set_fft_schedule(int mode) {
    g_err.clear();
    if (mode < 0 || mode > 3) return fail(QSM_EPARAM, "schedule mode must be 0, 1, 2 or 3");
    g_sched_mode.store(mode);
    return QSM_OK;
}

int qsm_fft_schedule(int axis, int n, int* sched) {
    g_err.clear();
    if (!sched || (axis != 1 && axis != 2)) return fail(QSM_EPARAM, "axis must be 1 (y) or 2 (z)");
    if (n <= 0 || n > kMaxAxis || !qsmfft::col2_ok(n)) {
        *sched = 1;  // only the staged transform exists for this length
        return QSM_OK;
    }
    const int mode = g_sched_mode.load();
    if (mode == 1 || mode == 2) {
        *sched = mode - 1;
        return QSM_OK;
    }
    if (mode == 0) {
        *sched = fixed_schedule(axis - 1, n);
        return QSM_OK;
    }
    std::lock_guard<std::mutex> lk(g_sched_mu);
    *sched = g_sched_tab[axis - 1][n];
    return QSM_OK;
}

int qsm_tv_batch_set_graphs(qsm_tv_batch* B, int enable) {
    g_err.clear();
    if (!B) return fail(QSM_EPARAM, "batch is null");
    if (!enable) {
        for (auto& g : B->gexec)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
        B->graph_state = -1;
    } else if (B->graph_state == -1) {
        B->graph_state = 0;  // captured at the next run
    }
    return QSM_OK;
}

int qsm_tv_batch_status(const qsm_tv_batch* B, size_t v, size_t* iters, double* rel) {
    if (!B || v >= (size_t)B->nvol) return fail(QSM_EPARAM, "bad batch or volume index");
    VolStatus s;
    hipError_t e = hipMemcpy(&s, B->d_status + v, sizeof s, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(QSM_EHIP, "hipMemcpy status: %s", hipGetErrorString(e));
    if (iters) *iters = (size_t)s.iters;
    if (rel) *rel = s.rel;
    return QSM_OK;
}

void qsm_tv_batch_destroy(qsm_tv_batch* batch) { batch_free(batch); }

int qsm_tv_batch_profile(qsm_tv_batch* batch, size_t iters, double* ms_out, size_t n_out) {
    g_err.clear();
    return tv_batch_profile_impl(batch, iters, ms_out, n_out);
}

int qsm_tv_admm_ex(const double* local_field, const uint8_t* mask, const qsm_grid* grid,
                   const double bdir[3], const qsm_tv_params* params, qsm_progress_fn progress,
                   void* user, double* chi_out, size_t* iters_out, double* rel_out) {
    g_err.clear();
    if (!local_field || !mask || !grid || !bdir || !params || !chi_out)
        return fail(QSM_EPARAM, "null argument");
    int rc = validate_grid(grid);
    if (rc) return rc;
    const size_t N = grid->nx * grid->ny * grid->nz;
    double *d_f = nullptr, *d_chi = nullptr;
    uint8_t* d_m = nullptr;
    qsm_tv_batch* B = nullptr;
    auto done = [&](int code) {
        if (B) batch_free(B);
        if (d_f) (void)hipFree(d_f);
        if (d_chi) (void)hipFree(d_chi);
        if (d_m) (void)hipFree(d_m);
        return code;
    };
    if ((rc = dev_alloc(&d_f, N)) || (rc = dev_alloc(&d_chi, N)) || (rc = dev_alloc(&d_m, N)))
        return done(rc);
    hipError_t e = hipMemcpy(d_f, local_field, N * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_m, mask, N, hipMemcpyHostToDevice);
    if (e != hipSuccess) return done(fail(QSM_EHIP, "hipMemcpy H2D: %s", hipGetErrorString(e)));
    qsm_tv_volume vol;
    vol.grid = *grid;
    vol.bdir[0] = bdir[0]; vol.bdir[1] = bdir[1]; vol.bdir[2] = bdir[2];
    vol.d_field = d_f; vol.d_mask = d_m; vol.d_chi = d_chi;
    if ((rc = tv_batch_create_impl(&vol, 1, params, nullptr, &B))) return done(rc);
    if ((rc = tv_batch_run_impl(B, progress, user))) return done(rc);
    e = hipMemcpy(chi_out, d_chi, N * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return done(fail(QSM_EHIP, "hipMemcpy D2H: %s", hipGetErrorString(e)));
    if (iters_out || rel_out) {
        size_t it = 0;
        double rel = 0;
        if ((rc = qsm_tv_batch_status(B, 0, &it, &rel))) return done(rc);
        if (iters_out) *iters_out = it;
        if (rel_out) *rel_out = rel;
    }
    return done(QSM_OK);
}

int qsm_tv_admm(const double* local_field, const uint8_t* mask, const qsm_grid* grid,
                const double bdir[3], const qsm_tv_params* params, qsm_progress_fn progress,
                void* user, double* chi_out) {
    return qsm_tv_admm_ex(local_field, mask, grid, bdir, params, progress, user, chi_out, nullptr,
                          nullptr);
}

int qsm_run_dipole_inversion(const double* local_field_ppm, size_t field_len, const uint8_t* mask,
                             size_t mask_len, const qsm_scan_metadata* md,
                             qsm_inversion_algorithm algorithm, const qsm_tv_params* tv_params,
                             qsm_progress_fn progress, void* user, double* chi_out) {
    g_err.clear();
    if (!md) return fail(QSM_EPARAM, "metadata is null");
    const size_t n = md->nx * md->ny * md->nz;
    // src/pipeline/inversion.rs:36-45
    if (field_len != n) return fail(QSM_EDIM, "dimension mismatch: expected %zu, got %zu", n, field_len);
    if (mask_len != n) return fail(QSM_EDIM, "dimension mismatch: expected %zu, got %zu", n, mask_len);
    if (algorithm == QSM_INV_TGV || algorithm == QSM_INV_QSMART)
        return fail(QSM_ECONFIG, "invalid config: %s should use run_tgv or run_qsmart",
                    algorithm == QSM_INV_TGV ? "Tgv" : "Qsmart");
    qsm_inversion_config cfg;
    qsm_inversion_config_default(&cfg);
    cfg.algorithm = algorithm;
    if (tv_params) cfg.tv = *tv_params;
    return qsm_run_dipole_inversion_cfg(local_field_ppm, field_len, mask, mask_len, md, &cfg, progress,
                                        user, chi_out);
}

int qsm_tv_unrolled_create(const qsm_tv_volume* vols, size_t n_vols, const qsm_tv_params* params,
                           void* stream, qsm_tv_unrolled** out) {
    g_err.clear();
    return unrolled_create_impl(vols, n_vols, params, stream, out);
}

int qsm_tv_unrolled_forward(qsm_tv_unrolled* h) {
    g_err.clear();
    return unrolled_forward_impl(h);
}

int qsm_tv_unrolled_backward(qsm_tv_unrolled* h, const double* const* d_chi_bar,
                             double* const* d_field_bar, double* lambda_bar) {
    g_err.clear();
    return unrolled_backward_impl(h, d_chi_bar, d_field_bar, lambda_bar);
}

void qsm_tv_unrolled_destroy(qsm_tv_unrolled* h) { unrolled_free(h); }

int qsm_dipole_forward(const double* d_chi, const qsm_grid* grid, const double bdir[3], double* d_field,
                       void* stream) {
    g_err.clear();
    return spectral_apply(P_Z_DIPOLE, 0, 0.0, d_chi, nullptr, grid, bdir, d_field, stream);
}

int qsm_dipole_adjoint(const double* d_field, const qsm_grid* grid, const double bdir[3], double* d_chi,
                       void* stream) {
    g_err.clear();
    return spectral_apply(P_Z_DIPOLE, 0, 0.0, d_field, nullptr, grid, bdir, d_chi, stream);
}

// ------------------------------------- dipole operator, no allocation ----
// Workspace of qsm_dipole_forward_noalloc in a file-scope device array: the
// volume descriptor, the twiddle tables (three lengths <= 1024) and one half
// spectrum.  Each device holds its own copy until the process ends.  Host
// code never passes `scratch` itself to a kernel or a copy: its device
// address comes from hipGetSymbolAddress, once per device.
namespace {
constexpr size_t kScratchDescBytes = 4096;
constexpr size_t kScratchTwBytes = sizeof(float2) * FixedTwiddleBank::kCap;
constexpr size_t kScratchSpecOff = 65536;
constexpr size_t kScratchSpecBytes = (size_t)72 << 20;
constexpr size_t kScratchBytes = kScratchSpecOff + kScratchSpecBytes;
static_assert(sizeof(VolDesc) <= kScratchDescBytes, "descriptor does not fit its scratch slot");
static_assert(kScratchDescBytes + kScratchTwBytes <= kScratchSpecOff, "twiddles overlap the spectrum");
constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the current device (looked up on the first
// call on each device).
static int scratch_address(unsigned char** out) {
    int dev = 0;
    QSM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchMaxDevices) return fail(QSM_EHIP, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(QSM_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    return QSM_OK;
}

int qsm_dipole_forward_noalloc(const double* d_chi, const qsm_grid* grid, const double bdir[3],
                               double* d_field, void* stream) {
    g_err.clear();
    if (!d_chi || !grid || !bdir || !d_field) return fail(QSM_EPARAM, "null argument");
    int rc = validate_grid(grid);
    if (rc) return rc;
    const double bn = std::sqrt(bdir[0] * bdir[0] + bdir[1] * bdir[1] + bdir[2] * bdir[2]);
    if (!(bn > 0)) return fail(QSM_EPARAM, "B0 direction has zero length");
    VolDesc V;
    std::memset(&V, 0, sizeof V);
    FixedTwiddleBank tb;
    PassBlocks pb = fill_geometry(V, *grid, tb);
    const size_t spec_bytes = (size_t)V.nxh * V.ny * V.nz * sizeof(float2);
    if (spec_bytes > kScratchSpecBytes)
        return fail(QSM_EPARAM,
                    "volume %zux%zux%zu needs %zu bytes of spectrum workspace; the fixed scratch of "
                    "qsm_dipole_forward_noalloc holds %zu (use qsm_dipole_forward)",
                    grid->nx, grid->ny, grid->nz, spec_bytes, kScratchSpecBytes);
    unsigned char* base = nullptr;
    if ((rc = scratch_address(&base))) return rc;
    VolDesc* d_vols = reinterpret_cast<VolDesc*>(base);
    float2* d_tw = reinterpret_cast<float2*>(base + kScratchDescBytes);
    V.bx = bdir[0] / bn; V.by = bdir[1] / bn; V.bz = bdir[2] / bn;
    V.field = d_chi;
    V.adj_out = d_field;
    V.out = d_field;
    V.spec[SPEC_WORK] = reinterpret_cast<float2*>(base + kScratchSpecOff);
    hipStream_t st = (hipStream_t)stream;
    // The R2C pass writes the whole spectrum, so nothing needs clearing.
    QSM_HIP(hipMemcpyAsync(d_vols, &V, sizeof V, hipMemcpyHostToDevice, st));
    QSM_HIP(hipMemcpyAsync(d_tw, tb.host, tb.count * sizeof(float2), hipMemcpyHostToDevice, st));
    enqueue_round_trip(P_R2C_F64, P_Z_DIPOLE, V, pb, d_vols, d_tw, st);
    QSM_CHECK_LAUNCH();
    // the host-side tables above live on this stack frame: wait before returning
    QSM_HIP(hipStreamSynchronize(st));
    return QSM_OK;
}

// Host-array wrapper of a closed-form inversion (filter kind, parameter).
static int filter_host(int filt, double fa, const double* field, const uint8_t* mask,
                       const qsm_grid* grid, const double bdir[3], double* chi_out,
                       int src_kind = P_R2C_F64) {
    if (!field || !mask || !grid || (!bdir && src_kind == P_R2C_F64) || !chi_out)
        return fail(QSM_EPARAM, "null argument");
    int rc = validate_grid(grid);
    if (rc) return rc;
    const size_t N = grid->nx * grid->ny * grid->nz;
    double *d_f = nullptr, *d_o = nullptr;
    uint8_t* d_m = nullptr;
    auto done = [&](int code) {
        if (d_f) (void)hipFree(d_f);
        if (d_o) (void)hipFree(d_o);
        if (d_m) (void)hipFree(d_m);
        return code;
    };
    if ((rc = dev_alloc(&d_f, N)) || (rc = dev_alloc(&d_o, N)) || (rc = dev_alloc(&d_m, N))) return done(rc);
    hipError_t e = hipMemcpy(d_f, field, N * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_m, mask, N, hipMemcpyHostToDevice);
    if (e != hipSuccess) return done(fail(QSM_EHIP, "hipMemcpy H2D: %s", hipGetErrorString(e)));
    if ((rc = spectral_apply(P_Z_FILTER, filt, fa, d_f, d_m, grid, bdir, d_o, nullptr, src_kind)))
        return done(rc);
    e = hipMemcpy(chi_out, d_o, N * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return done(fail(QSM_EHIP, "hipMemcpy D2H: %s", hipGetErrorString(e)));
    return done(QSM_OK);
}

// laplacian_unwrap, src/unwrap/laplacian.rs:135-161: Poisson solve of the
// masked wrapped Laplacian, masked.
int qsm_laplacian_unwrap(const double* phase, const uint8_t* mask, const qsm_grid* grid, double* out) {
    g_err.clear();
    return filter_host(6, 0.0, phase, mask, grid, nullptr, out, P_R2C_WRAPLAP);
}

// nltv, src/inversion/nltv.rs:63-133, on one volume: TV's kernels with
// weighted thresholds, a weight pass and a boundary step between Newton
// rounds.  The host synchronises after every step when it needs to know
// about convergence (tol > 0) or report progress.
void qsm_nltv_params_default(qsm_nltv_params* p) {
    if (!p) return;  // src/inversion/nltv.rs:38-49
    p->lambda = 1e-3;
    p->mu = 1.0;
    p->tol = 1e-3;
    p->max_iter = 250;
    p->newton_iter = 10;
}

int qsm_nltv(const double* local_field, const uint8_t* mask, const qsm_grid* grid, const double bdir[3],
             const qsm_nltv_params* params, qsm_progress_fn progress, void* user, double* chi_out) {
    g_err.clear();
    if (!local_field || !mask || !grid || !bdir || !params || !chi_out)
        return fail(QSM_EPARAM, "null argument");
    if (!(params->lambda > 0)) return fail(QSM_EPARAM, "NLTV needs lambda > 0 (rho = 100 lambda)");
    int rc = validate_grid(grid);
    if (rc) return rc;
    const size_t N = grid->nx * grid->ny * grid->nz;
    double *d_f = nullptr, *d_chi = nullptr;
    uint8_t* d_m = nullptr;
    qsm_tv_batch* B = nullptr;
    std::vector<void*> extra;
    auto done = [&](int code) {
        if (B) batch_free(B);
        for (void* p : extra) (void)hipFree(p);
        if (d_f) (void)hipFree(d_f);
        if (d_chi) (void)hipFree(d_chi);
        if (d_m) (void)hipFree(d_m);
        return code;
    };
    if ((rc = dev_alloc(&d_f, N)) || (rc = dev_alloc(&d_chi, N)) || (rc = dev_alloc(&d_m, N)))
        return done(rc);
    hipError_t e = hipMemcpy(d_f, local_field, N * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_m, mask, N, hipMemcpyHostToDevice);
    if (e != hipSuccess) return done(fail(QSM_EHIP, "hipMemcpy H2D: %s", hipGetErrorString(e)));
    qsm_tv_volume vol;
    vol.grid = *grid;
    vol.bdir[0] = bdir[0]; vol.bdir[1] = bdir[1]; vol.bdir[2] = bdir[2];
    vol.d_field = d_f; vol.d_mask = d_m; vol.d_chi = d_chi;
    qsm_tv_params tp;
    tp.lambda = params->lambda;
    tp.rho = 100.0 * params->lambda;  // src/inversion/nltv.rs:75
    tp.tol = params->tol;
    tp.max_iter = params->max_iter;
    if ((rc = tv_batch_create_impl(&vol, 1, &tp, nullptr, &B))) return done(rc);
    VolDesc& V = B->h[0];
    V.filt_b = params->mu;
    const int nb = (int)((N + qsmfft::kThreads * 4 - 1) / (qsmfft::kThreads * 4));
    for (int q = 0; q < 2; ++q) {
        if ((rc = dev_alloc(&V.nltv_w[q], N))) return done(rc);
        extra.push_back(V.nltv_w[q]);
    }
    if ((rc = dev_alloc(&V.nltv_wnorm, 2))) return done(rc);
    extra.push_back(V.nltv_wnorm);
    if ((rc = dev_alloc(&V.nltv_wpart, (size_t)nb))) return done(rc);
    extra.push_back(V.nltv_wpart);
    hipStream_t st = B->stream;
    QSM_HIP(hipMemcpy(B->d_vols, B->h.data(), sizeof(VolDesc), hipMemcpyHostToDevice));
    std::memset(B->h_status, 0, sizeof(VolStatus));
    if ((rc = enqueue_setup(B))) return done(rc);
    launch_nltv_init(st, B->d_vols, B->lt_out);
    const bool sync_each = progress != nullptr || params->tol > 0.0;
    const size_t total = params->max_iter * params->newton_iter;
    size_t S = 0, current = 0;
    int cw = 0;
    hipEvent_t ev = B->ev[0];
    for (size_t newton = 0; newton < params->newton_iter; ++newton) {
        for (size_t it = 1; it <= params->max_iter; ++it) {
            ++current;
            if (progress) progress(current, total, user);
            ++S;
            PassArgs a = B->args(B->d_status);
            a.parity = (int)((S - 1) & 1);
            if (newton > 0 && it == 1) {
                launch_nltv_weights(st, B->d_vols, B->lt_out, a.parity, cw ^ 1);
                launch_nltv_wmax(st, B->d_vols, 1, nb, cw ^ 1);
                PassArgs b = a;
                b.status_rw = B->d_status;
                b.ridx = cw;
                run_groups(P_R2C_NLTV_BND, B->gx, b);
                launch_status_clear_done(st, B->d_status, 1);
                cw ^= 1;
            } else {
                a.ridx = cw;
                run_groups(P_R2C_TVDIV_W, B->gx, a);
            }
            run_groups(P_Y_FWD, B->gy, a);
            run_groups(P_Z_TV_SOLVE, B->gz, a);
            run_groups(P_Y_INV, B->gy, a);
            run_groups(P_C2R_TV, B->gxc, a);
            launch_tv_reduce(st, B->d_vols, 1, B->d_status, B->h_status_dev);
            QSM_CHECK_LAUNCH();
            if (sync_each) {
                QSM_HIP(hipEventRecord(ev, st));
                QSM_HIP(hipEventSynchronize(ev));
                const volatile VolStatus* hs = B->h_status;
                if (hs->done) break;  // src/inversion/nltv.rs:105-107
            }
        }
    }
    launch_tv_output(st, B->d_vols, B->lt_out, B->d_status);
    QSM_CHECK_LAUNCH();
    QSM_HIP(hipStreamSynchronize(st));
    e = hipMemcpy(chi_out, d_chi, N * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return done(fail(QSM_EHIP, "hipMemcpy D2H: %s", hipGetErrorString(e)));
    return done(QSM_OK);
}

void qsm_tkd_params_default(qsm_tkd_params* p) {
    if (p) p->threshold = 0.15;  // src/inversion/tkd.rs:29-33
}

void qsm_tikhonov_params_default(qsm_tikhonov_params* p) {
    if (!p) return;
    p->lambda = 0.01;  // src/inversion/tikhonov.rs:45-49
    p->reg = QSM_REG_IDENTITY;
}

void qsm_rts_params_default(qsm_rts_params* p) {
    if (!p) return;  // src/inversion/rts.rs:41-52
    p->delta = 0.15;
    p->mu = 1e5;
    p->rho = 10.0;
    p->tol = 1e-2;
    p->max_iter = 20;
    p->lsmr_iter = 4;
}

void qsm_sharp_params_default(qsm_sharp_params* p) {
    if (!p) return;  // src/bgremove/sharp.rs:38-45
    p->threshold = 0.05;
    p->radius = 6.0;
}

int qsm_tkd(const double* local_field, const uint8_t* mask, const qsm_grid* grid, const double bdir[3],
            const qsm_tkd_params* params, double* chi_out) {
    g_err.clear();
    if (!params || !(params->threshold > 0)) return fail(QSM_EPARAM, "TKD threshold must be > 0");
    return filter_host(0, params->threshold, local_field, mask, grid, bdir, chi_out);
}

int qsm_tsvd(const double* local_field, const uint8_t* mask, const qsm_grid* grid, const double bdir[3],
             const qsm_tkd_params* params, double* chi_out) {
    g_err.clear();
    if (!params || std::isnan(params->threshold)) return fail(QSM_EPARAM, "TSVD threshold is NaN");
    return filter_host(1, params->threshold, local_field, mask, grid, bdir, chi_out);
}

int qsm_tikhonov(const double* local_field, const uint8_t* mask, const qsm_grid* grid,
                 const double bdir[3], const qsm_tikhonov_params* params, double* chi_out) {
    g_err.clear();
    if (!params || params->reg < QSM_REG_IDENTITY || params->reg > QSM_REG_LAPLACIAN)
        return fail(QSM_EPARAM, "bad Tikhonov parameters");
    return filter_host(2 + (int)params->reg, params->lambda, local_field, mask, grid, bdir, chi_out);
}

int qsm_rts(const double* local_field, const uint8_t* mask, const qsm_grid* grid, const double bdir[3],
            const qsm_rts_params* params, qsm_progress_fn progress, void* user, double* chi_out) {
    g_err.clear();
    if (!local_field || !mask || !grid || !bdir || !params || !chi_out)
        return fail(QSM_EPARAM, "null argument");
    int rc = validate_grid(grid);
    if (rc) return rc;
    const size_t N = grid->nx * grid->ny * grid->nz;
    double *d_f = nullptr, *d_chi = nullptr;
    uint8_t* d_m = nullptr;
    qsm_tv_batch* B = nullptr;
    auto done = [&](int code) {
        if (B) batch_free(B);
        if (d_f) (void)hipFree(d_f);
        if (d_chi) (void)hipFree(d_chi);
        if (d_m) (void)hipFree(d_m);
        return code;
    };
    if ((rc = dev_alloc(&d_f, N)) || (rc = dev_alloc(&d_chi, N)) || (rc = dev_alloc(&d_m, N)))
        return done(rc);
    hipError_t e = hipMemcpy(d_f, local_field, N * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_m, mask, N, hipMemcpyHostToDevice);
    if (e != hipSuccess) return done(fail(QSM_EHIP, "hipMemcpy H2D: %s", hipGetErrorString(e)));
    qsm_tv_volume vol;
    vol.grid = *grid;
    vol.bdir[0] = bdir[0]; vol.bdir[1] = bdir[1]; vol.bdir[2] = bdir[2];
    vol.d_field = d_f; vol.d_mask = d_m; vol.d_chi = d_chi;
    if ((rc = tv_batch_create_impl(&vol, 1, nullptr, nullptr, &B, params))) return done(rc);
    if ((rc = tv_batch_run_impl(B, progress, user))) return done(rc);
    e = hipMemcpy(chi_out, d_chi, N * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return done(fail(QSM_EHIP, "hipMemcpy D2H: %s", hipGetErrorString(e)));
    return done(QSM_OK);
}

int qsm_rts_batch_create(const qsm_tv_volume* vols, size_t n_vols, const qsm_rts_params* params,
                         void* stream, qsm_tv_batch** out) {
    g_err.clear();
    if (!params) return fail(QSM_EPARAM, "null argument");
    return tv_batch_create_impl(vols, n_vols, nullptr, stream, out, params);
}

int qsm_sharp(const double* field, const uint8_t* mask, const qsm_grid* grid,
              const qsm_sharp_params* params, double* local_out, uint8_t* mask_out) {
    g_err.clear();
    if (!params) return fail(QSM_EPARAM, "null argument");
    const double r = params->radius;
    return vsharp_host(field, mask, grid, &r, 1, params->threshold, nullptr, nullptr, local_out,
                       mask_out);
}

void qsm_inversion_config_default(qsm_inversion_config* c) {
    if (!c) return;  // src/pipeline/config.rs:248-264
    c->algorithm = QSM_INV_RTS;
    qsm_tkd_params_default(&c->tkd);
    qsm_tkd_params_default(&c->tsvd);
    qsm_tikhonov_params_default(&c->tikhonov);
    qsm_tv_params_default(&c->tv);
    qsm_rts_params_default(&c->rts);
    qsm_nltv_params_default(&c->nltv);
}

// run_dipole_inversion, src/pipeline/inversion.rs:28-114, for the arms with
// a GPU backend here (Tkd, Tsvd, Tikhonov, Tv, Rts, Nltv).
int qsm_run_dipole_inversion_cfg(const double* local_field_ppm, size_t field_len, const uint8_t* mask,
                                 size_t mask_len, const qsm_scan_metadata* md,
                                 const qsm_inversion_config* cfg, qsm_progress_fn progress, void* user,
                                 double* chi_out) {
    g_err.clear();
    if (!md || !cfg) return fail(QSM_EPARAM, "metadata or config is null");
    const size_t n = md->nx * md->ny * md->nz;
    if (field_len != n) return fail(QSM_EDIM, "dimension mismatch: expected %zu, got %zu", n, field_len);
    if (mask_len != n) return fail(QSM_EDIM, "dimension mismatch: expected %zu, got %zu", n, mask_len);
    qsm_grid g{md->nx, md->ny, md->nz, md->vsx, md->vsy, md->vsz};
    const double* b = md->b0_direction;
    switch (cfg->algorithm) {
        case QSM_INV_TKD: return qsm_tkd(local_field_ppm, mask, &g, b, &cfg->tkd, chi_out);
        case QSM_INV_TSVD: return qsm_tsvd(local_field_ppm, mask, &g, b, &cfg->tsvd, chi_out);
        case QSM_INV_TIKHONOV: return qsm_tikhonov(local_field_ppm, mask, &g, b, &cfg->tikhonov, chi_out);
        case QSM_INV_TV: return qsm_tv_admm(local_field_ppm, mask, &g, b, &cfg->tv, progress, user, chi_out);
        case QSM_INV_RTS: return qsm_rts(local_field_ppm, mask, &g, b, &cfg->rts, progress, user, chi_out);
        case QSM_INV_NLTV: return qsm_nltv(local_field_ppm, mask, &g, b, &cfg->nltv, progress, user, chi_out);
        case QSM_INV_TGV:
        case QSM_INV_QSMART:
            return fail(QSM_ECONFIG, "invalid config: %s should use run_tgv or run_qsmart",
                        cfg->algorithm == QSM_INV_TGV ? "Tgv" : "Qsmart");
        default:
            return fail(QSM_ECONFIG, "invalid config: algorithm %d has no GPU backend here",
                        (int)cfg->algorithm);
    }
}

int qsm_vsharp_batch_create(const qsm_vsharp_volume* vols, size_t n_vols, const double* radii,
                            size_t n_radii, double threshold, void* stream, qsm_vsharp_batch** out) {
    g_err.clear();
    return vs_create_impl(vols, n_vols, radii, n_radii, threshold, stream, out);
}

int qsm_vsharp_batch_run(qsm_vsharp_batch* batch, qsm_progress_fn progress, void* user) {
    g_err.clear();
    return vs_run_impl(batch, progress, user);
}

void qsm_vsharp_batch_destroy(qsm_vsharp_batch* batch) { vs_free(batch); }

int qsm_vsharp_with_radii(const double* field, const uint8_t* mask, const qsm_grid* grid,
                          const double* radii, size_t n_radii, double threshold,
                          qsm_progress_fn progress, void* user, double* local_out,
                          uint8_t* mask_out) {
    g_err.clear();
    return vsharp_host(field, mask, grid, radii, n_radii, threshold, progress, user, local_out,
                       mask_out);
}

int qsm_vsharp(const double* field, const uint8_t* mask, const qsm_grid* grid,
               const qsm_vsharp_params* params, qsm_progress_fn progress, void* user,
               double* local_out, uint8_t* mask_out) {
    g_err.clear();
    if (!params) return fail(QSM_EPARAM, "params is null");
    std::vector<double> radii;
    int rc = vsharp_radii(params, radii);
    if (rc) return rc;
    return vsharp_host(field, mask, grid, radii.data(), radii.size(), params->threshold, progress,
                       user, local_out, mask_out);
}

int qsm_run_bg_removal(const double* field_ppm, size_t field_len, const uint8_t* mask,
                       size_t mask_len, const qsm_scan_metadata* md,
                       qsm_bg_removal_algorithm algorithm, const qsm_vsharp_params* vsharp_params,
                       qsm_progress_fn progress, void* user, double* local_out, uint8_t* mask_out) {
    g_err.clear();
    if (!md) return fail(QSM_EPARAM, "metadata is null");
    const size_t n = md->nx * md->ny * md->nz;
    // src/pipeline/bg_removal.rs:30-35
    if (field_len != n) return fail(QSM_EDIM, "dimension mismatch: expected %zu, got %zu", n, field_len);
    if (mask_len != n) return fail(QSM_EDIM, "dimension mismatch: expected %zu, got %zu", n, mask_len);
    qsm_bg_removal_config cfg;
    qsm_bg_removal_config_default(&cfg);
    cfg.algorithm = algorithm;
    if (vsharp_params) cfg.vsharp = *vsharp_params;
    return qsm_run_bg_removal_cfg(field_ppm, field_len, mask, mask_len, md, &cfg, progress, user,
                                  local_out, mask_out);
}

void qsm_bg_removal_config_default(qsm_bg_removal_config* c) {
    if (!c) return;  // src/pipeline/config.rs:215-229
    c->algorithm = QSM_BG_VSHARP;
    qsm_vsharp_params_default(&c->vsharp);
    qsm_sharp_params_default(&c->sharp);
}

// run_bg_removal, src/pipeline/bg_removal.rs:20-92, for the Vsharp and Sharp
// arms.
int qsm_run_bg_removal_cfg(const double* field_ppm, size_t field_len, const uint8_t* mask,
                           size_t mask_len, const qsm_scan_metadata* md,
                           const qsm_bg_removal_config* cfg, qsm_progress_fn progress, void* user,
                           double* local_out, uint8_t* mask_out) {
    g_err.clear();
    if (!md || !cfg) return fail(QSM_EPARAM, "metadata or config is null");
    const size_t n = md->nx * md->ny * md->nz;
    if (field_len != n) return fail(QSM_EDIM, "dimension mismatch: expected %zu, got %zu", n, field_len);
    if (mask_len != n) return fail(QSM_EDIM, "dimension mismatch: expected %zu, got %zu", n, mask_len);
    qsm_grid g{md->nx, md->ny, md->nz, md->vsx, md->vsy, md->vsz};
    switch (cfg->algorithm) {
        case QSM_BG_VSHARP:
            return qsm_vsharp(field_ppm, mask, &g, &cfg->vsharp, progress, user, local_out, mask_out);
        case QSM_BG_SHARP:
            return qsm_sharp(field_ppm, mask, &g, &cfg->sharp, local_out, mask_out);
        default:
            return fail(QSM_ECONFIG, "invalid config: algorithm %d has no GPU backend here",
                        (int)cfg->algorithm);
    }
}

// Best rate of `reps` timed launches of `launch` moving `bytes` (GB/s),
// after one untimed launch.
extern "C++" {
template <class F>
static hipError_t best_rate(int reps, double bytes, F launch, double& best) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) { launch(); e = hipGetLastError(); }
    best = 0.0;
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        float ms = 0.f;
        if ((e = hipEventRecord(e0, nullptr)) != hipSuccess) break;
        launch();
        if ((e = hipGetLastError()) != hipSuccess) break;
        if ((e = hipEventRecord(e1, nullptr)) != hipSuccess) break;
        if ((e = hipEventSynchronize(e1)) != hipSuccess) break;
        if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
        const double g = bytes / (ms * 1e-3) / 1e9;
        if (g > best) best = g;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return e;
}
}  // extern "C++"

// Device-to-device copy rate of a plain one-element-per-thread kernel over
// `bytes` (read + write counted), best of `reps` timed launches.
int qsm_hbm_copy_probe(size_t bytes, int reps, double* gbps) {
    g_err.clear();
    if (!gbps || bytes < 16 || reps < 1) return fail(QSM_EPARAM, "bad probe arguments");
    const size_t n = bytes / sizeof(float2);
    float2 *a = nullptr, *b = nullptr;
    int rc;
    if ((rc = dev_alloc(&a, n)) || (rc = dev_alloc(&b, n))) {
        if (a) (void)hipFree(a);
        return rc;
    }
    launch_fill_zero(nullptr, reinterpret_cast<float*>(a), 2 * n);
    double best = 0.0;
    const hipError_t e = best_rate(reps, 2.0 * (double)n * sizeof(float2),
                                   [&] { launch_copy_probe(nullptr, a, b, n); }, best);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(QSM_EHIP, "copy probe: %s", hipGetErrorString(e));
    *gbps = best;
    return QSM_OK;
}

// Rate of a pass shape: `reads` arrays of n floats read and `writes` arrays
// written element by element (1..4 each), all bytes counted.
int qsm_hbm_stream_probe(size_t n, int reads, int writes, int reps, double* gbps) {
    g_err.clear();
    if (!gbps || n < 4096 || reps < 1 || reads < 1 || reads > 4 || writes < 1 || writes > 4)
        return fail(QSM_EPARAM, "bad probe arguments");
    float *in = nullptr, *out = nullptr;
    int rc;
    if ((rc = dev_alloc(&in, n * reads)) || (rc = dev_alloc(&out, n * writes))) {
        if (in) (void)hipFree(in);
        return rc;
    }
    launch_fill_zero(nullptr, in, n * reads);
    double best = 0.0;
    const hipError_t e = best_rate(reps, 4.0 * (double)n * (reads + writes),
                                   [&] { launch_stream_probe(nullptr, reads, writes, in, out, n); }, best);
    (void)hipFree(in);
    (void)hipFree(out);
    if (e != hipSuccess) return fail(QSM_EHIP, "stream probe: %s", hipGetErrorString(e));
    *gbps = best;
    return QSM_OK;
}

int qsm_device_alloc(size_t bytes, void** out) {
    g_err.clear();
    if (!out) return fail(QSM_EPARAM, "null output pointer");
    *out = nullptr;
    QSM_HIP(hipMalloc(out, bytes ? bytes : 1));
    return QSM_OK;
}

int qsm_device_free(void* ptr) {
    g_err.clear();
    if (ptr) QSM_HIP(hipFree(ptr));
    return QSM_OK;
}

int qsm_memcpy_htod(void* dst, const void* src, size_t bytes) {
    g_err.clear();
    if ((!dst || !src) && bytes) return fail(QSM_EPARAM, "null pointer");
    if (bytes) QSM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return QSM_OK;
}

int qsm_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    g_err.clear();
    if ((!dst || !src) && bytes) return fail(QSM_EPARAM, "null pointer");
    if (bytes) QSM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return QSM_OK;
}

int qsm_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    g_err.clear();
    if ((!dst || !src) && bytes) return fail(QSM_EPARAM, "null pointer");
    if (bytes) QSM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return QSM_OK;
}

int qsm_device_synchronize(void) {
    g_err.clear();
    QSM_HIP(hipDeviceSynchronize());
    return QSM_OK;
}

}  // extern "C"
