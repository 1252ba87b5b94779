// Host side of the C ABI: device context, resident agents, the MACE solve
// loop and the stateless operators.  Compiled with hipcc (host code only).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <unistd.h>

#include "mtr.h"
#include "mtr_kernels.h"

using namespace mtr;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? MTR_ERR_OOM : MTR_ERR_HIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                   \
    } while (0)

#define MTR_TRY(expr)             \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != MTR_OK) return rc_; \
    } while (0)

size_t elem_size(int prec) { return prec == MTR_PREC_FP32 ? 4 : 2; }

bool vol_valid(const mtr_volume_size *v)
{
    return v && v->X > 0 && v->Y > 0 && v->Z > 0 && v->Xend - v->Xstart == v->X &&
           v->Yend - v->Ystart == v->Y && v->Zend - v->Zstart == v->Z;
}

// Index of the first NaN or Inf in a[0..n) or -1: an all-ones exponent,
// tested on the bits so that the scan vectorises.
int64_t first_nonfinite(const float *a, int64_t n)
{
    const uint32_t *u = reinterpret_cast<const uint32_t *>(a);
    constexpr int64_t kBlock = 1 << 16;
    for (int64_t b = 0; b < n; b += kBlock) {
        const int64_t e = std::min(n, b + kBlock);
        uint32_t bad = 0;
        for (int64_t i = b; i < e; i++) bad |= (u[i] & 0x7f800000u) == 0x7f800000u;
        if (bad)
            for (int64_t i = b; i < e; i++)
                if ((u[i] & 0x7f800000u) == 0x7f800000u) return i;
    }
    return -1;
}

bool prj_valid(const mtr_projection_size *p)
{
    return p && p->X > 0 && p->Y > 0 && p->AngN > 0 && p->AngN <= 100000;
}

// Builds the ray-driven forward-projection plan of one view.
int make_plan(const double *xc, const double *yc, FpPlan &pl)
{
    std::memset(&pl, 0, sizeof pl);
    const double a[3] = {xc[1], xc[2], xc[3]};
    const double b[3] = {yc[1], yc[2], yc[3]};
    if (!std::isfinite(xc[0]) || !std::isfinite(yc[0])) return MTR_ERR_NONFINITE;
    for (int m = 0; m < 3; m++)
        if (!std::isfinite(a[m]) || !std::isfinite(b[m])) return MTR_ERR_NONFINITE;
    // Row-separable: v = yc0 + bW*W - 1 with bW = +-1, the other b zero and
    // yc0 integral -> every integer voxel lands on an integer row.
    int nzb = 0, W = -1;
    for (int m = 0; m < 3; m++)
        if (b[m] != 0.0) {
            nzb++;
            W = m;
        }
    if (nzb == 1 && std::fabs(b[W]) == 1.0 && std::floor(yc[0]) == yc[0] &&
        std::fabs(yc[0]) < 1e9) {
        const int A1 = W == 0 ? 1 : 0, A2 = W == 2 ? 1 : 2;
        // Line (drive) axis: the in-plane axis with the smaller |a|.
        int L = A1, M = A2;
        if (std::fabs(a[A1]) > std::fabs(a[A2])) {
            L = A2;
            M = A1;
        }
        const double EmM = a[M] != 0.0 ? 1.0 / std::fabs(a[M]) : 0.0;
        if (a[M] != 0.0 && std::floor(2.0 * EmM) + 1 <= 64) {
            pl.kind = 1;
            pl.W = W;
            pl.L = L;
            pl.M = M;
            pl.wsign = b[W] > 0 ? 1 : -1;
            pl.woff = (1.0 - yc[0]) * pl.wsign;
            pl.m0 = (1.0 - xc[0]) / a[M];
            pl.mi = 1.0 / a[M];
            pl.mw = -a[W] / a[M];
            pl.ml = -a[L] / a[M];
            pl.Em = 1.0 / std::fabs(a[M]);
            pl.ncand = (int)std::floor(2.0 * pl.Em) + 1;
            return MTR_OK;
        }
    }
    // General plan: driving axis = largest component of grad u x grad v.
    const double d[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    int D = 0;
    for (int m = 1; m < 3; m++)
        if (std::fabs(d[m]) > std::fabs(d[D])) D = m;
    if (!(std::fabs(d[D]) > 0.0)) return MTR_ERR_SHAPE;
    const int P = D == 0 ? 1 : 0, Q = D == 2 ? 1 : 2;
    const double det = a[P] * b[Q] - a[Q] * b[P];
    if (!(std::fabs(det) > 0.0)) return MTR_ERR_SHAPE;
    const double m00 = b[Q] / det, m01 = -a[Q] / det, m10 = -b[P] / det, m11 = a[P] / det;
    pl.kind = 0;
    pl.D = D;
    pl.P = P;
    pl.Q = Q;
    pl.p0 = m00 * (1.0 - xc[0]) + m01 * (1.0 - yc[0]);
    pl.pi = m00;
    pl.pk = m01;
    pl.pd = -(m00 * a[D] + m01 * b[D]);
    pl.q0 = m10 * (1.0 - xc[0]) + m11 * (1.0 - yc[0]);
    pl.qi = m10;
    pl.qk = m11;
    pl.qd = -(m10 * a[D] + m11 * b[D]);
    pl.Ep = std::fabs(m00) + std::fabs(m01);
    pl.Eq = std::fabs(m10) + std::fabs(m11);
    return MTR_OK;
}

// Column-plan forward projection: the eight waves of a block take eight
// views, one each, and meet at a barrier per round of columns; a view costs
// about the pixels a column reaches in it, |xc3| (Z - 1) + 3.  Views sorted
// by that count (ties in view order) make each block's views alike, so no
// wave waits long at the barrier for a costlier one.  pfx_view[g] is the
// view of sorted slot g (the sums of every view are unchanged).
void pfx_view_order(const double *xc, int A, int nz, std::vector<FpPlan> &plan)
{
    std::vector<int> ord(A);
    for (int v = 0; v < A; v++) ord[v] = v;
    auto npix = [&](int v) { return (int)(std::fabs(xc[4 * v + 3]) * (double)(nz - 1)) + 3; };
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return npix(a) > npix(b); });
    for (int g = 0; g < A; g++) plan[g].pfx_view = ord[g];
}

// Back-projection mode of an agent: the run-along-W kernel applies when
// every view is row-separable along the same axis W, u does not depend on
// W, and the tile footprint fits the fixed LDS pitch.
struct BpMode {
    bool sep = false;
    int W = 0, P1 = 0, P2 = 0;
    int rw = 32;  // voxels per thread along W
};

BpMode bp_mode(const std::vector<FpPlan> &plan, const double *xc, const double *yc, const mtr_volume_size &vol,
                const mtr_projection_size &prj)
{
    BpMode m;
    if (plan.empty()) return m;
    const int W = plan[0].W;
    for (size_t v = 0; v < plan.size(); v++)
        if (plan[v].kind != 1 || plan[v].W != W || xc[4 * v + 1 + W] != 0.0) return m;
    const int P1 = W == 0 ? 1 : 0, P2 = 3 - W - P1;
    const int lo[3] = {vol.Xstart, vol.Ystart, vol.Zstart}, hi[3] = {vol.Xend - 1, vol.Yend - 1, vol.Zend - 1};
    for (size_t v = 0; v < plan.size(); v++) {
        const double w = std::fabs(xc[4 * v + 1 + P1]) * 63 + std::fabs(xc[4 * v + 1 + P2]) * (bp_sep_waves() - 1) + 5;
        if (w > bp_sep_pitch()) return m;
        // every plane of the box must land on a detector row
        const double k0 = yc[4 * v] + yc[4 * v + 1 + W] * lo[W] - 1, k1 = yc[4 * v] + yc[4 * v + 1 + W] * hi[W] - 1;
        if (std::min(k0, k1) < 0 || std::max(k0, k1) > prj.Y - 1) return m;
    }
    m.sep = true;
    m.W = W;
    m.P1 = P1;
    m.P2 = P2;
    return m;
}

std::vector<BpSepView> bp_sep_views(const BpMode &m, const double *xc, const double *yc, int A)
{
    std::vector<BpSepView> out(m.sep ? A : 0);
    const int waves = bp_sep_waves();
    for (int v = 0; v < (int)out.size(); v++) {
        BpSepView &b = out[v];
        b.c0 = xc[4 * v] - 1.0;
        b.a1 = xc[4 * v + 1 + m.P1];
        b.a2 = xc[4 * v + 1 + m.P2];
        b.umin = std::min(0.0, b.a1 * 63) + std::min(0.0, b.a2 * (waves - 1));
        b.umax = std::max(0.0, b.a1 * 63) + std::max(0.0, b.a2 * (waves - 1));
        b.k0 = (int)(yc[4 * v] - 1.0);
        b.kstep = (int)yc[4 * v + 1 + m.W];
    }
    return out;
}

// Forward-projection streaming mode: every view row-separable along the
// same axis W with u independent of W.  Such an agent keeps its working
// volume with W as the contiguous axis, so the 64 lanes of a wave (64
// adjacent planes) read one coalesced row.  Returns the candidate count to
// instantiate (0 = not applicable) and the plane axis.
int fp_stream_nc(const std::vector<FpPlan> &plan, const double *xc, int *W_out)
{
    if (plan.empty() || plan[0].kind != 1) return 0;
    const int W = plan[0].W;
    int nc = 0;
    for (size_t v = 0; v < plan.size(); v++) {
        if (plan[v].kind != 1 || plan[v].W != W || xc[4 * v + 1 + W] != 0.0) return 0;
        nc = std::max(nc, plan[v].ncand);
    }
    if (nc < 2 || nc > 4) return 0;
    *W_out = W;
    return nc;
}

// Column plan: every view has v independent of Z (yc3 = 0) and the same
// v coefficients, so each (X, Y) column keeps its detector rows k0, k0 + 1
// and weights in all views.  The back-projection tile (64 x 8 columns, runs
// of rw along Z) must stage its rows times its u range within the LDS
// budget; lanes go along the axis that gives the smaller footprint.
struct ColMode {
    bool on = false;
    int p1y = 0;    // lanes along Y
    int pitch = 0;  // staged pairs per row (8 mod 16 where it fits, else odd)
    int rw = 32;
};

ColMode col_mode(const double *xc, const double *yc, int A, const mtr_projection_size &prj, int rw,
                 int only_p1y = -1)
{
    ColMode m;
    if (prj.X > fp_col_max_nu()) return m;
    for (int v = 0; v < A; v++)
        for (int q = 0; q < 4; q++)
            if (yc[4 * v + q] != yc[q]) return m;
    if (yc[3] != 0.0 || (yc[1] == 0.0 && yc[2] == 0.0)) return m;
    const int waves = bp_col_waves();
    int best = -1;
    for (int p1y = 0; p1y < 2; p1y++) {
        if (only_p1y >= 0 && p1y != only_p1y) continue;
        const double r1 = std::fabs(yc[1 + p1y]), r2 = std::fabs(yc[2 - p1y]);
        const int rows = (int)std::floor(63 * r1 + (waves - 1) * r2) + 3;
        double wmax = 0.0;
        for (int v = 0; v < A; v++) {
            const double w = 63 * std::fabs(xc[4 * v + 1 + p1y]) + (waves - 1) * std::fabs(xc[4 * v + 2 - p1y]) +
                             (rw - 1) * std::fabs(xc[4 * v + 3]);
            wmax = std::max(wmax, w);
        }
        // pitch = 8 (mod 16) pairs: a lane one row further reads 8 bank
        // pairs away, so the lanes of a wave, spread over rows and columns,
        // meet on fewer LDS banks (3% faster than an odd pitch at 45 and 60
        // deg; 1 (mod 16) was 2.3x slower at 45); odd where that does not fit
        int pitch = (int)std::floor(wmax) + 4;
        while (pitch % 16 != 8) pitch++;
        if (rows * pitch > col_elems()) {
            pitch = (int)std::floor(wmax) + 4;
            if (pitch % 2 == 0) pitch++;
        }
        const int elems = rows * pitch;
        if (elems <= col_elems() && (best < 0 || elems < best)) {
            best = elems;
            m.on = true;
            m.p1y = p1y;
            m.pitch = pitch;
            m.rw = rw;
        }
    }
    return m;
}

// Per-view u extremes of a column-plan back-projection tile relative to
// its origin (the tile spans 64 lanes along P1, the waves along P2 and rw
// planes along Z).
std::vector<BpSepView> bp_col_views(const ColMode &m, const double *xc, int A)
{
    std::vector<BpSepView> out(A);
    const int waves = bp_col_waves();
    for (int v = 0; v < A; v++) {
        BpSepView &b = out[v];
        std::memset(&b, 0, sizeof b);
        const double e[3] = {m.p1y ? (double)(waves - 1) : 63.0, m.p1y ? 63.0 : (double)(waves - 1), (double)(m.rw - 1)};
        for (int q = 0; q < 3; q++) {
            const double d = xc[4 * v + 1 + q] * e[q];
            b.umin += std::min(0.0, d);
            b.umax += std::max(0.0, d);
        }
        b.c0 = xc[4 * v];
    }
    return out;
}

// Element strides of a volume whose contiguous axis is `fast` (-1 or 0:
// the reference's [Z][Y][X]); the other two axes keep X-before-Y-before-Z
// order.
void layout_strides(int fast, int X, int Y, int Z, int64_t *st)
{
    if (fast == 1) {  // [Z][X][Y]
        st[1] = 1;
        st[0] = Y;
        st[2] = (int64_t)X * Y;
    } else if (fast == 2) {  // [Y][X][Z]
        st[2] = 1;
        st[0] = Z;
        st[1] = (int64_t)X * Z;
    } else {  // [Z][Y][X]
        st[0] = 1;
        st[1] = X;
        st[2] = (int64_t)X * Y;
    }
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

struct Agent {
    mtr_projection_size prj{};
    std::vector<double> xc, yc;
    bool lanes_k = false;  // forward projection with lanes along detector rows
    BpMode bpm;            // back-projection kernel choice
    int fp_nc = 0;         // > 0: streaming forward projection with this candidate count
    int fp_g = 1;          // streaming forward projection: detector columns per wave (1, 2)
    int fp_seg = 1, fp_seg_len = 0;  // streaming forward projection: Z segments and their planes
    float *d_seg = nullptr;          // per-segment partial ray sums [fp_seg][P] (fp_seg > 1)
    int *d_seg_ord = nullptr;        // per segment, (view, column group) pairs in crossing order
    int npartial = 0;      // per-block partials written by the last FP launch
    size_t dev_bytes = 0;  // device bytes the agent holds
    size_t dev_peak = 0;   // and at most held while it was added (staging included)
    int fast = 0;          // contiguous axis of the working volume (0 = X, canonical)
    int64_t st[3] = {0, 0, 0};
    void *d_tmp = nullptr; // canonical copy of the working volume (fast != 0)
    ViewCoef *d_coef = nullptr;
    FpPlan *d_plan = nullptr;
    BpSepView *d_bsv = nullptr;
    // Row-separable agents on the streaming FP and separable BP keep p, c
    // and r row-fastest ([A][nu][nv]): the 64 lanes of both kernels run
    // along detector rows.
    bool rowfast = false;
    void *d_p = nullptr;
    bool p_owned = true;
    float *d_p_std = nullptr;  // [A][nv][nu] copy of p for the plain form (rowfast only)
    ColMode col;               // column plan (oblique in-plane axes)
    bool tuned = false;        // back-projection tiles chosen by the autotuner
    int col_win = 0, col_jit = 0, col_span = 0;  // column FP row window (slots, power of two; 0: whole width)
    bool col_pfx = false;      // column FP from prefix moments (k_fp_pfx) rather than the segmented scan
    int2 *d_col_xy = nullptr;
    float2 *d_col_w = nullptr;
    int *d_strip = nullptr;
    float *d_pm = nullptr, *d_pn = nullptr;
    float *d_bpc = nullptr;    // column plan: BP normaliser per voxel
    bool col_cs = false;       // column plan: back-projection on strip tiles (k_bp_cs)
    ColGroup *d_col_grp = nullptr;
    int2 *d_cs_win = nullptr;  // strip-tiled BP: footprint window per block and view
    int2 *d_sep_win = nullptr;  // separable BP: footprint window per tile and view
    int ngrp = 0;
    float *d_c = nullptr;
    void *d_r = nullptr;
    void *d_x = nullptr;
    void *d_w = nullptr;
    double *d_partial = nullptr;
    double *d_scalar = nullptr;  // reduced ||r||^2
    int nfp = 0;
    bool c_ready = false;
    int64_t P() const { return (int64_t)prj.X * prj.Y * prj.AngN; }
};

}  // namespace

struct mtr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;  // false: the stream belongs to a server
    // A server's requests take their device memory from the server's pool,
    // ordered on the context's stream, so freeing it waits for no other
    // work on the device (hipFree does); other contexts use hipMalloc.
    hipMemPool_t pool = nullptr;
    // A second stream for the SIRT iterations of the agents that are not
    // column-plan ones, forked from and joined to `stream` by events.
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool kind_streams = true;
    int precision = MTR_PREC_FP32;
    mtr_volume_size vol{};
    bool vol_set = false;
    int64_t N = 0;
    void *d_wbar = nullptr;
    std::vector<Agent> agents;
    AgentDev *d_ags = nullptr;  // table in agent order
    // the same entries ordered by forward- and by back-projection kernel
    // kind, so that a whole-context SIRT iteration launches each kind once
    AgentDev *d_ags_fp = nullptr, *d_ags_bp = nullptr;
    std::vector<int> fp_order, bp_order;
    double *d_cons_partial = nullptr;
    double *d_cons_scalar = nullptr;
    int cons_blocks = 0;
    // plain form scratch
    float *d_s = nullptr, *d_cc = nullptr;
    int64_t sc_cap = 0;
    bool autotune = true;
    bool tune_timed = true;  // false: tiles from the tuning cache only, never timed here
    bool joint_tune = true;  // the joint run-length stage after the initial model
    std::string joint_key;   // agents it last ran for
    // graph of one fast MACE step, replayed for every step of a solve
    bool use_graphs = true;
    hipGraphExec_t step_graph = nullptr;
    int graph_S = -1;
    float graph_lambda = 0.f, graph_rho = 0.f;
    // tuning
    int bp_tz = 8;
    int lds_cap = 6144;
    // timing
    bool timing = false;
    struct Pending {
        int kernel;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    int64_t launches[MTR_K_COUNT] = {0};
    double ms[MTR_K_COUNT] = {0};
};

namespace {

// Device memory a context owns (see mtr_ctx::pool).
hipError_t dev_alloc(const mtr_ctx *ctx, void **p, size_t bytes)
{
    return ctx->pool ? hipMallocFromPoolAsync(p, bytes, ctx->pool, ctx->stream) : hipMalloc(p, bytes);
}

void dev_free(const mtr_ctx *ctx, void *p)
{
    if (!p) return;
    if (ctx->pool)
        (void)hipFreeAsync(p, ctx->stream);
    else
        (void)hipFree(p);
}

// A context's temporary, freed on scope exit.
struct CtxBuf {
    const mtr_ctx *ctx;
    void *p = nullptr;
    ~CtxBuf() { dev_free(ctx, p); }
};

int set_device(mtr_ctx *ctx)
{
    HIP_TRY(hipSetDevice(ctx->device));
    return MTR_OK;
}

hipEvent_t take_event(mtr_ctx *ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets one launch with events when timing is on.
struct Timed {
    mtr_ctx *ctx;
    int kernel;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(mtr_ctx *c, int k, bool on = true, hipStream_t s = nullptr) : ctx(c), kernel(k), st(s ? s : c->stream)
    {
        if (ctx->timing && on) {
            a = take_event(ctx);
            b = take_event(ctx);
            if (a && b) (void)hipEventRecord(a, st);
        }
    }
    ~Timed()
    {
        if (a && b) {
            (void)hipEventRecord(b, st);
            ctx->pending.push_back({kernel, a, b});
        }
    }
};

int collect_timing(mtr_ctx *ctx)
{
    if (ctx->pending.empty()) return MTR_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (auto &pd : ctx->pending) {
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, pd.a, pd.b));
        ctx->launches[pd.kernel]++;
        ctx->ms[pd.kernel] += t;
        ctx->event_pool.push_back(pd.a);
        ctx->event_pool.push_back(pd.b);
    }
    ctx->pending.clear();
    return MTR_OK;
}

// Planes per Z segment of the streaming FP (0: whole rays).
int fp_seg_planes()
{
    int len = 128;
    if (const char *g = std::getenv("MTR_FP_STREAM_SEG")) len = std::max(0, std::atoi(g));
    return len;
}

// Device bytes of the streaming FP's Z segments for an agent of P pixels
// and A views over nz planes, should it take that kernel: the block order
// of each segment (column groups of at least 4 columns) and the partial ray
// sums of a segmented box.
size_t fp_seg_bytes(size_t P, const mtr_projection_size &prj, int nz)
{
    const int len = fp_seg_planes();
    const size_t nseg = len > 0 && nz >= 2 * len ? (size_t)((nz + len - 1) / len) : 1;
    return 4 * nseg * (size_t)prj.AngN * (size_t)((prj.X + 3) / 4) + (nseg > 1 ? 4 * P * nseg : 0);
}

// Upper bound of one agent's device bytes over the kernel choices: geometry
// tables, p (plus a converted copy), c, r, the partial rows of the column
// plan, x, w and a canonical copy of x, the column plan's back-projection
// normaliser and tables, FP64 partial sums, the streaming FP's segments.
size_t agent_bytes(int precision, const mtr_volume_size &vol, const mtr_projection_size &prj)
{
    const size_t es = elem_size(precision);
    const size_t P = (size_t)prj.X * prj.Y * prj.AngN;
    const size_t N = (size_t)vol.X * vol.Y * vol.Z;
    const size_t cols = (size_t)vol.X * vol.Y;
    const size_t nfp = (size_t)((prj.X + 63) / 64 + 16) * ((prj.Y + 3) / 4 + 16) * prj.AngN;
    return (sizeof(ViewCoef) + sizeof(FpPlan) + sizeof(BpSepView)) * prj.AngN + 2 * 4 * P + 4 * P + es * P +
           2 * 4 * P + 3 * es * N + 4 * N + 12 * cols + 4 * (size_t)(prj.Y + 2) + 8 * nfp +
           sizeof(ColGroup) * (cols / 64 + (size_t)prj.Y + 2) + fp_seg_bytes(P, prj, vol.Z) +
           // strip-tiled BP windows: groups x blocks of 128 planes (one more
           // for a box off the 16-plane grid) x views; separable BP windows:
           // tiles of 64 x 8 lines x views
           sizeof(int2) * (cols / 64 + (size_t)prj.Y + 2) * ((size_t)vol.Z / 128 + 2) * prj.AngN +
           sizeof(int2) * ((size_t)std::max({vol.X, vol.Y, vol.Z}) / 64 + 1) *
               ((size_t)std::max({vol.X, vol.Y, vol.Z}) / 8 + 1) * prj.AngN;
}

// Host-to-device copy of a host buffer that is freed after the call,
// ordered on the context's stream (a null-stream copy would also wait for
// every other blocking stream of the device).
hipError_t upload_now(void *dst, const void *src, size_t bytes, hipStream_t st)
{
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    return e != hipSuccess ? e : hipStreamSynchronize(st);
}

// LDS row slots per wave of the column-plan FP.
int fp_col_slots(const Agent &ag) { return ag.col_win > 0 ? ag.col_win : ag.prj.X + 2; }

AgentDev make_dev(const mtr_ctx *ctx, const Agent &ag)
{
    AgentDev d{};
    d.nu = ag.prj.X;
    d.nv = ag.prj.Y;
    d.A = ag.prj.AngN;
    d.Xs = ctx->vol.Xstart;
    d.Xe = ctx->vol.Xend;
    d.Ys = ctx->vol.Ystart;
    d.Ye = ctx->vol.Yend;
    d.Zs = ctx->vol.Zstart;
    d.Ze = ctx->vol.Zend;
    d.X = ctx->vol.X;
    d.Y = ctx->vol.Y;
    d.Z = ctx->vol.Z;
    d.coef = ag.d_coef;
    d.plan = ag.d_plan;
    d.bsv = ag.d_bsv;
    d.pk = ag.rowfast ? 1 : ag.prj.X;
    d.pi = ag.rowfast ? ag.prj.Y : 1;
    d.p = ag.d_p;
    d.c = ag.d_c;
    d.r = ag.d_r;
    d.x = ag.d_x;
    d.partial = ag.d_partial;
    d.sx = ag.st[0];
    d.sy = ag.st[1];
    d.sz = ag.st[2];
    d.col_xy = ag.d_col_xy;
    d.col_w = ag.d_col_w;
    d.strip = ag.d_strip;
    d.nstrip = ag.col.on ? ag.prj.Y + 1 : 0;
    d.col_p1y = ag.col.p1y;
    d.col_pitch = ag.col.pitch;
    d.col_win = ag.col_win;
    d.col_jit = ag.col_jit;
    d.col_span = ag.col_span;
    d.pm = ag.d_pm;
    d.pn = ag.d_pn;
    d.bpc = ag.d_bpc;
    d.col_grp = ag.d_col_grp;
    d.cs_win = ag.d_cs_win;
    d.sep_win = ag.d_sep_win;
    d.ngrp = ag.ngrp;
    d.seg_part = ag.d_seg;
    d.nseg = ag.fp_seg;
    d.seg_len = ag.fp_seg_len;
    d.seg_order = ag.d_seg_ord;
    return d;
}

int fp_kind_key(const Agent &a);
int bp_kind_key(const Agent &a);

int upload_table(mtr_ctx *ctx)
{
    std::vector<AgentDev> t;
    for (auto &ag : ctx->agents) t.push_back(make_dev(ctx, ag));
    if (t.empty()) return MTR_OK;
    const int K = (int)t.size();
    auto order = [&](int (*key)(const Agent &)) {
        std::vector<int> o(K);
        for (int a = 0; a < K; a++) o[a] = a;
        std::stable_sort(o.begin(), o.end(),
                         [&](int p, int q) { return key(ctx->agents[p]) < key(ctx->agents[q]); });
        return o;
    };
    ctx->fp_order = order(fp_kind_key);
    ctx->bp_order = order(bp_kind_key);
    std::vector<AgentDev> tf(K), tb(K);
    for (int i = 0; i < K; i++) {
        tf[i] = t[ctx->fp_order[i]];
        tb[i] = t[ctx->bp_order[i]];
    }
    HIP_TRY(hipMemcpyAsync(ctx->d_ags, t.data(), sizeof(AgentDev) * K, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->d_ags_fp, tf.data(), sizeof(AgentDev) * K, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->d_ags_bp, tb.data(), sizeof(AgentDev) * K, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MTR_OK;
}

int check_agent(mtr_ctx *ctx, int agent)
{
    if (!ctx) return fail(MTR_ERR_ARG, "null context");
    if (agent < 0 || agent >= (int)ctx->agents.size())
        return fail(MTR_ERR_RANGE, "agent %d out of range (have %zu)", agent, ctx->agents.size());
    return MTR_OK;
}

// Geometry-only forward-projection normaliser of one agent.
int ensure_normaliser(mtr_ctx *ctx, int a)
{
    Agent &ag = ctx->agents[a];
    if (ag.c_ready) return MTR_OK;
    {
        Timed t(ctx, MTR_K_NORMALISER);
        if (ag.col.on) {
            const AgentDev d = make_dev(ctx, ag);
            if (ag.col_pfx)
                HIP_TRY(launch_fp_pfx(ctx->precision, ctx->d_ags + a, 1, fp_pfx_blocks(d),
                                      fp_pfx_lds(ag.col_win, ag.prj.X + 2, ctx->vol.Z, true), true, ag.col_win > 0,
                                      ctx->stream));
            else
                HIP_TRY(launch_fp_col(ctx->precision, ctx->d_ags + a, 1, fp_col_blocks(d), fp_col_slots(ag), true,
                                      ctx->stream));
            HIP_TRY(launch_fp_col_res(ctx->precision, ctx->d_ags + a, 1, fp_col_res_blocks(d), 2, ag.d_c,
                                      ctx->stream));
            // and the back-projection normaliser of every voxel
            if (ag.col_cs)
                HIP_TRY(launch_bp_cs(ctx->precision, ctx->d_ags + a, 1, bp_cs_blocks(d, 16), 2, 0.0f, ctx->stream));
            else
            HIP_TRY(launch_bp_col(ctx->precision, ctx->d_ags + a, 1, bp_col_blocks(d, ag.col.rw), 2, 0.0f, ag.col.rw,
                                  ctx->stream));
        } else {
            HIP_TRY(launch_fp(ctx->precision, ctx->d_ags + a, 1, ag.nfp, 2, ag.lanes_k, nullptr, ag.d_c,
                              ctx->stream));
        }
    }
    ag.c_ready = true;
    return MTR_OK;
}

// A batch of resident agents, possibly of several contexts (volumes):
// entry i is described by d_tab[i] on the device.  Launches go to the
// engine context's stream; consecutive entries sharing a kernel share one
// launch (blockIdx.y selects the entry, each entry sizes its own grid).
struct BatchEntry {
    mtr_ctx *owner;
    Agent *ag;
};
struct Batch {
    std::vector<BatchEntry> e;
    AgentDev *d_tab = nullptr;
};

// Column-plan kinds sort after every other kind (a suffix of a sorted
// batch: the SIRT loop runs them on a second stream).
int fp_kind_key(const Agent &a)
{
    return a.col.on ? 10100 + a.col.rw + (a.col_pfx ? 1000 : 0) + (a.col_win > 0 ? 2000 : 0)
                    : (a.fp_nc ? 10 + 8 * a.fp_nc + a.fp_g + (a.fp_seg > 1 ? 500 : 0) : (a.lanes_k ? 1 : 0));
}
int bp_kind_key(const Agent &a)
{
    return a.col.on ? 10100 + a.col.rw + (a.col_cs ? 1000 : 0)
                    : (a.bpm.sep ? 10 + 4 * a.bpm.W + (a.bpm.rw == 32 ? 1 : (a.bpm.rw == 16 ? 2 : 3)) : 0);
}

// FP + residual of entries [first, first + count), on stream st (the
// engine's stream when null).
int fp_batch(mtr_ctx *eng, const Batch &B, int first, int count, hipStream_t st = nullptr)
{
    if (!st) st = eng->stream;
    int s = first;
    const int end = first + count;
    while (s < end) {
        const Agent &a0 = *B.e[s].ag;
        int e = s + 1;
        while (e < end && fp_kind_key(*B.e[e].ag) == fp_kind_key(a0)) e++;
        int mb = 0, mr = 0, nu = 0, lds = 0;
        bool pfx = true;
        for (int q = s; q < e; q++) pfx = pfx && B.e[q].ag->col_pfx;
        const bool win = a0.col_win > 0;  // the kind key keeps windowed and whole-row agents apart
        for (int q = s; q < e; q++) {
            Agent &ag = *B.e[q].ag;
            const AgentDev d = make_dev(B.e[q].owner, ag);
            if (a0.col.on) {
                ag.npartial = fp_col_res_blocks(d);
                mb = std::max(mb, pfx ? fp_pfx_blocks(d) : fp_col_blocks(d));
                mr = std::max(mr, ag.npartial);
                nu = std::max(nu, fp_col_slots(ag));
                if (pfx) lds = std::max(lds, fp_pfx_lds(ag.col_win, ag.prj.X + 2, d.Z, false));
            } else {
                if (ag.fp_seg > 1) {
                    ag.npartial = fp_col_res_blocks(d);
                    mr = std::max(mr, ag.npartial);
                    mb = std::max(mb, fp_stream_blocks(d, ag.fp_g));
                } else {
                    ag.npartial = ag.fp_nc ? fp_stream_blocks(d, ag.fp_g) : ag.nfp;
                    mb = std::max(mb, ag.npartial);
                }
            }
        }
        Timed t(eng, MTR_K_FP, true, st);
        if (a0.col.on) {
            {
                Timed tk(eng, MTR_K_FP_COLUMN, true, st);
                if (pfx)
                    HIP_TRY(launch_fp_pfx(eng->precision, B.d_tab + s, e - s, mb, lds, false, win, st));
                else HIP_TRY(launch_fp_col(eng->precision, B.d_tab + s, e - s, mb, nu, false, st));
            }
            Timed tk(eng, MTR_K_FP_COLUMN_RES, true, st);
            HIP_TRY(launch_fp_col_res(eng->precision, B.d_tab + s, e - s, mr, 0, nullptr, st));
        } else if (a0.fp_nc) {
            {
                Timed tk(eng, MTR_K_FP_STREAM, true, st);
                HIP_TRY(launch_fp_stream(eng->precision, B.d_tab + s, e - s, mb, a0.fp_nc, a0.fp_g, st));
            }
            // the kind key keeps segmented and whole-ray agents apart
            if (a0.fp_seg > 1) {
                Timed tk(eng, MTR_K_FP_STREAM_RES, true, st);
                HIP_TRY(launch_fp_seg_res(eng->precision, B.d_tab + s, e - s, mr, st));
            }
        } else {
            Timed tk(eng, MTR_K_FP_GENERAL, true, st);
            HIP_TRY(launch_fp(eng->precision, B.d_tab + s, e - s, mb, 0, a0.lanes_k, nullptr, nullptr, st));
        }
        s = e;
    }
    return MTR_OK;
}

// BP (bpt: initial model) of entries [first, first + count), on stream st
// (the engine's stream when null).
int bp_batch(mtr_ctx *eng, const Batch &B, int first, int count, bool bpt, float lambda, hipStream_t st = nullptr)
{
    if (!st) st = eng->stream;
    int s = first;
    const int end = first + count;
    while (s < end) {
        const Agent &a0 = *B.e[s].ag;
        int e = s + 1;
        while (e < end && bp_kind_key(*B.e[e].ag) == bp_kind_key(a0)) e++;
        const BpMode &m = a0.bpm;
        const ColMode &cm = a0.col;
        int mb = 0;
        for (int q = s; q < e; q++) {
            const AgentDev d = make_dev(B.e[q].owner, *B.e[q].ag);
            mb = std::max(mb, cm.on ? (a0.col_cs ? bp_cs_blocks(d, 16) : bp_col_blocks(d, cm.rw))
                                    : (m.sep ? bp_sep_blocks(d, m.W, m.P1, m.P2, m.rw) : bp_blocks(d, eng->bp_tz)));
        }
        Timed t(eng, bpt ? MTR_K_BPT : MTR_K_BP, true, st);
        Timed tk(eng, cm.on ? MTR_K_BP_COLUMN : (m.sep ? MTR_K_BP_SEPARABLE : MTR_K_BP_GENERAL), !bpt, st);
        if (cm.on) {
            if (a0.col_cs)
                HIP_TRY(launch_bp_cs(eng->precision, B.d_tab + s, e - s, mb, bpt ? 1 : 0, lambda, st));
            else
                HIP_TRY(launch_bp_col(eng->precision, B.d_tab + s, e - s, mb, bpt ? 1 : 0, lambda, cm.rw, st));
        } else if (m.sep) {
            HIP_TRY(launch_bp_sep(eng->precision, B.d_tab + s, e - s, mb, bpt, lambda, m.W, m.P1, m.P2, m.rw,
                                  st));
        } else {
            HIP_TRY(launch_bp(eng->precision, B.d_tab + s, e - s, mb, eng->bp_tz, bpt, lambda, nullptr, eng->lds_cap,
                              st));
        }
        s = e;
    }
    return MTR_OK;
}

// The agents [first, first + count) of one context as a batch.
Batch ctx_batch(mtr_ctx *ctx, int first, int count)
{
    Batch B;
    for (int a = first; a < first + count; a++) B.e.push_back({ctx, &ctx->agents[a]});
    B.d_tab = ctx->d_ags + first;
    return B;
}

// Every agent of the context in one of the kind-sorted tables.
Batch ctx_sorted(mtr_ctx *ctx, const std::vector<int> &order, AgentDev *d_tab)
{
    Batch B;
    for (int a : order) B.e.push_back({ctx, &ctx->agents[a]});
    B.d_tab = d_tab;
    return B;
}

int bp_group(mtr_ctx *ctx, int first, int count, bool bpt, float lambda)
{
    if (first == 0 && count == (int)ctx->agents.size() && (int)ctx->bp_order.size() == count)
        return bp_batch(ctx, ctx_sorted(ctx, ctx->bp_order, ctx->d_ags_bp), 0, count, bpt, lambda);
    return bp_batch(ctx, ctx_batch(ctx, first, count), 0, count, bpt, lambda);
}

// ---- autotuner ------------------------------------------------------
// Back-projection tiles (run length along the plane axis or Z, and for the
// column plan the lane axis) are chosen per agent by timing each candidate
// on the agent's own data: the initial-model BPT, whose tiles and staging
// are those of the SIRT back-projection, with HIP events.  Choices are
// cached per shape and geometry for the process, and in the file named by
// MTR_TUNE_CACHE when set ("key<TAB>run lanes" lines).
struct TuneCache {
    std::mutex mu;
    bool loaded = false;
    std::map<std::string, std::pair<int, int>> kv;
};

TuneCache &tune_cache()
{
    static TuneCache c;
    return c;
}

uint64_t fnv(const void *p, size_t n, uint64_t h = 1469598103934665603ull)
{
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

std::string tune_key(const mtr_ctx *ctx, const Agent &ag)
{
    char buf[256];
    uint64_t h = fnv(ag.xc.data(), sizeof(double) * ag.xc.size());
    h = fnv(ag.yc.data(), sizeof(double) * ag.yc.size(), h);
    std::snprintf(buf, sizeof buf, "%s W%d box%dx%dx%d det%dx%dx%d prec%d geom%016llx", ag.col.on ? "col" : "sep",
                  ag.bpm.W, ctx->vol.X, ctx->vol.Y, ctx->vol.Z, ag.prj.X, ag.prj.Y, ag.prj.AngN, ctx->precision,
                  (unsigned long long)h);
    return buf;
}

void tune_cache_load(TuneCache &c)
{
    if (c.loaded) return;
    c.loaded = true;
    const char *path = std::getenv("MTR_TUNE_CACHE");
    if (!path) return;
    FILE *f = std::fopen(path, "r");
    if (!f) return;
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        char *tab = std::strchr(line, '\t');
        if (!tab) continue;
        *tab = 0;
        int run = 0, lanes = 0;
        if (std::sscanf(tab + 1, "%d %d", &run, &lanes) == 2) c.kv[line] = {run, lanes};
    }
    std::fclose(f);
}

void tune_cache_store(TuneCache &c, const std::string &key, int run, int lanes)
{
    c.kv[key] = {run, lanes};
    if (const char *path = std::getenv("MTR_TUNE_CACHE")) {
        if (FILE *f = std::fopen(path, "a")) {
            std::fprintf(f, "%s\t%d %d\n", key.c_str(), run, lanes);
            std::fclose(f);
        }
    }
}

// Applies a tile choice to agent a (run length, lane axis for the column
// plan) and refreshes its device tables.
int apply_tiles(mtr_ctx *ctx, int a, int run, int lanes)
{
    Agent &ag = ctx->agents[a];
    if (ag.col.on) {
        const ColMode m = col_mode(ag.xc.data(), ag.yc.data(), ag.prj.AngN, ag.prj, run, lanes);
        if (!m.on) return fail(MTR_ERR_SHAPE, "tile choice does not fit");
        ag.col = m;
        const std::vector<BpSepView> bsv = bp_col_views(m, ag.xc.data(), ag.prj.AngN);
        HIP_TRY(hipMemcpyAsync(ag.d_bsv, bsv.data(), sizeof(BpSepView) * bsv.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    } else {
        ag.bpm.rw = run;
    }
    if (ctx->step_graph) {
        (void)hipGraphExecDestroy(ctx->step_graph);
        ctx->step_graph = nullptr;
    }
    return upload_table(ctx);
}

// Chooses the back-projection tiles of agent a by timing SIRT
// back-projections of a zeroed residual (x += 0).  Clobbers the agent's
// residual: call only before a BPT, which needs none.
int tune_agent(mtr_ctx *ctx, int a)
{
    Agent &ag = ctx->agents[a];
    if (ag.tuned || !ctx->autotune || !(ag.col.on || ag.bpm.sep) || ag.col_cs) {
        ag.tuned = true;
        return MTR_OK;
    }
    std::vector<std::pair<int, int>> cand;
    if (ag.col.on) {
        for (int run : {16})
            for (int lanes : {0, 1})
                if (col_mode(ag.xc.data(), ag.yc.data(), ag.prj.AngN, ag.prj, run, lanes).on) cand.push_back({run, lanes});
    } else {
        for (int run : {32, 16, 8}) cand.push_back({run, 0});
    }
    TuneCache &tc = tune_cache();
    const std::string key = tune_key(ctx, ag);
    {
        std::lock_guard<std::mutex> lk(tc.mu);
        tune_cache_load(tc);
        auto it = tc.kv.find(key);
        if (it != tc.kv.end()) {
            ag.tuned = true;
            return apply_tiles(ctx, a, it->second.first, it->second.second);
        }
    }
    if (!ctx->tune_timed) {  // the heuristic tiles bp_mode chose
        ag.tuned = true;
        return MTR_OK;
    }
    const bool timing = ctx->timing;
    ctx->timing = false;
    HIP_TRY(launch_fill(ctx->precision, ag.P(), ag.d_r, 0.0f, ctx->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    std::pair<int, int> pick = cand.front();
    int rc = MTR_OK;
    for (const auto &c : cand) {
        if ((rc = apply_tiles(ctx, a, c.first, c.second)) != MTR_OK) break;
        if ((rc = bp_group(ctx, a, 1, false, 0.0f)) != MTR_OK) break;  // warm-up
        float t = 1e30f;
        for (int rep = 0; rep < 3 && rc == MTR_OK; rep++) {
            float tr = 0.f;
            if (hipEventRecord(e0, ctx->stream) != hipSuccess || bp_group(ctx, a, 1, false, 0.0f) != MTR_OK ||
                hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&tr, e0, e1) != hipSuccess)
                rc = fail(MTR_ERR_HIP, "autotuner timing failed");
            t = std::min(t, tr);
        }
        if (rc != MTR_OK) break;
        if (t < best) {
            best = t;
            pick = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->timing = timing;
    if (rc != MTR_OK) return rc;
    ag.tuned = true;
    {
        std::lock_guard<std::mutex> lk(tc.mu);
        tune_cache_store(tc, key, pick.first, pick.second);
    }
    return apply_tiles(ctx, a, pick.first, pick.second);
}

int tune_all(mtr_ctx *ctx)
{
    for (int a = 0; a < (int)ctx->agents.size(); a++) MTR_TRY(tune_agent(ctx, a));
    return MTR_OK;
}

int ensure_plain_scratch(mtr_ctx *ctx, int64_t P)
{
    if (ctx->sc_cap >= P) return MTR_OK;
    dev_free(ctx, ctx->d_s);
    dev_free(ctx, ctx->d_cc);
    ctx->d_s = ctx->d_cc = nullptr;
    ctx->sc_cap = 0;
    HIP_TRY(dev_alloc(ctx, (void **)&ctx->d_s, sizeof(float) * P));
    HIP_TRY(dev_alloc(ctx, (void **)&ctx->d_cc, sizeof(float) * P));
    ctx->sc_cap = P;
    return MTR_OK;
}

// Projections in the [A][nv][nu] layout the plain form keeps (FP32).
int plain_p(mtr_ctx *ctx, Agent &ag, const float **out)
{
    if (!ag.rowfast) {
        *out = (const float *)ag.d_p;
        return MTR_OK;
    }
    if (!ag.d_p_std) {
        HIP_TRY(dev_alloc(ctx, (void **)&ag.d_p_std, sizeof(float) * ag.P()));
        ag.dev_bytes += sizeof(float) * ag.P();
        ag.dev_peak = std::max(ag.dev_peak, ag.dev_bytes);
        const int64_t ss[3] = {ag.prj.Y, 1, (int64_t)ag.prj.X * ag.prj.Y};
        const int64_t ds[3] = {1, ag.prj.X, (int64_t)ag.prj.X * ag.prj.Y};
        HIP_TRY(launch_permute(MTR_PREC_FP32, ag.d_p, ag.d_p_std, ag.prj.X, ag.prj.Y, ag.prj.AngN, ss, ds,
                               ctx->stream));
    }
    *out = ag.d_p_std;
    return MTR_OK;
}

// Plain form: one agent, splat FP with atomics into zeroed s/c (c
// recomputed every iteration), separate residual kernel, direct-gather BP,
// host synchronisation after each iteration.
int sirt_plain(mtr_ctx *ctx, int a, int iters, float lambda)
{
    Agent &ag = ctx->agents[a];
    if (ctx->precision != MTR_PREC_FP32)
        return fail(MTR_ERR_STATE, "plain form runs FP32 storage only");
    MTR_TRY(ensure_plain_scratch(ctx, ag.P()));
    const float *pstd = nullptr;
    MTR_TRY(plain_p(ctx, ag, &pstd));
    const AgentDev d = make_dev(ctx, ag);
    for (int it = 0; it < iters; it++) {
        HIP_TRY(hipMemsetAsync(ctx->d_s, 0, sizeof(float) * ag.P(), ctx->stream));
        HIP_TRY(hipMemsetAsync(ctx->d_cc, 0, sizeof(float) * ag.P(), ctx->stream));
        {
            Timed t(ctx, MTR_K_SPLAT);
            HIP_TRY(launch_plain_splat(d, ctx->d_s, ctx->d_cc, ctx->stream));
        }
        {
            Timed t(ctx, MTR_K_ELEMENTWISE);
            HIP_TRY(launch_residual(ctx->precision, ag.P(), pstd, ctx->d_s, ctx->d_cc, ag.d_r, ctx->stream));
        }
        {
            Timed t(ctx, MTR_K_BP);
            HIP_TRY(launch_plain_bp(d, (const float *)ag.d_r, lambda, 0, ctx->stream));
        }
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return MTR_OK;
}

int sirt_fast(mtr_ctx *ctx, const std::vector<int> &ids, int iters, float lambda)
{
    for (int a : ids) MTR_TRY(ensure_normaliser(ctx, a));
    const int K = (int)ctx->agents.size();
    bool all = (int)ids.size() == K && (int)ctx->fp_order.size() == K;
    for (int i = 0; all && i < K; i++) all = ids[i] == i;
    if (all) {
        // the whole context: each kernel kind once per iteration
        const Batch F = ctx_sorted(ctx, ctx->fp_order, ctx->d_ags_fp);
        const Batch Bk = ctx_sorted(ctx, ctx->bp_order, ctx->d_ags_bp);
        // The column-plan agents (a suffix of both kind orders) iterate on
        // the context's stream and the others on the second one: their
        // kernels are bound by different units (VALU and LDS against the
        // vector-memory path), and blocks of both kinds sharing the CUs
        // overlap.  Every agent's iterations touch only its own buffers.
        // Per-kernel timing keeps one stream, so each kernel is timed alone.
        int fs = 0, bs = 0;
        while (fs < K && !F.e[fs].ag->col.on) fs++;
        while (bs < K && !Bk.e[bs].ag->col.on) bs++;
        if (ctx->aux && ctx->kind_streams && !ctx->timing && fs > 0 && fs < K && fs == bs) {
            HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->stream));
            HIP_TRY(hipStreamWaitEvent(ctx->aux, ctx->ev_fork, 0));
            for (int it = 0; it < iters; it++) {
                MTR_TRY(fp_batch(ctx, F, fs, K - fs));
                MTR_TRY(bp_batch(ctx, Bk, bs, K - bs, false, lambda));
                MTR_TRY(fp_batch(ctx, F, 0, fs, ctx->aux));
                MTR_TRY(bp_batch(ctx, Bk, 0, bs, false, lambda, ctx->aux));
            }
            HIP_TRY(hipEventRecord(ctx->ev_join, ctx->aux));
            HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
            return MTR_OK;
        }
        for (int it = 0; it < iters; it++) {
            MTR_TRY(fp_batch(ctx, F, 0, K));
            MTR_TRY(bp_batch(ctx, Bk, 0, K, false, lambda));
        }
        return MTR_OK;
    }
    // one batch per contiguous run of agents
    std::vector<Batch> runs;
    size_t s = 0;
    while (s < ids.size()) {
        size_t e = s + 1;
        while (e < ids.size() && ids[e] == ids[e - 1] + 1) e++;
        runs.push_back(ctx_batch(ctx, ids[s], (int)(e - s)));
        s = e;
    }
    for (int it = 0; it < iters; it++) {
        for (const Batch &B : runs) MTR_TRY(fp_batch(ctx, B, 0, (int)B.e.size()));
        for (const Batch &B : runs) MTR_TRY(bp_batch(ctx, B, 0, (int)B.e.size(), false, lambda));
    }
    return MTR_OK;
}

// Row window of the column-plan FP (ColRows in the kernels).  For every
// view and strip, the columns in walk order span slots sa .. sb (slot =
// detector column + 1, from the kernel's u at the planes of smallest and
// largest u).  jit is the largest drop of a column's sa below the largest
// sa before it, span the largest sb - sa (slots past the detector do not
// count); the window needs more than span + jit slots.  The reverse walk is
// chosen when it needs a smaller window (u falling along the memory order).
// win = 0 when whole rows (L slots) already leave room for six blocks per
// CU in its 160 KiB of LDS (the window's bookkeeping then costs more than the
// extra waves gain: 4.29 -> 4.43 ms at L = 770), or when no window smaller
// than L suffices; rev: walk each strip backwards.
void col_window(const std::vector<int2> &xy, const std::vector<int> &cnt, const double *xc, int A,
                const mtr_volume_size &vol, int nu, bool pfx, int *win, int *jit, int *span, bool *rev)
{
    const int L = nu + 2, ns = (int)cnt.size() - 1;
    *win = 0;
    *rev = false;
    // the prefix-moment kernel keeps eight waves' rows beside the staged
    // moments, and always takes a window smaller than whole rows
    if (!pfx && (160 << 10) / (fp_col_waves() * L * (int)sizeof(float2)) >= 6) return;
    const int zs = vol.Zstart, nz = vol.Z;
    // views [v0, v1): (largest drop, largest span)
    auto pass = [&](bool back, int v0, int v1) {
        std::pair<int, int> out{0, 0};
        for (int v = v0; v < v1; v++) {
            const double *c = xc + 4 * v;
            const double zlo = c[3] >= 0.0 ? (double)zs : (double)(zs + nz - 1);
            const double zhi = c[3] >= 0.0 ? (double)(zs + nz - 1) : (double)zs;
            for (int s = 0; s < ns; s++) {
                int runmax = -0x40000000;
                const int n = cnt[s + 1] - cnt[s];
                for (int q = 0; q < n; q++) {
                    const int2 p = xy[back ? cnt[s + 1] - 1 - q : cnt[s] + q];
                    const double bxy = std::fma(c[2], (double)p.y, std::fma(c[1], (double)p.x, c[0]));
                    const int sa = (int)std::floor(std::fma(c[3], zlo, bxy) - 1.0) + 1;
                    const int sb = (int)std::floor(std::fma(c[3], zhi, bxy) - 1.0) + 2;
                    runmax = std::max(runmax, sa);
                    out.first = std::max(out.first, runmax - sa);
                    out.second = std::max(out.second, std::min(sb, L - 1) - sa);
                }
            }
        }
        return out;
    };
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (int r = 0; r < 2; r++) {
        std::vector<std::pair<int, int>> part(nth);
        std::vector<std::thread> th;
        for (int t = 0; t < nth; t++) {
            auto job = [&, t] { part[t] = pass(r == 1, A * t / nth, A * (t + 1) / nth); };
            try {
                th.emplace_back(job);
            } catch (...) {
                job();  // no thread to spare: this share of the views here
            }
        }
        for (auto &x : th) x.join();
        int jt = 0, sp = 0;
        for (auto &q : part) {
            jt = std::max(jt, q.first);
            sp = std::max(sp, q.second);
        }
        jt += 2;  // margins against any rounding difference of the kernel's u
        sp += 2;
        int w = 64;
        while (w <= sp + jt) w *= 2;
        if (w < L && (*win == 0 || w < *win)) {
            *win = w;
            *jit = jt;
            *span = sp;
            *rev = r == 1;
        }
    }
}

// p_checked: p_host was already scanned for non-finite values (a request
// the server validated at submission).
int agent_add_common(mtr_ctx *ctx, const mtr_projection_size *prj, const double *xc, const double *yc,
                     const float *p_host, const float *p_dev, int *agent_id, bool p_checked = false)
{
    if (!ctx || !prj || !xc || !yc || (!p_host && !p_dev)) return fail(MTR_ERR_ARG, "null argument");
    if (!ctx->vol_set) return fail(MTR_ERR_STATE, "mtr_ctx_set_volume must precede mtr_agent_add");
    if (!prj_valid(prj)) return fail(MTR_ERR_SHAPE, "invalid projection size %d x %d x %d", prj->X, prj->Y, prj->AngN);
    if ((int)ctx->agents.size() >= kMaxAgents)
        return fail(MTR_ERR_RANGE, "at most %d agents per context", kMaxAgents);
    if (p_host && !p_checked) {
        // a NaN or Inf pixel would reach every voxel its views touch
        const int64_t i = first_nonfinite(p_host, (int64_t)prj->X * prj->Y * prj->AngN);
        if (i >= 0)
            return fail(MTR_ERR_NONFINITE, "projection value %lld (view %lld) is not finite", (long long)i,
                        (long long)(i / ((int64_t)prj->X * prj->Y)));
    }
    MTR_TRY(set_device(ctx));
    const int A = prj->AngN;
    std::vector<ViewCoef> coef(A);
    std::vector<FpPlan> plan(A);
    for (int v = 0; v < A; v++) {
        for (int m = 0; m < 4; m++) {
            coef[v].c[m] = xc[4 * v + m];
            coef[v].c[4 + m] = yc[4 * v + m];
        }
        int rc = make_plan(xc + 4 * v, yc + 4 * v, plan[v]);
        if (rc != MTR_OK) return fail(rc, "view %d: degenerate or non-finite geometry", v);
    }
    pfx_view_order(xc, A, ctx->vol.Z, plan);
    ctx->agents.emplace_back();
    Agent &ag = ctx->agents.back();
    auto undo = [&](int rc) {
        Agent &g = ctx->agents.back();
        for (void *q : {(void *)g.d_coef, (void *)g.d_plan, (void *)g.d_bsv, g.p_owned ? (void *)g.d_p : nullptr, (void *)g.d_c,
                        g.d_r, g.d_x, g.d_w, g.d_tmp, (void *)g.d_partial, (void *)g.d_scalar, (void *)g.d_p_std,
                        (void *)g.d_col_xy, (void *)g.d_col_w, (void *)g.d_strip, (void *)g.d_pm, (void *)g.d_pn,
                        (void *)g.d_bpc, (void *)g.d_col_grp, (void *)g.d_cs_win, (void *)g.d_sep_win, (void *)g.d_seg, (void *)g.d_seg_ord})
            dev_free(ctx, q);
        ctx->agents.pop_back();
        return rc;
    };
    ag.prj = *prj;
    ag.xc.assign(xc, xc + 4 * A);
    ag.yc.assign(yc, yc + 4 * A);
    // Lanes along detector rows when every row is one X plane: the 64
    // lanes of a wave then walk the same (Y, Z) path in adjacent planes.
    bool all_wx = true;
    for (int v = 0; v < A; v++) all_wx = all_wx && plan[v].kind == 1 && plan[v].W == 0;
    ag.lanes_k = all_wx;
    ag.bpm = bp_mode(plan, xc, yc, ctx->vol, *prj);
    int sW = 0;
    ag.fp_nc = fp_stream_nc(plan, xc, &sW);
    // the streaming FP addresses the volume with 32-bit byte offsets
    if ((uint64_t)ctx->vol.X * ctx->vol.Y * ctx->vol.Z * elem_size(ctx->precision) > 0xFFFFFFFFull) ag.fp_nc = 0;
    ag.fp_g = 2;
    if (const char *g = std::getenv("MTR_FP_COLS")) ag.fp_g = std::atoi(g) == 1 ? 1 : 2;
    ag.fast = ag.fp_nc ? sW : 0;
    // Streaming FP of a slab (plane axis X or Y) in Z segments of 128
    // planes when the box has at least two: the voxels that the rays in
    // flight on an XCD share then mostly stay in its L2 (64 planes halve
    // that working set but double the partial sums; measured slower at
    // 1024 x 1024 x 256).  MTR_FP_STREAM_SEG sets the planes per segment
    // (0: whole rays).
    if (ag.fp_nc && sW != 2) {
        const int len = fp_seg_planes(), nz = ctx->vol.Zend - ctx->vol.Zstart;
        if (len > 0 && nz >= 2 * len) {
            ag.fp_seg = (nz + len - 1) / len;
            ag.fp_seg_len = len;
        }
    }
    // Block order of the streaming FP: in each segment, the (view, column
    // group) pairs sorted by the coordinate H (the in-plane axis other than
    // Z) where the group's middle ray crosses the segment's middle plane,
    // so the blocks in flight on an XCD read one window of H for all views.
    // MTR_FP_STREAM_ORDER=0 keeps the view-fastest order.
    std::vector<int> seg_ord;
    if (ag.fp_nc && sW != 2) {
        const char *o = std::getenv("MTR_FP_STREAM_ORDER");
        if (!o || std::atoi(o) != 0) {
            const int G4 = 4 * ag.fp_g, nti = (prj->X + G4 - 1) / G4, nva = nti * A, H = sW == 0 ? 2 : 1;
            const double hmid = 0.5 * (sW == 0 ? ctx->vol.Ystart + ctx->vol.Yend - 1 : ctx->vol.Xstart + ctx->vol.Xend - 1);
            const int len = ag.fp_seg > 1 ? ag.fp_seg_len : ctx->vol.Zend - ctx->vol.Zstart;
            std::vector<std::pair<double, int>> key(nva);
            for (int sg = 0; sg < ag.fp_seg; sg++) {
                const double zc = ctx->vol.Zstart + sg * len + 0.5 * (std::min(len, ctx->vol.Zend - ctx->vol.Zstart - sg * len) - 1);
                for (int bi = 0; bi < nti; bi++)
                    for (int v = 0; v < A; v++) {
                        const double *c = xc + 4 * v;
                        // u = c0 + cH H + cZ Z - 1 at the group's middle column
                        const double u = bi * G4 + 0.5 * (G4 - 1);
                        const double h = std::fabs(c[H]) > 1e-6 ? (u + 1.0 - c[0] - c[3] * zc) / c[H] : hmid;
                        key[bi * A + v] = {h, v + A * bi};
                    }
                std::stable_sort(key.begin(), key.end(),
                                 [](const std::pair<double, int> &p, const std::pair<double, int> &q) { return p.first < q.first; });
                for (int q = 0; q < nva; q++) seg_ord.push_back(key[q].second);
                for (int q = 0; q < nva; q++) key[q] = {0.0, 0};
            }
        }
    }
    if (!ag.fp_nc && !ag.bpm.sep) {
        ag.col = col_mode(xc, yc, A, *prj, 16);
        if (ag.col.on) ag.fast = 2;  // Z-contiguous columns
    }
    layout_strides(ag.fast, ctx->vol.X, ctx->vol.Y, ctx->vol.Z, ag.st);
    ag.rowfast = ag.fp_nc > 0 && ag.bpm.sep;
    const size_t es = elem_size(ctx->precision);
    const int64_t P = ag.P();
#define ALLOC(ptr, bytes)                                                                 \
    do {                                                                                  \
        hipError_t e_ = dev_alloc(ctx, (void **)&(ptr), (bytes));                         \
        if (e_ != hipSuccess)                                                             \
            return undo(fail(MTR_ERR_OOM, "hipMalloc %zu bytes: %s", (size_t)(bytes),     \
                             hipGetErrorString(e_)));                                     \
        ag.dev_bytes += (size_t)(bytes);                                                  \
        ag.dev_peak = std::max(ag.dev_peak, ag.dev_bytes);                                \
    } while (0)
    ALLOC(ag.d_coef, sizeof(ViewCoef) * A);
    ALLOC(ag.d_plan, sizeof(FpPlan) * A);
    const std::vector<BpSepView> bsv = ag.col.on ? bp_col_views(ag.col, xc, A) : bp_sep_views(ag.bpm, xc, yc, A);
    if (!bsv.empty()) {
        ALLOC(ag.d_bsv, sizeof(BpSepView) * A);
        if (hipMemcpyAsync(ag.d_bsv, bsv.data(), sizeof(BpSepView) * A, hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess)
            return undo(fail(MTR_ERR_HIP, "geometry upload failed"));
    }
    if (!ag.col.on && !bsv.empty()) {
        const std::vector<int2> sw = bp_sep_windows(bsv.data(), A, make_dev(ctx, ag), ag.bpm.P1, ag.bpm.P2);
        ALLOC(ag.d_sep_win, sizeof(int2) * sw.size());
        if (upload_now(ag.d_sep_win, sw.data(), sizeof(int2) * sw.size(), ctx->stream) != hipSuccess)
            return undo(fail(MTR_ERR_HIP, "back-projection window upload failed"));
    }
    ALLOC(ag.d_c, sizeof(float) * P);
    ALLOC(ag.d_r, es * P);
    ALLOC(ag.d_x, es * ctx->N);
    ALLOC(ag.d_w, es * ctx->N);
    if (ag.fast != 0) ALLOC(ag.d_tmp, es * ctx->N);
    if (ag.col.on) {
        // strips: columns grouped by k0 = floor(v) in [-1, nv - 1], each
        // strip in memory order; v is the reference's chain with yc3 = 0
        const mtr_volume_size &vb = ctx->vol;
        const int nstrip = prj->Y + 1;
        std::vector<int> cnt(nstrip + 1, 0);
        std::vector<int> key((size_t)vb.X * vb.Y);
        std::vector<double> dvs((size_t)vb.X * vb.Y);
        for (int Y = vb.Ystart; Y < vb.Yend; Y++)
            for (int X = vb.Xstart; X < vb.Xend; X++) {
                const double v = std::fma(yc[2], (double)Y, std::fma(yc[1], (double)X, yc[0])) - 1.0;
                const double k0 = std::floor(v);
                const size_t c = (size_t)(Y - vb.Ystart) * vb.X + (X - vb.Xstart);
                key[c] = (k0 < -1.0 || k0 > prj->Y - 1) ? -1 : (int)k0 + 1;
                dvs[c] = v - k0;
                if (key[c] >= 0) cnt[key[c] + 1]++;
            }
        for (int q = 0; q < nstrip; q++) cnt[q + 1] += cnt[q];
        std::vector<int2> xy(std::max(cnt[nstrip], 1));
        std::vector<float2> wts(std::max(cnt[nstrip], 1));
        std::vector<int> pos(cnt.begin(), cnt.end() - 1);
        for (int Y = vb.Ystart; Y < vb.Yend; Y++)
            for (int X = vb.Xstart; X < vb.Xend; X++) {
                const size_t c = (size_t)(Y - vb.Ystart) * vb.X + (X - vb.Xstart);
                if (key[c] < 0) continue;
                const int at = pos[key[c]]++;
                xy[at] = make_int2(X, Y);
                wts[at] = make_float2((float)(1.0 - dvs[c]), (float)dvs[c]);
            }
        bool rev = false;
        const char *pf = std::getenv("MTR_FP_COL_PFX");
        ag.col_pfx = !(pf && std::atoi(pf) == 0);
        col_window(xy, cnt, xc, A, vb, prj->X, ag.col_pfx, &ag.col_win, &ag.col_jit, &ag.col_span, &rev);
        if (const char *w = std::getenv("MTR_FP_COL_WINDOW"); w && std::atoi(w) == 0) {
            ag.col_win = 0;
            rev = false;
        }
        // the moments of one column and the rows must fit one block's LDS
        if (ag.col_pfx && fp_pfx_lds(ag.col_win, prj->X + 2, vb.Z, false) == 0) ag.col_pfx = false;
        if (rev)
            for (int q = 0; q < nstrip; q++) {
                std::reverse(xy.begin() + cnt[q], xy.begin() + cnt[q + 1]);
                std::reverse(wts.begin() + cnt[q], wts.begin() + cnt[q + 1]);
            }
        // columns whose rows lie off the detector follow the strips: the
        // forward projection never reads them, the strip-tiled
        // back-projection gives them C = 0 (x = 0 at the BPT)
        xy.resize(cnt[nstrip]);
        wts.resize(cnt[nstrip]);
        for (int Y = vb.Ystart; Y < vb.Yend; Y++)
            for (int X = vb.Xstart; X < vb.Xend; X++)
                if (key[(size_t)(Y - vb.Ystart) * vb.X + (X - vb.Xstart)] < 0) {
                    xy.push_back(make_int2(X, Y));
                    wts.push_back(make_float2(0.0f, 0.0f));
                }
        const int ncols = (int)xy.size();  // strip columns, then the off-detector ones
        if (xy.empty()) {
            xy.push_back(make_int2(0, 0));
            wts.push_back(make_float2(0.0f, 0.0f));
        }
        ALLOC(ag.d_col_xy, sizeof(int2) * xy.size());
        ALLOC(ag.d_col_w, sizeof(float2) * wts.size());
        ALLOC(ag.d_strip, sizeof(int) * cnt.size());
        ALLOC(ag.d_pm, sizeof(float) * P);
        ALLOC(ag.d_pn, sizeof(float) * P);
        ALLOC(ag.d_bpc, sizeof(float) * ctx->N);
        // strip tiles of the back-projection: 64 consecutive columns of a
        // strip, usable when every tile's footprint in every view is narrow
        // enough to stage (the kernel's own FP64 bounds, over any block of
        // bp_cs_planes() planes)
        {
            std::vector<ColGroup> grp;
            for (int q = 0; q <= nstrip; q++) {
                // q = nstrip: the columns off the detector's rows (k0 = -2)
                const int ja = q < nstrip ? cnt[q] : cnt[nstrip], jb = q < nstrip ? cnt[q + 1] : ncols;
                for (int j = ja; j < jb; j += 64) {
                    ColGroup g{};
                    g.j0 = j;
                    g.n = std::min(64, jb - j);
                    g.xmin = g.ymin = 0x7fffffff;
                    g.xmax = g.ymax = -0x7fffffff;
                    for (int i = j; i < j + g.n; i++) {
                        g.xmin = std::min(g.xmin, xy[i].x);
                        g.xmax = std::max(g.xmax, xy[i].x);
                        g.ymin = std::min(g.ymin, xy[i].y);
                        g.ymax = std::max(g.ymax, xy[i].y);
                    }
                    g.k0 = q < nstrip ? q - 1 : -2;
                    grp.push_back(g);
                }
            }
            const char *cs = std::getenv("MTR_BP_COL_STRIP");
            ag.col_cs = !grp.empty() && !(cs && std::atoi(cs) == 0);
            const double zb1 = (double)(bp_cs_planes() - 1);
            for (size_t gi = 0; gi < grp.size() && ag.col_cs; gi++) {
                const ColGroup &g = grp[gi];
                if (g.k0 < -1) continue;  // no footprint
                for (int v = 0; v < A && ag.col_cs; v++) {
                    const double *c = xc + 4 * v;
                    const double ux = std::fmin(c[1] * g.xmin, c[1] * g.xmax), vx = std::fmax(c[1] * g.xmin, c[1] * g.xmax);
                    const double uy = std::fmin(c[2] * g.ymin, c[2] * g.ymax), vy = std::fmax(c[2] * g.ymin, c[2] * g.ymax);
                    const double uz = std::fmin(0.0, c[3] * zb1), vz = std::fmax(0.0, c[3] * zb1);
                    const int w = (int)(std::floor(vx + vy + vz) - std::floor(ux + uy + uz)) + 3;
                    if (w > cs_max_width() - 2) ag.col_cs = false;
                }
            }
            if (ag.col_cs) {
                ag.ngrp = (int)grp.size();
                ALLOC(ag.d_col_grp, sizeof(ColGroup) * grp.size());
                if (upload_now(ag.d_col_grp, grp.data(), sizeof(ColGroup) * grp.size(), ctx->stream) != hipSuccess)
                    return undo(fail(MTR_ERR_HIP, "column group upload failed"));
                const std::vector<int2> win =
                    bp_cs_windows(grp.data(), ag.ngrp, xc, A, ctx->vol.Zstart, ctx->vol.Zend);
                ALLOC(ag.d_cs_win, sizeof(int2) * win.size());
                if (upload_now(ag.d_cs_win, win.data(), sizeof(int2) * win.size(), ctx->stream) != hipSuccess)
                    return undo(fail(MTR_ERR_HIP, "back-projection window upload failed"));
            }
        }
        if (upload_now(ag.d_col_xy, xy.data(), sizeof(int2) * xy.size(), ctx->stream) != hipSuccess ||
            upload_now(ag.d_col_w, wts.data(), sizeof(float2) * wts.size(), ctx->stream) != hipSuccess ||
            upload_now(ag.d_strip, cnt.data(), sizeof(int) * cnt.size(), ctx->stream) != hipSuccess)
            return undo(fail(MTR_ERR_HIP, "column table upload failed"));
    }
    ag.nfp = fp_blocks(make_dev(ctx, ag), ag.lanes_k);
    ALLOC(ag.d_partial, sizeof(double) * std::max(std::max(std::max(ag.nfp, fp_stream_blocks(make_dev(ctx, ag), 1)),
                                                           fp_col_res_blocks(make_dev(ctx, ag))),
                                                  1));
    ALLOC(ag.d_scalar, sizeof(double));
    if (ag.fp_seg > 1) ALLOC(ag.d_seg, sizeof(float) * P * ag.fp_seg);
    if (!seg_ord.empty()) {
        ALLOC(ag.d_seg_ord, sizeof(int) * seg_ord.size());
        if (upload_now(ag.d_seg_ord, seg_ord.data(), sizeof(int) * seg_ord.size(), ctx->stream) != hipSuccess)
            return undo(fail(MTR_ERR_HIP, "block order upload failed"));
    }
    if (hipMemcpyAsync(ag.d_coef, coef.data(), sizeof(ViewCoef) * A, hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess ||
        hipMemcpyAsync(ag.d_plan, plan.data(), sizeof(FpPlan) * A, hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess)
        return undo(fail(MTR_ERR_HIP, "geometry upload failed"));
    if (ctx->precision == MTR_PREC_FP32 && p_dev && !ag.rowfast) {
        ag.d_p = const_cast<float *>(p_dev);
        ag.p_owned = false;
    } else {
        ALLOC(ag.d_p, es * P);
        ag.p_owned = true;
        // p arrives FP32 [A][nv][nu]: convert to storage precision and, for
        // row-fastest agents, transpose each view; both through a staging
        // buffer where needed
        const bool direct = ctx->precision == MTR_PREC_FP32 && !ag.rowfast;
        if (p_host && direct) {
            if (hipMemcpyAsync(ag.d_p, p_host, sizeof(float) * P, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
                return undo(fail(MTR_ERR_HIP, "projection upload failed"));
        } else {
            const float *src = p_dev;
            float *tmp = nullptr;
            void *conv = nullptr;
            auto cleanup = [&]() {
                (void)hipStreamSynchronize(ctx->stream);
                dev_free(ctx, tmp);
                dev_free(ctx, conv);
            };
            if (p_host) {
                if (dev_alloc(ctx, (void **)&tmp, sizeof(float) * P) != hipSuccess)
                    return undo(fail(MTR_ERR_OOM, "staging buffer"));
                ag.dev_peak = std::max(ag.dev_peak, ag.dev_bytes + sizeof(float) * P);
                if (hipMemcpyAsync(tmp, p_host, sizeof(float) * P, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
                    cleanup();
                    return undo(fail(MTR_ERR_HIP, "projection upload failed"));
                }
                src = tmp;
            }
            hipError_t e = hipSuccess;
            if (ag.rowfast) {
                if (ctx->precision != MTR_PREC_FP32) {
                    if (dev_alloc(ctx, &conv, es * P) != hipSuccess) {
                        cleanup();
                        return undo(fail(MTR_ERR_OOM, "staging buffer"));
                    }
                    ag.dev_peak = std::max(ag.dev_peak, ag.dev_bytes + (tmp ? sizeof(float) * P : 0) + es * P);
                    e = launch_to_storage(ctx->precision, P, src, conv, ctx->stream);
                }
                const int64_t ss[3] = {1, prj->X, (int64_t)prj->X * prj->Y};
                const int64_t ds[3] = {prj->Y, 1, (int64_t)prj->X * prj->Y};
                if (e == hipSuccess)
                    e = launch_permute(ctx->precision, conv ? conv : (const void *)src, ag.d_p, prj->X, prj->Y,
                                       prj->AngN, ss, ds, ctx->stream);
            } else {
                e = launch_to_storage(ctx->precision, P, src, ag.d_p, ctx->stream);
            }
            cleanup();
            if (e != hipSuccess) return undo(fail(MTR_ERR_HIP, "projection conversion failed"));
        }
    }
#undef ALLOC
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return undo(fail(MTR_ERR_HIP, "upload sync failed"));
    if (launch_fill(ctx->precision, ctx->N, ag.d_x, 0.f, ctx->stream) != hipSuccess ||
        launch_fill(ctx->precision, ctx->N, ag.d_w, 0.f, ctx->stream) != hipSuccess)
        return undo(fail(MTR_ERR_HIP, "volume init failed"));
    int rc = upload_table(ctx);
    if (rc != MTR_OK) return undo(rc);
    if (ctx->step_graph) {
        (void)hipGraphExecDestroy(ctx->step_graph);
        ctx->step_graph = nullptr;
    }
    if (agent_id) *agent_id = (int)ctx->agents.size() - 1;
    return MTR_OK;
}

// Temporary single-agent device state for the stateless operators (FP32).
struct TmpAgent {
    DevBuf coef, plan, bsv, sepw, x, s, c, r, tab;
    AgentDev dev{};
    bool lanes_k = false;
    BpMode bpm;
    int nfp = 0;
};

int tmp_agent(mtr_ctx *ctx, const mtr_projection_size *prj, const mtr_volume_size *vol, const double *xc,
              const double *yc, TmpAgent &t)
{
    if (!prj_valid(prj)) return fail(MTR_ERR_SHAPE, "invalid projection size");
    if (!vol_valid(vol)) return fail(MTR_ERR_SHAPE, "invalid volume size");
    const int A = prj->AngN;
    std::vector<ViewCoef> coef(A);
    std::vector<FpPlan> plan(A);
    bool all_wx = true;
    for (int v = 0; v < A; v++) {
        for (int m = 0; m < 4; m++) {
            coef[v].c[m] = xc[4 * v + m];
            coef[v].c[4 + m] = yc[4 * v + m];
        }
        int rc = make_plan(xc + 4 * v, yc + 4 * v, plan[v]);
        if (rc != MTR_OK) return fail(rc, "view %d: degenerate or non-finite geometry", v);
        plan[v].pfx_view = v;
        all_wx = all_wx && plan[v].kind == 1 && plan[v].W == 0;
    }
    t.lanes_k = all_wx;
    t.bpm = bp_mode(plan, xc, yc, *vol, *prj);
    const int64_t P = (int64_t)prj->X * prj->Y * A;
    const int64_t N = (int64_t)vol->X * vol->Y * vol->Z;
    HIP_TRY(hipMalloc(&t.coef.p, sizeof(ViewCoef) * A));
    HIP_TRY(hipMalloc(&t.plan.p, sizeof(FpPlan) * A));
    HIP_TRY(hipMalloc(&t.x.p, sizeof(float) * N));
    HIP_TRY(hipMalloc(&t.s.p, sizeof(float) * P));
    HIP_TRY(hipMalloc(&t.c.p, sizeof(float) * P));
    HIP_TRY(hipMalloc(&t.r.p, sizeof(float) * P));
    HIP_TRY(hipMalloc(&t.tab.p, sizeof(AgentDev)));
    HIP_TRY(hipMemcpyAsync(t.coef.p, coef.data(), sizeof(ViewCoef) * A, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(t.plan.p, plan.data(), sizeof(FpPlan) * A, hipMemcpyHostToDevice, ctx->stream));
    const std::vector<BpSepView> bsv = bp_sep_views(t.bpm, xc, yc, A);
    if (!bsv.empty()) {
        HIP_TRY(hipMalloc(&t.bsv.p, sizeof(BpSepView) * A));
        HIP_TRY(hipMemcpyAsync(t.bsv.p, bsv.data(), sizeof(BpSepView) * A, hipMemcpyHostToDevice, ctx->stream));
    }
    AgentDev &d = t.dev;
    d.nu = prj->X;
    d.nv = prj->Y;
    d.A = A;
    d.Xs = vol->Xstart;
    d.Xe = vol->Xend;
    d.Ys = vol->Ystart;
    d.Ye = vol->Yend;
    d.Zs = vol->Zstart;
    d.Ze = vol->Zend;
    d.X = vol->X;
    d.Y = vol->Y;
    d.Z = vol->Z;
    d.coef = (const ViewCoef *)t.coef.p;
    d.plan = (const FpPlan *)t.plan.p;
    d.bsv = (const BpSepView *)t.bsv.p;
    d.pk = prj->X;
    d.pi = 1;
    d.p = t.r.p;  // the stateless BP/BPT read their input from r
    d.c = (const float *)t.c.p;
    d.r = t.r.p;
    d.x = t.x.p;
    d.partial = nullptr;
    d.sx = 1;
    d.sy = vol->X;
    d.sz = (int64_t)vol->X * vol->Y;
    std::vector<int2> sepw;
    if (!bsv.empty()) {
        sepw = bp_sep_windows(bsv.data(), A, d, t.bpm.P1, t.bpm.P2);
        HIP_TRY(hipMalloc(&t.sepw.p, sizeof(int2) * sepw.size()));
        HIP_TRY(hipMemcpyAsync(t.sepw.p, sepw.data(), sizeof(int2) * sepw.size(), hipMemcpyHostToDevice, ctx->stream));
        d.sep_win = (const int2 *)t.sepw.p;
    }
    HIP_TRY(hipMemcpyAsync(t.tab.p, &d, sizeof(AgentDev), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));  // the host tables are freed on return
    t.nfp = fp_blocks(d, t.lanes_k);
    return MTR_OK;
}

// Canonical ([Z][Y][X]) view of an agent's working volume: the volume
// itself, or its canonical copy for agents stored with another fast axis.
void *work_canon(Agent &ag) { return ag.fast ? ag.d_tmp : ag.d_x; }

int to_canon(mtr_ctx *ctx, Agent &ag)
{
    if (!ag.fast) return MTR_OK;
    int64_t cs[3];
    layout_strides(0, ctx->vol.X, ctx->vol.Y, ctx->vol.Z, cs);
    Timed t(ctx, MTR_K_ELEMENTWISE);
    HIP_TRY(launch_permute(ctx->precision, ag.d_x, ag.d_tmp, ctx->vol.X, ctx->vol.Y, ctx->vol.Z, ag.st, cs,
                           ctx->stream));
    return MTR_OK;
}

int from_canon(mtr_ctx *ctx, Agent &ag)
{
    if (!ag.fast) return MTR_OK;
    int64_t cs[3];
    layout_strides(0, ctx->vol.X, ctx->vol.Y, ctx->vol.Z, cs);
    Timed t(ctx, MTR_K_ELEMENTWISE);
    HIP_TRY(launch_permute(ctx->precision, ag.d_tmp, ag.d_x, ctx->vol.X, ctx->vol.Y, ctx->vol.Z, cs, ag.st,
                           ctx->stream));
    return MTR_OK;
}

void *vol_ptr(mtr_ctx *ctx, int which, int agent)
{
    if (which == MTR_VOL_WBAR) return ctx->d_wbar;
    Agent &ag = ctx->agents[agent];
    return which == MTR_VOL_WORK ? work_canon(ag) : ag.d_w;
}

int copy_volume_out(mtr_ctx *ctx, const void *dev, float *host)
{
    if (ctx->precision == MTR_PREC_FP32) {
        HIP_TRY(hipMemcpyAsync(host, dev, sizeof(float) * ctx->N, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return MTR_OK;
    }
    CtxBuf tmp{ctx};
    HIP_TRY(dev_alloc(ctx, &tmp.p, sizeof(float) * ctx->N));
    HIP_TRY(launch_from_storage(ctx->precision, ctx->N, dev, (float *)tmp.p, ctx->stream));
    HIP_TRY(hipMemcpyAsync(host, tmp.p, sizeof(float) * ctx->N, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MTR_OK;
}

int copy_volume_in(mtr_ctx *ctx, void *dev, const float *host)
{
    if (ctx->precision == MTR_PREC_FP32) {
        HIP_TRY(hipMemcpyAsync(dev, host, sizeof(float) * ctx->N, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return MTR_OK;
    }
    CtxBuf tmp{ctx};
    HIP_TRY(dev_alloc(ctx, &tmp.p, sizeof(float) * ctx->N));
    HIP_TRY(hipMemcpyAsync(tmp.p, host, sizeof(float) * ctx->N, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_to_storage(ctx->precision, ctx->N, (const float *)tmp.p, dev, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MTR_OK;
}

// Device tables of a context: the agent table in agent order and sorted
// by kernel kind, and the consensus partial sums.
int scratch_init(int device);  // the allocation-free entry point's per-device scratch

int alloc_tables(mtr_ctx *c)
{
    if (dev_alloc(c, (void **)&c->d_ags, sizeof(AgentDev) * kMaxAgents) != hipSuccess ||
        dev_alloc(c, (void **)&c->d_ags_fp, sizeof(AgentDev) * kMaxAgents) != hipSuccess ||
        dev_alloc(c, (void **)&c->d_ags_bp, sizeof(AgentDev) * kMaxAgents) != hipSuccess ||
        dev_alloc(c, (void **)&c->d_cons_partial, sizeof(double) * 8192) != hipSuccess ||
        dev_alloc(c, (void **)&c->d_cons_scalar, sizeof(double)) != hipSuccess)
        return fail(MTR_ERR_OOM, "context tables");
    return MTR_OK;
}

}  // namespace

// ======================================================================
extern "C" {

int mtr_abi_version(void) { return MTR_ABI_VERSION; }

const char *mtr_last_error(void) { return g_err.c_str(); }

int mtr_ctx_create(int device, mtr_ctx **out)
{
    if (!out) return fail(MTR_ERR_ARG, "null out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(MTR_ERR_NODEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(MTR_ERR_RANGE, "device %d out of range (%d visible)", device, n);
    mtr_ctx *ctx = new mtr_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return fail(MTR_ERR_HIP, "stream creation failed");
    }
    if (hipStreamCreate(&ctx->aux) != hipSuccess || hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess) {
        mtr_ctx_destroy(ctx);
        return fail(MTR_ERR_HIP, "stream creation failed");
    }
    if (const char *ks = std::getenv("MTR_KIND_STREAMS")) ctx->kind_streams = std::atoi(ks) != 0;
    if (const int rc = alloc_tables(ctx); rc != MTR_OK) {
        mtr_ctx_destroy(ctx);
        return rc;
    }
    if (const int rc = scratch_init(device); rc != MTR_OK) {
        mtr_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return MTR_OK;
}

int mtr_ctx_destroy(mtr_ctx *ctx)
{
    if (!ctx) return MTR_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &ag : ctx->agents) {
        for (void *q : {(void *)ag.d_coef, (void *)ag.d_plan, (void *)ag.d_bsv, ag.p_owned ? ag.d_p : nullptr, (void *)ag.d_c, ag.d_r,
                        ag.d_x, ag.d_w, ag.d_tmp, (void *)ag.d_partial, (void *)ag.d_scalar, (void *)ag.d_p_std,
                        (void *)ag.d_col_xy, (void *)ag.d_col_w, (void *)ag.d_strip, (void *)ag.d_pm, (void *)ag.d_pn,
                        (void *)ag.d_bpc, (void *)ag.d_col_grp, (void *)ag.d_cs_win, (void *)ag.d_sep_win, (void *)ag.d_seg, (void *)ag.d_seg_ord})
            dev_free(ctx, q);
    }
    for (void *q : {ctx->d_wbar, (void *)ctx->d_ags, (void *)ctx->d_ags_fp, (void *)ctx->d_ags_bp,
                    (void *)ctx->d_cons_partial, (void *)ctx->d_cons_scalar, (void *)ctx->d_s, (void *)ctx->d_cc})
        dev_free(ctx, q);
    for (auto &pd : ctx->pending) {
        (void)hipEventDestroy(pd.a);
        (void)hipEventDestroy(pd.b);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->step_graph) (void)hipGraphExecDestroy(ctx->step_graph);
    if (ctx->aux) {
        (void)hipStreamSynchronize(ctx->aux);
        (void)hipStreamDestroy(ctx->aux);
    }
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->stream && ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MTR_OK;
}

int mtr_ctx_autotune(mtr_ctx *ctx, int on)
{
    if (!ctx) return fail(MTR_ERR_ARG, "null context");
    ctx->autotune = on != 0;
    return MTR_OK;
}

int mtr_agent_tiles(const mtr_ctx *ctx, int agent, int *run, int *lanes_axis)
{
    MTR_TRY(check_agent(const_cast<mtr_ctx *>(ctx), agent));
    const Agent &ag = ctx->agents[agent];
    if (run) *run = ag.col.on ? ag.col.rw : (ag.bpm.sep ? ag.bpm.rw : 0);
    if (lanes_axis) *lanes_axis = ag.col.on ? ag.col.p1y : (ag.bpm.sep ? ag.bpm.P1 : 0);
    return MTR_OK;
}

int mtr_agent_fp_window(const mtr_ctx *ctx, int agent, int *slots)
{
    MTR_TRY(check_agent(const_cast<mtr_ctx *>(ctx), agent));
    if (!slots) return fail(MTR_ERR_ARG, "null slots");
    *slots = ctx->agents[agent].col.on ? ctx->agents[agent].col_win : 0;
    return MTR_OK;
}

int mtr_ctx_use_graphs(mtr_ctx *ctx, int on)
{
    if (!ctx) return fail(MTR_ERR_ARG, "null context");
    ctx->use_graphs = on != 0;
    return MTR_OK;
}

int mtr_ctx_set_precision(mtr_ctx *ctx, int precision)
{
    if (!ctx) return fail(MTR_ERR_ARG, "null context");
    if (precision != MTR_PREC_FP32 && precision != MTR_PREC_BF16 && precision != MTR_PREC_FP16)
        return fail(MTR_ERR_ARG, "unknown precision %d", precision);
    if (ctx->vol_set) return fail(MTR_ERR_STATE, "precision must be set before the volume");
    ctx->precision = precision;
    return MTR_OK;
}

int mtr_ctx_set_volume(mtr_ctx *ctx, const mtr_volume_size *vol)
{
    if (!ctx) return fail(MTR_ERR_ARG, "null context");
    if (!vol_valid(vol)) return fail(MTR_ERR_SHAPE, "invalid volume box");
    if (ctx->vol_set) return fail(MTR_ERR_STATE, "volume already set");
    MTR_TRY(set_device(ctx));
    ctx->vol = *vol;
    ctx->N = (int64_t)vol->X * vol->Y * vol->Z;
    HIP_TRY(dev_alloc(ctx, &ctx->d_wbar, elem_size(ctx->precision) * ctx->N));
    HIP_TRY(launch_fill(ctx->precision, ctx->N, ctx->d_wbar, 0.f, ctx->stream));
    ctx->vol_set = true;
    return MTR_OK;
}

int mtr_agent_add(mtr_ctx *ctx, const mtr_projection_size *prj, const double *xc, const double *yc,
                  const float *p_host, int *agent_id)
{
    return agent_add_common(ctx, prj, xc, yc, p_host, nullptr, agent_id);
}

int mtr_agent_add_device(mtr_ctx *ctx, const mtr_projection_size *prj, const double *xc, const double *yc,
                         const float *p_dev, int *agent_id)
{
    return agent_add_common(ctx, prj, xc, yc, nullptr, p_dev, agent_id);
}

int mtr_agent_count(const mtr_ctx *ctx, int *count)
{
    if (!ctx || !count) return fail(MTR_ERR_ARG, "null argument");
    *count = (int)ctx->agents.size();
    return MTR_OK;
}

int mtr_agent_footprint(const mtr_ctx *ctx, const mtr_projection_size *prj, size_t *bytes)
{
    if (!ctx || !bytes) return fail(MTR_ERR_ARG, "null argument");
    if (!ctx->vol_set) return fail(MTR_ERR_STATE, "volume not set");
    if (!prj_valid(prj)) return fail(MTR_ERR_SHAPE, "invalid projection size");
    *bytes = agent_bytes(ctx->precision, ctx->vol, *prj);
    return MTR_OK;
}

int mtr_agent_device_bytes(const mtr_ctx *ctx, int agent, size_t *resident, size_t *peak)
{
    MTR_TRY(check_agent(const_cast<mtr_ctx *>(ctx), agent));
    if (!resident || !peak) return fail(MTR_ERR_ARG, "null argument");
    *resident = ctx->agents[agent].dev_bytes;
    *peak = ctx->agents[agent].dev_peak;
    return MTR_OK;
}

int mtr_agent_kernels(const mtr_ctx *ctx, int agent, int *fp_kind, int *bp_kind)
{
    MTR_TRY(check_agent(const_cast<mtr_ctx *>(ctx), agent));
    const Agent &ag = ctx->agents[agent];
    if (fp_kind) *fp_kind = ag.col.on ? MTR_FPK_COLUMN : (ag.fp_nc ? MTR_FPK_STREAM : MTR_FPK_GENERAL);
    if (bp_kind) *bp_kind = ag.col.on ? MTR_BPK_COLUMN : (ag.bpm.sep ? MTR_BPK_SEPARABLE : MTR_BPK_GENERAL);
    return MTR_OK;
}

int mtr_bpt(mtr_ctx *ctx, int agent, float *out_host)
{
    MTR_TRY(check_agent(ctx, agent));
    MTR_TRY(set_device(ctx));
    MTR_TRY(ensure_normaliser(ctx, agent));
    MTR_TRY(tune_agent(ctx, agent));
    MTR_TRY(bp_group(ctx, agent, 1, true, 1.0f));
    if (out_host) {
        MTR_TRY(to_canon(ctx, ctx->agents[agent]));
        MTR_TRY(copy_volume_out(ctx, work_canon(ctx->agents[agent]), out_host));
    }
    return collect_timing(ctx);
}

int mtr_sirt(mtr_ctx *ctx, int agent, const float *x0_host, int iters, float lambda, float *out_host)
{
    MTR_TRY(check_agent(ctx, agent));
    if (iters < 0) return fail(MTR_ERR_ARG, "iters < 0");
    MTR_TRY(set_device(ctx));
    Agent &ag = ctx->agents[agent];
    if (x0_host) {
        MTR_TRY(copy_volume_in(ctx, work_canon(ag), x0_host));
        MTR_TRY(from_canon(ctx, ag));
    }
    MTR_TRY(sirt_fast(ctx, {agent}, iters, lambda));
    if (out_host) {
        MTR_TRY(to_canon(ctx, ag));
        MTR_TRY(copy_volume_out(ctx, work_canon(ag), out_host));
    }
    return collect_timing(ctx);
}

int mtr_form_z(mtr_ctx *ctx, int agent)
{
    MTR_TRY(check_agent(ctx, agent));
    MTR_TRY(set_device(ctx));
    Agent &ag = ctx->agents[agent];
    {
        Timed t(ctx, MTR_K_ELEMENTWISE);
        HIP_TRY(launch_form_z(ctx->precision, ctx->N, ag.d_w, ctx->d_wbar, work_canon(ag), ctx->stream));
    }
    MTR_TRY(from_canon(ctx, ag));
    return collect_timing(ctx);
}

int mtr_mann(mtr_ctx *ctx, int agent, float rho)
{
    MTR_TRY(check_agent(ctx, agent));
    MTR_TRY(set_device(ctx));
    Agent &ag = ctx->agents[agent];
    MTR_TRY(to_canon(ctx, ag));
    {
        Timed t(ctx, MTR_K_ELEMENTWISE);
        HIP_TRY(launch_mann_z(ctx->precision, ctx->N, work_canon(ag), ag.d_w, ctx->d_wbar, rho, ctx->stream));
    }
    return collect_timing(ctx);
}

int mtr_mean(mtr_ctx *ctx)
{
    if (!ctx) return fail(MTR_ERR_ARG, "null context");
    if (ctx->agents.empty()) return fail(MTR_ERR_STATE, "no agents");
    MTR_TRY(set_device(ctx));
    std::vector<const void *> w;
    for (auto &ag : ctx->agents) w.push_back(ag.d_w);
    {
        Timed t(ctx, MTR_K_ELEMENTWISE);
        HIP_TRY(launch_mean(ctx->precision, (int)w.size(), ctx->N, w.data(), ctx->d_wbar, ctx->stream));
    }
    return collect_timing(ctx);
}

// B = sum_a BPT_a into w-bar and every w_a, first working volumes z_a = B.
// The working volumes must already hold BPT_a (agent layout).
static int init_from_bpt(mtr_ctx *ctx)
{
    const int K = (int)ctx->agents.size();
    std::vector<void *> w, x;
    for (auto &ag : ctx->agents) {
        MTR_TRY(to_canon(ctx, ag));
        w.push_back(ag.d_w);
        x.push_back(work_canon(ag));
    }
    {
        Timed t(ctx, MTR_K_ELEMENTWISE);
        HIP_TRY(launch_init_sum(ctx->precision, K, ctx->N, w.data(), x.data(), ctx->d_wbar, ctx->stream));
    }
    for (auto &ag : ctx->agents) MTR_TRY(from_canon(ctx, ag));
    return MTR_OK;
}

// Joint choice of the separable back-projections' run lengths.  In the
// whole-context solve each agent's back-projection follows the forward
// projection of every agent and the back-projections sorted before it, and
// the cache state this leaves can favour another run length than the one an
// agent is fastest with alone (at 512x512x128 with axes at 0 and 90 deg:
// runs of 16 for both agents, against 32 chosen one agent at a time).  Run
// after the initial model: it times whole-context SIRT iterations with
// lambda = 0, whose update adds fl32(S/C) * 0 = 0 to every voxel, so the
// working volumes keep their values; the residuals are rewritten by the
// solve's first forward projection.  Candidates are the per-agent choice and
// every separable agent at one common run length, timed in turns (only the
// separable back-projections, each after a whole-context forward
// projection), and the fastest is kept and cached under a hash of the keys
// of all the context's agents.
static int tune_joint(mtr_ctx *ctx)
{
    const int K = (int)ctx->agents.size();
    std::vector<int> sep;
    for (int a = 0; a < K; a++)
        if (!ctx->agents[a].col.on && ctx->agents[a].bpm.sep) sep.push_back(a);
    if (!ctx->autotune || !ctx->joint_tune || sep.empty()) return MTR_OK;
    std::string all;
    for (const auto &ag : ctx->agents) all += tune_key(ctx, ag) + "|";
    char key[64];
    std::snprintf(key, sizeof key, "joint K%d %016llx", K, (unsigned long long)fnv(all.data(), all.size()));
    if (ctx->joint_key == key) return MTR_OK;
    std::vector<int> own;
    for (int a : sep) own.push_back(ctx->agents[a].bpm.rw);
    auto apply = [&](int run) -> int {
        for (size_t q = 0; q < sep.size(); q++) MTR_TRY(apply_tiles(ctx, sep[q], run ? run : own[q], 0));
        return MTR_OK;
    };
    TuneCache &tc = tune_cache();
    {
        std::lock_guard<std::mutex> lk(tc.mu);
        tune_cache_load(tc);
        auto it = tc.kv.find(key);
        if (it != tc.kv.end()) {
            ctx->joint_key = key;
            return apply(it->second.first);
        }
    }
    std::vector<int> cand = {0};  // 0: the per-agent choice
    for (int run : {16, 32}) {
        bool same = true;
        for (int r : own) same = same && r == run;
        if (!same) cand.push_back(run);
    }
    const bool timing = ctx->timing;
    ctx->timing = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    std::vector<float> best(cand.size(), 1e30f);
    int rc = MTR_OK;
    for (int rep = 0; rep < 4 && rc == MTR_OK; rep++)
        for (size_t c = 0; c < cand.size() && rc == MTR_OK; c++) {
            if ((rc = apply(cand[c])) != MTR_OK) break;
            // the separable agents are contiguous in the kind-sorted table
            int s0 = K, s1 = 0;
            for (int i = 0; i < K; i++) {
                const Agent &ag = ctx->agents[ctx->bp_order[i]];
                if (!ag.col.on && ag.bpm.sep) {
                    s0 = std::min(s0, i);
                    s1 = i + 1;
                }
            }
            const Batch F = ctx_sorted(ctx, ctx->fp_order, ctx->d_ags_fp);
            const Batch Bk = ctx_sorted(ctx, ctx->bp_order, ctx->d_ags_bp);
            float t = 0.f;
            if (fp_batch(ctx, F, 0, K) != MTR_OK || bp_batch(ctx, Bk, 0, s0, false, 0.0f) != MTR_OK ||
                hipEventRecord(e0, ctx->stream) != hipSuccess || bp_batch(ctx, Bk, s0, s1 - s0, false, 0.0f) != MTR_OK ||
                hipEventRecord(e1, ctx->stream) != hipSuccess || bp_batch(ctx, Bk, s1, K - s1, false, 0.0f) != MTR_OK ||
                hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)
                rc = fail(MTR_ERR_HIP, "joint autotuner timing failed");
            best[c] = std::min(best[c], t);
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->timing = timing;
    if (rc != MTR_OK) return rc;
    size_t pick = 0;
    for (size_t c = 1; c < cand.size(); c++)
        if (best[c] < best[pick]) pick = c;
    {
        std::lock_guard<std::mutex> lk(tc.mu);
        tune_cache_store(tc, key, cand[pick], 0);
    }
    ctx->joint_key = key;
    return apply(cand[pick]);
}

int mtr_mace_init(mtr_ctx *ctx)
{
    if (!ctx) return fail(MTR_ERR_ARG, "null context");
    if (ctx->agents.empty()) return fail(MTR_ERR_STATE, "no agents");
    MTR_TRY(set_device(ctx));
    for (int a = 0; a < (int)ctx->agents.size(); a++) MTR_TRY(ensure_normaliser(ctx, a));
    MTR_TRY(tune_all(ctx));
    // BPT of every agent into its working volume, batched per kernel kind.
    MTR_TRY(bp_group(ctx, 0, (int)ctx->agents.size(), true, 1.0f));
    MTR_TRY(init_from_bpt(ctx));
    return tune_joint(ctx);
}

// One fast MACE step (MultiAxis.cpp:40-71): S SIRT iterations of every
// agent, then the fused consensus in the canonical layout.  Launches only:
// no allocation, synchronisation or memset, so it can be captured.
static int mace_step(mtr_ctx *ctx, int S, float lambda, float rho)
{
    const int K = (int)ctx->agents.size();
    std::vector<int> ids(K);
    for (int a = 0; a < K; a++) ids[a] = a;
    std::vector<void *> w, x;
    for (auto &ag : ctx->agents) {
        w.push_back(ag.d_w);
        x.push_back(work_canon(ag));
    }
    MTR_TRY(sirt_fast(ctx, ids, S, lambda));
    for (auto &ag : ctx->agents) MTR_TRY(to_canon(ctx, ag));
    {
        Timed t(ctx, MTR_K_CONSENSUS);
        HIP_TRY(launch_consensus(ctx->precision, K, ctx->N, rho, 1, w.data(), x.data(), ctx->d_wbar,
                                 ctx->d_cons_partial, consensus_blocks(ctx->N), ctx->stream));
    }
    for (auto &ag : ctx->agents) MTR_TRY(from_canon(ctx, ag));
    return MTR_OK;
}

// Captures mace_step into a graph (kept until the agents or the solve
// parameters change) so that a solve replays one graph per step.
static int step_graph(mtr_ctx *ctx, int S, float lambda, float rho)
{
    if (ctx->step_graph && ctx->graph_S == S && ctx->graph_lambda == lambda && ctx->graph_rho == rho)
        return MTR_OK;
    if (ctx->step_graph) {
        (void)hipGraphExecDestroy(ctx->step_graph);
        ctx->step_graph = nullptr;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = mace_step(ctx, S, lambda, rho);
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc != MTR_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    HIP_TRY(e);
    const hipError_t ei = hipGraphInstantiate(&ctx->step_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_TRY(ei);
    ctx->graph_S = S;
    ctx->graph_lambda = lambda;
    ctx->graph_rho = rho;
    return MTR_OK;
}

// The FP64 consensus difference of the last step (mtr_consensus_diff2).
static int consensus_d2(mtr_ctx *ctx, double *out)
{
    HIP_TRY(launch_reduce(ctx->d_cons_partial, ctx->cons_blocks, ctx->d_cons_scalar, ctx->stream));
    HIP_TRY(hipMemcpyAsync(out, ctx->d_cons_scalar, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MTR_OK;
}

int mtr_mace_continue(mtr_ctx *ctx, int M, int S, float lambda, float rho, double rel_tol, int *steps_run,
                      float *wbar_host)
{
    if (!ctx) return fail(MTR_ERR_ARG, "null context");
    if (ctx->agents.empty()) return fail(MTR_ERR_STATE, "no agents");
    if (M < 0 || S < 0) return fail(MTR_ERR_ARG, "negative iteration count");
    if (!(rel_tol >= 0.0)) return fail(MTR_ERR_ARG, "rel_tol must be >= 0");
    MTR_TRY(set_device(ctx));
    for (int a = 0; a < (int)ctx->agents.size(); a++) MTR_TRY(ensure_normaliser(ctx, a));
    MTR_TRY(tune_all(ctx));
    // working volumes from the state, z_a = 2 w-bar - w_a: the formula and
    // rounding a consensus step uses for the next step's input
    for (auto &ag : ctx->agents) {
        {
            Timed t(ctx, MTR_K_ELEMENTWISE);
            HIP_TRY(launch_form_z(ctx->precision, ctx->N, ag.d_w, ctx->d_wbar, work_canon(ag), ctx->stream));
        }
        MTR_TRY(from_canon(ctx, ag));
    }
    const int cb = consensus_blocks(ctx->N);
    const bool graph = ctx->use_graphs && !ctx->timing && M > 0;
    if (graph) MTR_TRY(step_graph(ctx, S, lambda, rho));
    int m = 0;
    double prev = -1.0;
    while (m < M) {
        if (graph) HIP_TRY(hipGraphLaunch(ctx->step_graph, ctx->stream));
        else MTR_TRY(mace_step(ctx, S, lambda, rho));
        m++;
        ctx->cons_blocks = cb;
        if (rel_tol > 0.0) {
            double d2 = 0.0;
            MTR_TRY(consensus_d2(ctx, &d2));
            if (prev >= 0.0 && std::fabs(d2 - prev) <= rel_tol * prev) break;
            prev = d2;
        }
    }
    if (steps_run) *steps_run = m;
    if (wbar_host) MTR_TRY(copy_volume_out(ctx, ctx->d_wbar, wbar_host));
    return collect_timing(ctx);
}

int mtr_mace_solve(mtr_ctx *ctx, int M, int S, float lambda, float rho, int form, float *wbar_host)
{
    if (!ctx) return fail(MTR_ERR_ARG, "null context");
    if (ctx->agents.empty()) return fail(MTR_ERR_STATE, "no agents");
    if (M < 0 || S < 0) return fail(MTR_ERR_ARG, "negative iteration count");
    if (form != MTR_FORM_FAST && form != MTR_FORM_PLAIN) return fail(MTR_ERR_ARG, "unknown form %d", form);
    MTR_TRY(set_device(ctx));
    const int K = (int)ctx->agents.size();
    if (form == MTR_FORM_PLAIN) {
        if (ctx->precision != MTR_PREC_FP32) return fail(MTR_ERR_STATE, "plain form runs FP32 storage only");
        // BPT per agent, sum, then per step: z, SIRT, Mann per agent, mean.
        for (int a = 0; a < K; a++) {
            const float *pstd = nullptr;
            MTR_TRY(plain_p(ctx, ctx->agents[a], &pstd));
            Timed t(ctx, MTR_K_BPT);
            HIP_TRY(launch_plain_bp(make_dev(ctx, ctx->agents[a]), pstd, 1.0f, 1, ctx->stream));
        }
        MTR_TRY(init_from_bpt(ctx));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::vector<const void *> wc;
        for (auto &ag : ctx->agents) wc.push_back(ag.d_w);
        for (int m = 0; m < M; m++) {
            for (int a = 0; a < K; a++) {
                Agent &ag = ctx->agents[a];
                {
                    Timed t(ctx, MTR_K_ELEMENTWISE);
                    HIP_TRY(launch_form_z(ctx->precision, ctx->N, ag.d_w, ctx->d_wbar, work_canon(ag), ctx->stream));
                }
                MTR_TRY(from_canon(ctx, ag));
                MTR_TRY(sirt_plain(ctx, a, S, lambda));
                MTR_TRY(to_canon(ctx, ag));
                {
                    Timed t(ctx, MTR_K_ELEMENTWISE);
                    HIP_TRY(launch_mann_z(ctx->precision, ctx->N, work_canon(ag), ag.d_w, ctx->d_wbar, rho,
                                          ctx->stream));
                }
                HIP_TRY(hipStreamSynchronize(ctx->stream));
            }
            {
                Timed t(ctx, MTR_K_ELEMENTWISE);
                HIP_TRY(launch_mean(ctx->precision, K, ctx->N, wc.data(), ctx->d_wbar, ctx->stream));
            }
            HIP_TRY(hipStreamSynchronize(ctx->stream));
        }
    } else {
        MTR_TRY(mtr_mace_init(ctx));
        for (int a = 0; a < K; a++) MTR_TRY(ensure_normaliser(ctx, a));
        const int cb = consensus_blocks(ctx->N);
        // Per-launch timing records events around every kernel, which a
        // graph does not hold: timed solves launch eagerly.
        if (ctx->use_graphs && !ctx->timing && M > 0) {
            MTR_TRY(step_graph(ctx, S, lambda, rho));
            for (int m = 0; m < M; m++) HIP_TRY(hipGraphLaunch(ctx->step_graph, ctx->stream));
        } else {
            for (int m = 0; m < M; m++) MTR_TRY(mace_step(ctx, S, lambda, rho));
        }
        ctx->cons_blocks = M > 0 ? cb : 0;
    }
    if (wbar_host) MTR_TRY(copy_volume_out(ctx, ctx->d_wbar, wbar_host));
    return collect_timing(ctx);
}

int mtr_get_volume(mtr_ctx *ctx, int which, int agent, float *host)
{
    if (!ctx || !host) return fail(MTR_ERR_ARG, "null argument");
    if (!ctx->vol_set) return fail(MTR_ERR_STATE, "volume not set");
    if (which != MTR_VOL_WBAR) MTR_TRY(check_agent(ctx, agent));
    if (which < 0 || which > 2) return fail(MTR_ERR_ARG, "unknown volume %d", which);
    MTR_TRY(set_device(ctx));
    if (which == MTR_VOL_WORK) MTR_TRY(to_canon(ctx, ctx->agents[agent]));
    return copy_volume_out(ctx, vol_ptr(ctx, which, agent), host);
}

int mtr_set_volume(mtr_ctx *ctx, int which, int agent, const float *host)
{
    if (!ctx || !host) return fail(MTR_ERR_ARG, "null argument");
    if (!ctx->vol_set) return fail(MTR_ERR_STATE, "volume not set");
    if (which != MTR_VOL_WBAR) MTR_TRY(check_agent(ctx, agent));
    if (which < 0 || which > 2) return fail(MTR_ERR_ARG, "unknown volume %d", which);
    MTR_TRY(set_device(ctx));
    MTR_TRY(copy_volume_in(ctx, vol_ptr(ctx, which, agent), host));
    if (which == MTR_VOL_WORK) MTR_TRY(from_canon(ctx, ctx->agents[agent]));
    return MTR_OK;
}

// ---- MACE checkpoint file ------------------------------------------
// A 64-byte header, then the K + 1 state volumes w_0 .. w_{K-1}, w-bar as
// float32 [Z][Y][X], little-endian.  The geometry hash ties a checkpoint
// to the inputs it was taken from (every agent's detector size and
// projection coefficients, in agent order).
namespace {

struct CkptHeader {
    char magic[8];
    int32_t K, X, Y, Z, Xstart, Ystart, Zstart;
    int32_t steps_done;
    int32_t precision;
    int32_t reserved0;
    uint64_t geometry;
    uint64_t reserved1;
};
static_assert(sizeof(CkptHeader) == 64, "checkpoint header is 64 bytes");
constexpr char kCkptMagic[8] = {'M', 'T', 'R', 'C', 'K', 'P', 'T', '1'};

uint64_t geometry_hash(const mtr_ctx *ctx)
{
    uint64_t h = 1469598103934665603ull;
    for (const Agent &ag : ctx->agents) {
        const int32_t d[3] = {ag.prj.X, ag.prj.Y, ag.prj.AngN};
        h = fnv(d, sizeof d, h);
        h = fnv(ag.xc.data(), sizeof(double) * ag.xc.size(), h);
        h = fnv(ag.yc.data(), sizeof(double) * ag.yc.size(), h);
    }
    return h;
}

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) std::fclose(f);
    }
};

}  // namespace

int mtr_checkpoint_save(mtr_ctx *ctx, const char *path, int steps_done)
{
    if (!ctx || !path) return fail(MTR_ERR_ARG, "null argument");
    if (ctx->agents.empty()) return fail(MTR_ERR_STATE, "no agents");
    if (steps_done < 0) return fail(MTR_ERR_ARG, "negative step count");
    MTR_TRY(set_device(ctx));
    CkptHeader h{};
    std::memcpy(h.magic, kCkptMagic, sizeof h.magic);
    h.K = (int32_t)ctx->agents.size();
    h.X = ctx->vol.X, h.Y = ctx->vol.Y, h.Z = ctx->vol.Z;
    h.Xstart = ctx->vol.Xstart, h.Ystart = ctx->vol.Ystart, h.Zstart = ctx->vol.Zstart;
    h.steps_done = steps_done;
    h.precision = ctx->precision;
    h.geometry = geometry_hash(ctx);
    // written beside the target and renamed over it once complete and on
    // disk, so an interrupted save leaves any earlier checkpoint intact
    const std::string tmp = std::string(path) + ".tmp";
    int rc = MTR_OK;
    {
        File out;
        if (!(out.f = std::fopen(tmp.c_str(), "wb"))) return fail(MTR_ERR_ARG, "cannot create %s", tmp.c_str());
        if (std::fwrite(&h, sizeof h, 1, out.f) != 1) rc = fail(MTR_ERR_ARG, "write failed: %s", tmp.c_str());
        std::vector<float> buf((size_t)ctx->N);
        for (int v = 0; v <= h.K && rc == MTR_OK; v++) {
            rc = copy_volume_out(ctx, v < h.K ? ctx->agents[v].d_w : ctx->d_wbar, buf.data());
            if (rc == MTR_OK && std::fwrite(buf.data(), sizeof(float), buf.size(), out.f) != buf.size())
                rc = fail(MTR_ERR_ARG, "write failed: %s", tmp.c_str());
        }
        if (rc == MTR_OK && (std::fflush(out.f) != 0 || fsync(fileno(out.f)) != 0))
            rc = fail(MTR_ERR_ARG, "write failed: %s", tmp.c_str());
        const int cr = std::fclose(out.f);
        out.f = nullptr;
        if (rc == MTR_OK && cr != 0) rc = fail(MTR_ERR_ARG, "write failed: %s", tmp.c_str());
    }
    if (rc == MTR_OK && std::rename(tmp.c_str(), path) != 0)
        rc = fail(MTR_ERR_ARG, "cannot replace %s", path);
    if (rc != MTR_OK) (void)std::remove(tmp.c_str());
    return rc;
}

int mtr_checkpoint_load(mtr_ctx *ctx, const char *path, int *steps_done)
{
    if (!ctx || !path) return fail(MTR_ERR_ARG, "null argument");
    if (ctx->agents.empty()) return fail(MTR_ERR_STATE, "no agents");
    MTR_TRY(set_device(ctx));
    File in;
    if (!(in.f = std::fopen(path, "rb"))) return fail(MTR_ERR_ARG, "cannot open %s", path);
    CkptHeader h{};
    if (std::fread(&h, sizeof h, 1, in.f) != 1 || std::memcmp(h.magic, kCkptMagic, sizeof h.magic) != 0)
        return fail(MTR_ERR_SHAPE, "%s is not a MACE checkpoint", path);
    if (h.K != (int32_t)ctx->agents.size() || h.X != ctx->vol.X || h.Y != ctx->vol.Y || h.Z != ctx->vol.Z ||
        h.Xstart != ctx->vol.Xstart || h.Ystart != ctx->vol.Ystart || h.Zstart != ctx->vol.Zstart)
        return fail(MTR_ERR_SHAPE, "checkpoint %s holds %d agents on a %dx%dx%d box, the context %d on %dx%dx%d",
                    path, h.K, h.X, h.Y, h.Z, (int)ctx->agents.size(), ctx->vol.X, ctx->vol.Y, ctx->vol.Z);
    if (h.geometry != geometry_hash(ctx))
        return fail(MTR_ERR_SHAPE, "checkpoint %s was taken with other projection geometry", path);
    if (h.steps_done < 0) return fail(MTR_ERR_SHAPE, "checkpoint %s: bad step count", path);
    // a state rounded to another storage type would not continue bit-exactly
    if (h.precision != ctx->precision)
        return fail(MTR_ERR_SHAPE, "checkpoint %s was written in storage precision %d, the context uses %d", path,
                    (int)h.precision, ctx->precision);
    // the whole payload must be there before any state is replaced
    const long long need = (long long)sizeof h + (long long)(h.K + 1) * ctx->N * (long long)sizeof(float);
    if (std::fseek(in.f, 0, SEEK_END) != 0 || std::ftell(in.f) != need)
        return fail(MTR_ERR_SHAPE, "checkpoint %s is truncated or too long", path);
    // pass 1 checks every value, pass 2 uploads
    std::vector<float> buf((size_t)ctx->N);
    for (int pass = 0; pass < 2; pass++) {
        if (std::fseek(in.f, sizeof h, SEEK_SET) != 0) return fail(MTR_ERR_ARG, "read failed: %s", path);
        for (int v = 0; v <= h.K; v++) {
            if (std::fread(buf.data(), sizeof(float), buf.size(), in.f) != buf.size())
                return fail(MTR_ERR_ARG, "read failed: %s", path);
            if (pass == 0 && first_nonfinite(buf.data(), ctx->N) >= 0)
                return fail(MTR_ERR_NONFINITE, "checkpoint %s: NaN or Inf in volume %d", path, v);
            if (pass == 1) MTR_TRY(copy_volume_in(ctx, v < h.K ? ctx->agents[v].d_w : ctx->d_wbar, buf.data()));
        }
    }
    if (steps_done) *steps_done = h.steps_done;
    return MTR_OK;
}

int mtr_wbar_device(mtr_ctx *ctx, void **dev_ptr)
{
    if (!ctx || !dev_ptr) return fail(MTR_ERR_ARG, "null argument");
    if (!ctx->vol_set) return fail(MTR_ERR_STATE, "volume not set");
    *dev_ptr = ctx->d_wbar;
    return MTR_OK;
}

int mtr_residual_norm2(mtr_ctx *ctx, int agent, double *out)
{
    MTR_TRY(check_agent(ctx, agent));
    if (!out) return fail(MTR_ERR_ARG, "null out");
    MTR_TRY(set_device(ctx));
    Agent &ag = ctx->agents[agent];
    if (ag.npartial <= 0) return fail(MTR_ERR_STATE, "no SIRT iteration has run on agent %d", agent);
    HIP_TRY(launch_reduce(ag.d_partial, ag.npartial, ag.d_scalar, ctx->stream));
    HIP_TRY(hipMemcpyAsync(out, ag.d_scalar, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MTR_OK;
}

int mtr_consensus_diff2(mtr_ctx *ctx, double *out)
{
    if (!ctx || !out) return fail(MTR_ERR_ARG, "null argument");
    if (ctx->cons_blocks <= 0) return fail(MTR_ERR_STATE, "no consensus step has run");
    MTR_TRY(set_device(ctx));
    MTR_TRY(consensus_d2(ctx, out));
    return MTR_OK;
}

int mtr_op_fp(mtr_ctx *ctx, const mtr_projection_size *prj, const mtr_volume_size *vol, const double *xc,
              const double *yc, const float *x, float *s, float *c)
{
    if (!ctx || !prj || !vol || !xc || !yc || !x || !s || !c) return fail(MTR_ERR_ARG, "null argument");
    MTR_TRY(set_device(ctx));
    TmpAgent t;
    MTR_TRY(tmp_agent(ctx, prj, vol, xc, yc, t));
    const int64_t P = (int64_t)prj->X * prj->Y * prj->AngN, N = (int64_t)vol->X * vol->Y * vol->Z;
    HIP_TRY(hipMemcpyAsync(t.x.p, x, sizeof(float) * N, hipMemcpyHostToDevice, ctx->stream));
    {
        Timed tm(ctx, MTR_K_FP);
        HIP_TRY(launch_fp(MTR_PREC_FP32, (const AgentDev *)t.tab.p, 1, t.nfp, 1, t.lanes_k, (float *)t.s.p,
                          (float *)t.c.p, ctx->stream));
    }
    HIP_TRY(hipMemcpyAsync(s, t.s.p, sizeof(float) * P, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(c, t.c.p, sizeof(float) * P, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return collect_timing(ctx);
}

// ---- allocation-free forward projection --------------------------------
// mtr_op_fp keeps the agent record and the per-view tables (coefficients and
// ray plans) in buffers it allocates per call; mtr_op_fp_noalloc stages them
// in `scratch`, one fixed array per GPU held for the life of the process.
// Host code never uses the array's name as an address: the address of each
// device's copy comes from hipGetSymbolAddress, taken once per device with
// that device current (at context creation, or at the entry point's first
// call on the device) and kept in g_scratch_ptr.
namespace {
constexpr size_t round256(size_t n) { return (n + 255) / 256 * 256; }
constexpr size_t kNaCoefOff = round256(sizeof(AgentDev));
constexpr size_t kNaPlanOff = kNaCoefOff + round256(sizeof(ViewCoef) * MTR_NOALLOC_MAX_VIEWS);
constexpr size_t kNaBytes = kNaPlanOff + round256(sizeof(FpPlan) * MTR_NOALLOC_MAX_VIEWS);
constexpr int kNaMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kNaMaxDevices];

// Address of this device's copy of `scratch`; `device` must be current.
int scratch_init(int device)
{
    if (device < 0 || device >= kNaMaxDevices) return fail(MTR_ERR_RANGE, "device %d: no scratch slot", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return MTR_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(MTR_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<unsigned char *>(p);
    return MTR_OK;
}

unsigned char *scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kNaMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

// MTR_OK when q is device memory of `device`; clears the error state a
// failed query leaves, so the next launch does not report it.
int check_device_buffer(const void *q, int device, const char *what)
{
    hipPointerAttribute_t at{};
    const hipError_t e = hipPointerGetAttributes(&at, q);
    if (e != hipSuccess) (void)hipGetLastError();
    if (e != hipSuccess || at.type != hipMemoryTypeDevice || at.device != device)
        return fail(MTR_ERR_ARG, "%s is not device memory of device %d", what, device);
    return MTR_OK;
}
}  // namespace

int mtr_op_fp_noalloc(mtr_ctx *ctx, const mtr_projection_size *prj, const mtr_volume_size *vol, const double *xc,
                      const double *yc, const float *x_dev, float *s_dev, float *c_dev)
{
    if (!ctx || !prj || !vol || !xc || !yc || !x_dev || !s_dev || !c_dev) return fail(MTR_ERR_ARG, "null argument");
    if (!prj_valid(prj)) return fail(MTR_ERR_SHAPE, "invalid projection size");
    if (!vol_valid(vol)) return fail(MTR_ERR_SHAPE, "invalid volume size");
    const int A = prj->AngN;
    if (A > MTR_NOALLOC_MAX_VIEWS)
        return fail(MTR_ERR_SHAPE, "%d views: the allocation-free forward projection takes at most %d", A,
                    MTR_NOALLOC_MAX_VIEWS);
    MTR_TRY(set_device(ctx));
    unsigned char *base = scratch_ptr(ctx->device);
    if (!base) {
        MTR_TRY(scratch_init(ctx->device));
        base = scratch_ptr(ctx->device);
    }
    MTR_TRY(check_device_buffer(x_dev, ctx->device, "x"));
    MTR_TRY(check_device_buffer(s_dev, ctx->device, "s"));
    MTR_TRY(check_device_buffer(c_dev, ctx->device, "c"));
    // host staging on the stack (no heap): 61 KB at the view maximum
    ViewCoef coef[MTR_NOALLOC_MAX_VIEWS];
    FpPlan plan[MTR_NOALLOC_MAX_VIEWS];
    bool lanes_k = true;
    for (int v = 0; v < A; v++) {
        for (int m = 0; m < 4; m++) {
            coef[v].c[m] = xc[4 * v + m];
            coef[v].c[4 + m] = yc[4 * v + m];
        }
        const int rc = make_plan(xc + 4 * v, yc + 4 * v, plan[v]);
        plan[v].pfx_view = v;
        if (rc != MTR_OK) return fail(rc, "view %d: degenerate or non-finite geometry", v);
        lanes_k = lanes_k && plan[v].kind == 1 && plan[v].W == 0;
    }
    AgentDev d{};
    d.nu = prj->X;
    d.nv = prj->Y;
    d.A = A;
    d.Xs = vol->Xstart;
    d.Xe = vol->Xend;
    d.Ys = vol->Ystart;
    d.Ye = vol->Yend;
    d.Zs = vol->Zstart;
    d.Ze = vol->Zend;
    d.X = vol->X;
    d.Y = vol->Y;
    d.Z = vol->Z;
    d.coef = reinterpret_cast<const ViewCoef *>(base + kNaCoefOff);
    d.plan = reinterpret_cast<const FpPlan *>(base + kNaPlanOff);
    d.pk = prj->X;
    d.pi = 1;
    d.x = const_cast<float *>(x_dev);
    d.sx = 1;
    d.sy = vol->X;
    d.sz = (int64_t)vol->X * vol->Y;
    // every pixel of s and c is written by the launch (one thread per
    // pixel), so nothing in the scratch or the outputs needs clearing
    HIP_TRY(hipMemcpyAsync(base + kNaCoefOff, coef, sizeof(ViewCoef) * A, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(base + kNaPlanOff, plan, sizeof(FpPlan) * A, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(base, &d, sizeof d, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_fp(MTR_PREC_FP32, reinterpret_cast<const AgentDev *>(base), 1, fp_blocks(d, lanes_k), 1, lanes_k,
                      s_dev, c_dev, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MTR_OK;
}

int mtr_op_residual(mtr_ctx *ctx, int64_t n, const float *p, const float *s, const float *c, float *r)
{
    if (!ctx || !p || !s || !c || !r || n < 0) return fail(MTR_ERR_ARG, "bad argument");
    if (n == 0) return MTR_OK;
    MTR_TRY(set_device(ctx));
    DevBuf dp, ds, dc, dr;
    HIP_TRY(hipMalloc(&dp.p, sizeof(float) * n));
    HIP_TRY(hipMalloc(&ds.p, sizeof(float) * n));
    HIP_TRY(hipMalloc(&dc.p, sizeof(float) * n));
    HIP_TRY(hipMalloc(&dr.p, sizeof(float) * n));
    HIP_TRY(hipMemcpyAsync(dp.p, p, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ds.p, s, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dc.p, c, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_residual(MTR_PREC_FP32, n, dp.p, (const float *)ds.p, (const float *)dc.p, dr.p, ctx->stream));
    HIP_TRY(hipMemcpyAsync(r, dr.p, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MTR_OK;
}

int mtr_op_form_z(mtr_ctx *ctx, int64_t n, const float *w, const float *wbar, float *z)
{
    if (!ctx || !w || !wbar || !z || n < 0) return fail(MTR_ERR_ARG, "bad argument");
    if (n == 0) return MTR_OK;
    MTR_TRY(set_device(ctx));
    DevBuf dw, db, dz;
    HIP_TRY(hipMalloc(&dw.p, sizeof(float) * n));
    HIP_TRY(hipMalloc(&db.p, sizeof(float) * n));
    HIP_TRY(hipMalloc(&dz.p, sizeof(float) * n));
    HIP_TRY(hipMemcpyAsync(dw.p, w, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(db.p, wbar, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_form_z(MTR_PREC_FP32, n, dw.p, db.p, dz.p, ctx->stream));
    HIP_TRY(hipMemcpyAsync(z, dz.p, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MTR_OK;
}

int mtr_op_mann(mtr_ctx *ctx, int64_t n, const float *v, float *w, const float *z, float rho)
{
    if (!ctx || !v || !w || !z || n < 0) return fail(MTR_ERR_ARG, "bad argument");
    if (n == 0) return MTR_OK;
    MTR_TRY(set_device(ctx));
    DevBuf dv, dw, dz;
    HIP_TRY(hipMalloc(&dv.p, sizeof(float) * n));
    HIP_TRY(hipMalloc(&dw.p, sizeof(float) * n));
    HIP_TRY(hipMalloc(&dz.p, sizeof(float) * n));
    HIP_TRY(hipMemcpyAsync(dv.p, v, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dw.p, w, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dz.p, z, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_mann(MTR_PREC_FP32, n, dv.p, dw.p, dz.p, rho, ctx->stream));
    HIP_TRY(hipMemcpyAsync(w, dw.p, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MTR_OK;
}

int mtr_op_mean(mtr_ctx *ctx, int64_t n, int K, const float *const *w, float *wbar)
{
    if (!ctx || !w || !wbar || n < 0 || K <= 0 || K > kMaxAgents) return fail(MTR_ERR_ARG, "bad argument");
    if (n == 0) return MTR_OK;
    MTR_TRY(set_device(ctx));
    DevBuf dw[kMaxAgents], db;
    std::vector<const void *> ptrs;
    for (int a = 0; a < K; a++) {
        if (!w[a]) return fail(MTR_ERR_ARG, "null agent volume %d", a);
        HIP_TRY(hipMalloc(&dw[a].p, sizeof(float) * n));
        HIP_TRY(hipMemcpyAsync(dw[a].p, w[a], sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
        ptrs.push_back(dw[a].p);
    }
    HIP_TRY(hipMalloc(&db.p, sizeof(float) * n));
    HIP_TRY(launch_mean(MTR_PREC_FP32, K, n, ptrs.data(), db.p, ctx->stream));
    HIP_TRY(hipMemcpyAsync(wbar, db.p, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MTR_OK;
}

static int op_bp_common(mtr_ctx *ctx, const mtr_projection_size *prj, const mtr_volume_size *vol,
                        const double *xc, const double *yc, const float *r, float *x, float lambda, bool bpt)
{
    if (!ctx || !prj || !vol || !xc || !yc || !r || !x) return fail(MTR_ERR_ARG, "null argument");
    MTR_TRY(set_device(ctx));
    TmpAgent t;
    MTR_TRY(tmp_agent(ctx, prj, vol, xc, yc, t));
    const int64_t P = (int64_t)prj->X * prj->Y * prj->AngN, N = (int64_t)vol->X * vol->Y * vol->Z;
    HIP_TRY(hipMemcpyAsync(t.r.p, r, sizeof(float) * P, hipMemcpyHostToDevice, ctx->stream));
    if (!bpt) HIP_TRY(hipMemcpyAsync(t.x.p, x, sizeof(float) * N, hipMemcpyHostToDevice, ctx->stream));
    {
        Timed tm(ctx, bpt ? MTR_K_BPT : MTR_K_BP);
        if (t.bpm.sep)
            HIP_TRY(launch_bp_sep(MTR_PREC_FP32, (const AgentDev *)t.tab.p, 1,
                                  bp_sep_blocks(t.dev, t.bpm.W, t.bpm.P1, t.bpm.P2, t.bpm.rw), bpt, lambda, t.bpm.W,
                                  t.bpm.P1, t.bpm.P2, t.bpm.rw, ctx->stream));
        else
            HIP_TRY(launch_bp(MTR_PREC_FP32, (const AgentDev *)t.tab.p, 1, bp_blocks(t.dev, ctx->bp_tz), ctx->bp_tz,
                              bpt, lambda, t.r.p, ctx->lds_cap, ctx->stream));
    }
    HIP_TRY(hipMemcpyAsync(x, t.x.p, sizeof(float) * N, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return collect_timing(ctx);
}

int mtr_op_bp(mtr_ctx *ctx, const mtr_projection_size *prj, const mtr_volume_size *vol, const double *xc,
              const double *yc, const float *r, float *x, float lambda)
{
    return op_bp_common(ctx, prj, vol, xc, yc, r, x, lambda, false);
}

int mtr_op_bpt(mtr_ctx *ctx, const mtr_projection_size *prj, const mtr_volume_size *vol, const double *xc,
               const double *yc, const float *p, float *x)
{
    return op_bp_common(ctx, prj, vol, xc, yc, p, x, 1.0f, true);
}

int mtr_divide_slice(int myid, int nparts, int Z, int Zstart, mtr_slice *out)
{
    if (!out || nparts <= 0 || Z < 0 || myid < 0 || myid >= nparts) return fail(MTR_ERR_ARG, "bad partition");
    // First Z % nparts parts carry one extra slice (MultiAxis.cpp:13-31).
    const int rem = Z % nparts, base = Z / nparts;
    const int h = base + (myid < rem ? 1 : 0);
    out->z_start = Zstart + myid * base + std::min(myid, rem);
    out->z_end = out->z_start + h;
    return MTR_OK;
}

int mtr_timing_enable(mtr_ctx *ctx, int on)
{
    if (!ctx) return fail(MTR_ERR_ARG, "null context");
    MTR_TRY(collect_timing(ctx));
    ctx->timing = on != 0;
    return MTR_OK;
}

int mtr_timing_reset(mtr_ctx *ctx)
{
    if (!ctx) return fail(MTR_ERR_ARG, "null context");
    MTR_TRY(collect_timing(ctx));
    for (int k = 0; k < MTR_K_COUNT; k++) {
        ctx->launches[k] = 0;
        ctx->ms[k] = 0;
    }
    return MTR_OK;
}

int mtr_timing_get(mtr_ctx *ctx, int kernel, int64_t *launches, double *ms)
{
    if (!ctx || !launches || !ms) return fail(MTR_ERR_ARG, "null argument");
    if (kernel < 0 || kernel >= MTR_K_COUNT) return fail(MTR_ERR_RANGE, "unknown kernel class %d", kernel);
    MTR_TRY(collect_timing(ctx));
    *launches = ctx->launches[kernel];
    *ms = ctx->ms[kernel];
    return MTR_OK;
}

int mtr_calibrate_copy(mtr_ctx *ctx, size_t bytes, double *ms)
{
    if (!ctx || !ms || bytes < 4) return fail(MTR_ERR_ARG, "bad argument");
    MTR_TRY(set_device(ctx));
    DevBuf a, b;
    const int64_t n = (int64_t)(bytes / 4);
    HIP_TRY(hipMalloc(&a.p, sizeof(float) * n));
    HIP_TRY(hipMalloc(&b.p, sizeof(float) * n));
    HIP_TRY(launch_fill(MTR_PREC_FP32, n, a.p, 1.0f, ctx->stream));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, ctx->stream));
    HIP_TRY(launch_to_storage(MTR_PREC_FP32, n, (const float *)a.p, b.p, ctx->stream));
    HIP_TRY(hipEventRecord(e1, ctx->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t;
    return MTR_OK;
}

int mtr_sync(mtr_ctx *ctx)
{
    if (!ctx) return fail(MTR_ERR_ARG, "null context");
    MTR_TRY(set_device(ctx));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return collect_timing(ctx);
}

}  // extern "C"

/* ---- request server ------------------------------------------------- */

namespace {

enum { REQ_PENDING = 0, REQ_ACTIVE = 1, REQ_DONE = 2, REQ_FAILED = 3, REQ_RETIRING = 4 };
// host-side preparation of a pending request (its context, tables and
// uploads) by the server's worker thread
enum { PREP_NONE = 0, PREP_QUEUED = 1, PREP_RUNNING = 2, PREP_DONE = 3, PREP_FAILED = 4 };

struct ServerReq {
    int64_t id = 0;
    int state = REQ_PENDING;
    mtr_volume_size vol{};
    std::vector<mtr_projection_size> prj;
    std::vector<std::vector<double>> xc, yc;
    std::vector<std::vector<float>> p;  // host copies until admission
    int M = 0, S = 0;
    float lambda = 0.f, rho = 0.f;
    int step = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    size_t bytes = 0;
    double work = 0;  // voxel-view updates of the whole solve: (2SM + 1) N sum_a A_a
    mtr_ctx *ctx = nullptr;
    // From REQ_RETIRING on, state, result and t_done are written under
    // mtr_server::mu (by the retirer).
    std::unique_ptr<float[]> result;  // w-bar of a finished request, result_n values
    size_t result_n = 0;
    std::string error;
    int code = MTR_ERR_HIP;  // status of a failed request
    // PREP_*: written under mtr_server::mu (the tick thread sets QUEUED and
    // NONE, the worker RUNNING, DONE and FAILED)
    int prep = PREP_NONE;
    int prep_rc = MTR_OK;
    std::string prep_err;
};

}  // namespace

constexpr int kPrepWorkers = 4;  // request preparation threads of a server

struct mtr_server {
    int device = 0;
    int precision = MTR_PREC_FP32;
    size_t budget = 0;
    int max_active = 0;
    int admit_policy = MTR_ADMIT_ARRIVAL;
    double max_wait = 0;  // seconds, MTR_ADMIT_SMALLEST only
    mtr_ctx *eng = nullptr;  // engine context: the stream every request shares
    size_t used = 0;  // under mu
    int64_t next_id = 1;
    int n_done = 0;
    std::deque<ServerReq *> pending;
    std::vector<ServerReq *> active;
    std::map<int64_t, std::unique_ptr<ServerReq>> reqs;
    AgentDev *d_tab = nullptr;
    int tab_cap = 0;
    std::vector<AgentDev> h_tab;
    // A tick's launches are in flight until the next step synchronizes:
    // the host's work between ticks (submissions, admission) overlaps the
    // GPU's.  The requests that finished with it go to a second thread, the
    // retirer, which copies their w-bar to the host (on out_stream, through
    // pinned staging: a copy into pageable memory runs at a fraction of the
    // link's rate beside a tick's kernels) and frees their device memory
    // while the next ticks run.
    bool inflight = false;
    hipStream_t out_stream = nullptr;
    hipMemPool_t pool = nullptr;  // the requests' device memory
    float *h_stage = nullptr;  // the retirer's
    size_t stage_cap = 0;
    std::thread retirer;
    std::deque<ServerReq *> ret_q;
    int retiring = 0;  // handed to the retirer and not yet retired (under mu)
    std::condition_variable cv_ret, cv_retired;
    bool ret_quit = false;
    // Background preparation: worker threads build the contexts of the
    // requests the admission policy takes next (allocations, geometry
    // tables, projection uploads, each worker on a stream of its own) while
    // the GPU runs ticks; admission then only launches their normalisers
    // and initial models.
    bool prefetch = true;
    std::vector<hipStream_t> prep_streams;
    std::vector<std::thread> workers;
    mutable std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServerReq *> prep_q;
    bool quit = false;
    std::chrono::steady_clock::time_point t0;
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

size_t request_bytes(int precision, const mtr_volume_size &vol, int K, const mtr_projection_size *prj)
{
    const size_t es = elem_size(precision);
    const size_t N = (size_t)vol.X * vol.Y * vol.Z;
    size_t b = es * N + 3 * sizeof(AgentDev) * kMaxAgents + 8 * 8193;  // w-bar, tables, consensus partials
    for (int a = 0; a < K; a++) b += agent_bytes(precision, vol, prj[a]);
    return b;
}

void release(mtr_server *srv, ServerReq *r)
{
    if (r->ctx) {
        mtr_ctx_destroy(r->ctx);
        r->ctx = nullptr;
    }
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->used -= std::min(srv->used, r->bytes);
}

// Host side of admission: the request's context on stream st, its
// agents' geometry tables, allocations and projection uploads; nothing is
// launched on the engine's stream unless st is that stream.
int prepare(mtr_server *srv, ServerReq *r, hipStream_t st)
{
    mtr_ctx *c = new mtr_ctx();
    c->device = srv->device;
    c->stream = st;
    c->own_stream = false;
    c->pool = srv->pool;
    c->precision = srv->precision;
    c->joint_tune = false;  // its agents share launches with other requests' agents
    c->tune_timed = false;  // no timed tuning inside a tick: cached tiles or the heuristic ones
    r->ctx = c;
    MTR_TRY(set_device(c));
    MTR_TRY(alloc_tables(c));
    MTR_TRY(mtr_ctx_set_volume(c, &r->vol));
    for (size_t a = 0; a < r->prj.size(); a++) {
        int id = 0;
        MTR_TRY(agent_add_common(c, &r->prj[a], r->xc[a].data(), r->yc[a].data(), r->p[a].data(), nullptr, &id,
                                 true));
        std::vector<float>().swap(r->p[a]);  // resident now
    }
    HIP_TRY(hipStreamSynchronize(st));
    return MTR_OK;
}

// GPU side of admission, on the engine's stream: normalisers and the
// BPT-sum initial model (MultiAxis.cpp:163-200).
int activate(mtr_server *srv, ServerReq *r)
{
    mtr_ctx *c = r->ctx;
    c->stream = srv->eng->stream;
    for (int a = 0; a < (int)c->agents.size(); a++) MTR_TRY(ensure_normaliser(c, a));
    MTR_TRY(mtr_mace_init(c));
    return MTR_OK;
}

void prep_worker(mtr_server *srv, hipStream_t st)
{
    (void)hipSetDevice(srv->device);
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        srv->cv_work.wait(lk, [&] { return srv->quit || !srv->prep_q.empty(); });
        if (srv->quit) return;
        ServerReq *r = srv->prep_q.front();
        srv->prep_q.pop_front();
        r->prep = PREP_RUNNING;
        lk.unlock();
        const int rc = prepare(srv, r, st);
        std::string err = rc != MTR_OK ? g_err : std::string();
        lk.lock();
        r->prep_rc = rc;
        r->prep_err = std::move(err);
        r->prep = rc == MTR_OK ? PREP_DONE : PREP_FAILED;
        srv->cv_done.notify_all();
    }
}

// The pending requests in the order the admission loop would take them
// now: the oldest, or (smallest first) the least work unless the oldest
// has waited longer than max_wait.
std::vector<ServerReq *> policy_order(const mtr_server *srv)
{
    std::vector<ServerReq *> left(srv->pending.begin(), srv->pending.end()), out;
    const double now = srv->now();
    while (!left.empty()) {
        size_t pick = 0;
        if (srv->admit_policy == MTR_ADMIT_SMALLEST && now - left[0]->t_submit <= srv->max_wait)
            for (size_t q = 1; q < left.size(); q++)
                if (left[q]->work < left[pick]->work) pick = q;
        out.push_back(left[pick]);
        left.erase(left.begin() + (long)pick);
    }
    return out;
}

// Queues for preparation the requests the policy takes next, for the slots
// the active ones leave free and one more per worker, while the HBM budget
// allows (their bytes count as in use from here on).
void plan_prep(mtr_server *srv)
{
    if (!srv->prefetch) return;
    const std::vector<ServerReq *> order = policy_order(srv);
    const int want = std::max(srv->max_active - (int)srv->active.size(), 0) + (int)srv->workers.size();
    std::lock_guard<std::mutex> lk(srv->mu);
    for (int q = 0; q < (int)order.size() && q < want; q++) {
        ServerReq *r = order[q];
        if (r->prep != PREP_NONE) continue;
        if (srv->used + r->bytes > srv->budget) break;
        srv->used += r->bytes;
        r->prep = PREP_QUEUED;
        srv->prep_q.push_back(r);
    }
    if (!srv->prep_q.empty()) srv->cv_work.notify_all();
}

// The admitted request's prepared context: waits for the worker if it is
// preparing it, else prepares it here on the engine's stream.
int take_prepared(mtr_server *srv, ServerReq *r)
{
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        if (r->prep == PREP_QUEUED) {
            srv->prep_q.erase(std::find(srv->prep_q.begin(), srv->prep_q.end(), r));
            r->prep = PREP_NONE;
        } else if (r->prep == PREP_RUNNING) {
            srv->cv_done.wait(lk, [&] { return r->prep == PREP_DONE || r->prep == PREP_FAILED; });
        }
        if (r->prep == PREP_FAILED) return fail(r->prep_rc, "%s", r->prep_err.c_str());
        if (r->prep == PREP_DONE) return MTR_OK;
    }
    return prepare(srv, r, srv->eng->stream);
}

// Hands a finished request's w-bar to the host and frees its device
// memory (on the retirer, or on the tick thread when there is none).
void retire_one(mtr_server *srv, ServerReq *r)
{
    const size_t n = (size_t)r->ctx->N;
    int rc = MTR_OK;
    if (n > srv->stage_cap) {
        if (srv->h_stage) (void)hipHostFree(srv->h_stage);
        srv->h_stage = nullptr;
        srv->stage_cap = 0;
        if (hipHostMalloc((void **)&srv->h_stage, sizeof(float) * n, hipHostMallocDefault) == hipSuccess)
            srv->stage_cap = n;
        else
            rc = fail(MTR_ERR_OOM, "pinned staging of %zu floats", n);
    }
    if (rc == MTR_OK) {
        r->ctx->stream = srv->out_stream;  // its destruction then waits for this copy only
        rc = copy_volume_out(r->ctx, r->ctx->d_wbar, srv->h_stage);
    }
    std::unique_ptr<float[]> res;
    if (rc == MTR_OK) {
        res.reset(new (std::nothrow) float[n]);
        if (res)
            std::memcpy(res.get(), srv->h_stage, sizeof(float) * n);
        else
            rc = fail(MTR_ERR_OOM, "host result of %zu floats", n);
    }
    const std::string err = rc != MTR_OK ? g_err : std::string();
    // done once the w-bar is on the host: before the memory it frees can
    // admit another request
    const double t_done = srv->now();
    release(srv, r);
    std::lock_guard<std::mutex> lk(srv->mu);
    r->result = std::move(res);
    r->result_n = n;
    r->t_done = t_done;
    if (rc != MTR_OK) {
        r->error = err;
        r->code = rc;
        r->state = REQ_FAILED;
    } else {
        r->state = REQ_DONE;
    }
    if (srv->retiring > 0) srv->retiring--;
    srv->cv_retired.notify_all();
}

void retire_worker(mtr_server *srv)
{
    (void)hipSetDevice(srv->device);
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        srv->cv_ret.wait(lk, [&] { return srv->ret_quit || !srv->ret_q.empty(); });
        if (srv->ret_quit) return;  // the server releases what is left
        ServerReq *r = srv->ret_q.front();
        srv->ret_q.pop_front();
        lk.unlock();
        retire_one(srv, r);
        lk.lock();
    }
}

}  // namespace

extern "C" {

int mtr_server_create(int device, int precision, size_t hbm_budget, int max_active, mtr_server **out)
{
    if (!out) return fail(MTR_ERR_ARG, "null out");
    *out = nullptr;
    if (precision != MTR_PREC_FP32 && precision != MTR_PREC_BF16 && precision != MTR_PREC_FP16)
        return fail(MTR_ERR_ARG, "unknown precision %d", precision);
    if (max_active <= 0 || hbm_budget == 0) return fail(MTR_ERR_ARG, "empty server");
    mtr_ctx *eng = nullptr;
    MTR_TRY(mtr_ctx_create(device, &eng));
    eng->precision = precision;
    mtr_server *srv = new mtr_server();
    srv->device = device;
    srv->precision = precision;
    srv->budget = hbm_budget;
    srv->max_active = max_active;
    srv->eng = eng;
    srv->t0 = std::chrono::steady_clock::now();
    if (hipStreamCreateWithFlags(&srv->out_stream, hipStreamNonBlocking) != hipSuccess) {
        mtr_ctx_destroy(eng);
        delete srv;
        return fail(MTR_ERR_HIP, "server result stream");
    }
    {
        // freed request memory stays in the pool (up to a quarter of the
        // budget) for the next requests rather than going back to the
        // driver at each synchronization; without a pool (a device that
        // has none) requests use hipMalloc and hipFree
        hipMemPoolProps pp = {};
        pp.allocType = hipMemAllocationTypePinned;
        pp.location.type = hipMemLocationTypeDevice;
        pp.location.id = device;
        uint64_t keep = hbm_budget / 4;
        if (hipMemPoolCreate(&srv->pool, &pp) != hipSuccess) {
            srv->pool = nullptr;
            (void)hipGetLastError();
        } else if (hipMemPoolSetAttribute(srv->pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) {
            (void)hipMemPoolDestroy(srv->pool);
            srv->pool = nullptr;
            (void)hipGetLastError();
        }
    }
    try {
        srv->retirer = std::thread(retire_worker, srv);
    } catch (...) {  // no thread to spare: retire inside the steps
    }
    if (const char *pf = std::getenv("MTR_SERVER_PREFETCH")) srv->prefetch = std::atoi(pf) != 0;
    // as many workers as threads can be had, up to kPrepWorkers; with none,
    // requests are prepared inside the ticks
    for (int w = 0; srv->prefetch && w < kPrepWorkers; w++) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) break;
        try {
            srv->workers.emplace_back(prep_worker, srv, st);
        } catch (...) {
            (void)hipStreamDestroy(st);
            break;
        }
        srv->prep_streams.push_back(st);
    }
    if (srv->workers.empty()) srv->prefetch = false;
    *out = srv;
    return MTR_OK;
}

int mtr_server_destroy(mtr_server *srv)
{
    if (!srv) return MTR_OK;
    (void)hipSetDevice(srv->device);
    if (!srv->workers.empty()) {
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->quit = true;
        }
        srv->cv_work.notify_all();
        for (auto &w : srv->workers) w.join();
    }
    if (srv->retirer.joinable()) {
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->ret_quit = true;
        }
        srv->cv_ret.notify_all();
        srv->retirer.join();
    }
    (void)hipStreamSynchronize(srv->eng->stream);
    for (auto &kv : srv->reqs) release(srv, kv.second.get());
    // the frees just queued on the server's streams, then the pool
    (void)hipStreamSynchronize(srv->eng->stream);
    (void)hipStreamSynchronize(srv->out_stream);
    for (hipStream_t st : srv->prep_streams) (void)hipStreamSynchronize(st);
    if (srv->pool) (void)hipMemPoolDestroy(srv->pool);
    for (hipStream_t st : srv->prep_streams) (void)hipStreamDestroy(st);
    if (srv->out_stream) (void)hipStreamDestroy(srv->out_stream);
    if (srv->h_stage) (void)hipHostFree(srv->h_stage);
    if (srv->d_tab) (void)hipFree(srv->d_tab);
    mtr_ctx_destroy(srv->eng);
    delete srv;
    return MTR_OK;
}

int mtr_server_footprint(const mtr_server *srv, const mtr_volume_size *vol, int K, const mtr_projection_size *prj,
                         size_t *bytes)
{
    if (!srv || !vol || !prj || !bytes) return fail(MTR_ERR_ARG, "null argument");
    if (!vol_valid(vol)) return fail(MTR_ERR_SHAPE, "invalid volume box");
    if (K <= 0 || K > kMaxAgents) return fail(MTR_ERR_RANGE, "agent count %d outside 1..%d", K, kMaxAgents);
    for (int a = 0; a < K; a++)
        if (!prj_valid(&prj[a])) return fail(MTR_ERR_SHAPE, "invalid projection size of agent %d", a);
    *bytes = request_bytes(srv->precision, *vol, K, prj);
    return MTR_OK;
}

int mtr_server_submit(mtr_server *srv, const mtr_volume_size *vol, int K, const mtr_projection_size *prj,
                      const double *const *xc, const double *const *yc, const float *const *p, int M, int S,
                      float lambda, float rho, int64_t *request_id)
{
    if (!srv || !xc || !yc || !p) return fail(MTR_ERR_ARG, "null argument");
    size_t bytes = 0;
    MTR_TRY(mtr_server_footprint(srv, vol, K, prj, &bytes));
    if (M < 0 || S < 0) return fail(MTR_ERR_ARG, "negative iteration count");
    if (bytes > srv->budget)
        return fail(MTR_ERR_OOM, "request needs %zu bytes, the server's budget is %zu", bytes, srv->budget);
    auto r = std::make_unique<ServerReq>();
    r->id = srv->next_id++;
    r->vol = *vol;
    r->prj.assign(prj, prj + K);
    for (int a = 0; a < K; a++) {
        if (!xc[a] || !yc[a] || !p[a]) return fail(MTR_ERR_ARG, "null series %d", a);
        const size_t A = (size_t)prj[a].AngN, P = (size_t)prj[a].X * prj[a].Y * prj[a].AngN;
        // rejected here, not at admission: a queued request always runs
        for (size_t i = 0; i < 4 * A; i++)
            if (!std::isfinite(xc[a][i]) || !std::isfinite(yc[a][i]))
                return fail(MTR_ERR_NONFINITE, "series %d: geometry coefficient %zu is not finite", a, i);
        const int64_t bad = first_nonfinite(p[a], (int64_t)P);
        if (bad >= 0)
            return fail(MTR_ERR_NONFINITE, "series %d: projection value %lld is not finite", a, (long long)bad);
        r->xc.emplace_back(xc[a], xc[a] + 4 * A);
        r->yc.emplace_back(yc[a], yc[a] + 4 * A);
        r->p.emplace_back(p[a], p[a] + P);
    }
    r->M = M;
    r->S = S;
    r->lambda = lambda;
    r->rho = rho;
    r->bytes = bytes;
    {
        double views = 0;
        for (int a = 0; a < K; a++) views += prj[a].AngN;
        r->work = (2.0 * S * M + 1.0) * (double)vol->X * vol->Y * vol->Z * views;
    }
    r->t_submit = srv->now();
    if (request_id) *request_id = r->id;
    srv->pending.push_back(r.get());
    srv->reqs[r->id] = std::move(r);
    plan_prep(srv);
    return MTR_OK;
}

int mtr_server_set_admission(mtr_server *srv, int policy, double max_wait_s)
{
    if (!srv) return fail(MTR_ERR_ARG, "null server");
    if (policy != MTR_ADMIT_ARRIVAL && policy != MTR_ADMIT_SMALLEST)
        return fail(MTR_ERR_ARG, "unknown admission policy %d", policy);
    if (!(max_wait_s >= 0) || !std::isfinite(max_wait_s))
        return fail(MTR_ERR_ARG, "maximum wait must be finite and non-negative");
    srv->admit_policy = policy;
    srv->max_wait = max_wait_s;
    return MTR_OK;
}

namespace {

// Completes the tick in flight: waits for its launches; the requests that
// reached their M steps leave the active set (their slots are free for
// this step's admissions) and count as done; the retirer takes them over.
int finish_tick(mtr_server *srv, int *done)
{
    if (!srv->inflight) return MTR_OK;
    HIP_TRY(hipStreamSynchronize(srv->eng->stream));
    srv->inflight = false;
    std::vector<ServerReq *> still, fin;
    for (ServerReq *r : srv->active) (r->step < r->M ? still : fin).push_back(r);
    srv->active.swap(still);
    *done += (int)fin.size();
    srv->n_done += (int)fin.size();
    if (srv->retirer.joinable()) {
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            for (ServerReq *r : fin) {
                r->state = REQ_RETIRING;
                srv->ret_q.push_back(r);
            }
            srv->retiring += (int)fin.size();
        }
        if (!fin.empty()) srv->cv_ret.notify_one();
    } else {
        for (ServerReq *r : fin) retire_one(srv, r);
    }
    return collect_timing(srv->eng);
}

}  // namespace

// One scheduling tick: complete the previous one (its finished requests go
// to the retirer), admit pending requests while the HBM budget and the
// slot count allow (in arrival order, or smallest solve first with a bound
// on the wait), and launch one MACE step of every active request with their
// SIRT iterations batched into shared launches.
int mtr_server_step(mtr_server *srv, int *completed)
{
    if (!srv) return fail(MTR_ERR_ARG, "null server");
    if (completed) *completed = 0;
    MTR_TRY(set_device(srv->eng));
    int done = 0;
    MTR_TRY(finish_tick(srv, &done));
    while (!srv->pending.empty() && (int)srv->active.size() < srv->max_active) {
        // arrival order: the oldest request; smallest first: the oldest one
        // that has waited longer than max_wait, else the least work (the
        // earlier arrival on a tie).  The pick waits for HBM rather than be
        // overtaken, so a large request is never starved.
        auto pick = srv->pending.begin();
        if (srv->admit_policy == MTR_ADMIT_SMALLEST && srv->now() - (*pick)->t_submit <= srv->max_wait)
            for (auto it = srv->pending.begin(); it != srv->pending.end(); ++it)
                if ((*it)->work < (*pick)->work) pick = it;
        ServerReq *r = *pick;
        // a request queued for preparation already counts in `used`
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            const size_t need = r->prep == PREP_NONE ? r->bytes : 0;
            // with nothing else to launch, wait for memory the retirer is
            // about to free rather than return an empty tick
            if (srv->active.empty())
                srv->cv_retired.wait(lk, [&] { return srv->used + need <= srv->budget || srv->retiring == 0; });
            if (srv->used + need > srv->budget) break;
            srv->used += need;
        }
        srv->pending.erase(pick);
        r->t_admit = srv->now();
        int rc = take_prepared(srv, r);
        if (rc == MTR_OK) rc = activate(srv, r);
        if (rc != MTR_OK) {
            // finished, with its error for mtr_server_result
            r->error = g_err;
            r->code = rc;
            r->state = REQ_FAILED;
            r->t_done = srv->now();
            release(srv, r);
            done++;
            srv->n_done++;
            continue;
        }
        r->state = REQ_ACTIVE;
        srv->active.push_back(r);
    }
    // batch: every agent of every active request with steps left, the
    // column-plan agents after the others (the two lanes of the tick, on two
    // streams), each lane most SIRT iterations first, then by relaxation and
    // kernel
    Batch B;
    std::vector<const ServerReq *> owner;
    for (ServerReq *r : srv->active)
        if (r->step < r->M)
            for (auto &ag : r->ctx->agents) {
                B.e.push_back({r->ctx, &ag});
                owner.push_back(r);
            }
    std::vector<int> order(B.e.size());
    for (size_t q = 0; q < order.size(); q++) order[q] = (int)q;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (B.e[x].ag->col.on != B.e[y].ag->col.on) return B.e[y].ag->col.on;
        if (owner[x]->S != owner[y]->S) return owner[x]->S > owner[y]->S;
        if (owner[x]->lambda != owner[y]->lambda) return owner[x]->lambda < owner[y]->lambda;
        const int fx = fp_kind_key(*B.e[x].ag), fy = fp_kind_key(*B.e[y].ag);
        if (fx != fy) return fx < fy;
        return bp_kind_key(*B.e[x].ag) < bp_kind_key(*B.e[y].ag);
    });
    {
        Batch sorted;
        std::vector<const ServerReq *> own2;
        for (int q : order) {
            sorted.e.push_back(B.e[q]);
            own2.push_back(owner[q]);
        }
        B = std::move(sorted);
        owner = std::move(own2);
    }
    const int n = (int)B.e.size();
    // from here on this tick's work is on the stream (admissions above
    // already launched their initial models)
    srv->inflight = true;
    if (n > 0) {
        if (n > srv->tab_cap) {
            if (srv->d_tab) (void)hipFree(srv->d_tab);
            srv->d_tab = nullptr;
            srv->tab_cap = 0;
            HIP_TRY(hipMalloc(&srv->d_tab, sizeof(AgentDev) * n));
            srv->tab_cap = n;
        }
        srv->h_tab.resize(n);
        for (int q = 0; q < n; q++) srv->h_tab[q] = make_dev(B.e[q].owner, *B.e[q].ag);
        HIP_TRY(hipMemcpyAsync(srv->d_tab, srv->h_tab.data(), sizeof(AgentDev) * n, hipMemcpyHostToDevice,
                               srv->eng->stream));
        B.d_tab = srv->d_tab;
        // lanes [0, nb) and [nb, n): the others on the engine's second
        // stream, the column-plan agents on its own (timing keeps one)
        int nb = 0;
        while (nb < n && !B.e[nb].ag->col.on) nb++;
        mtr_ctx *eng = srv->eng;
        const bool two = eng->aux && eng->kind_streams && !eng->timing && nb > 0 && nb < n;
        if (two) {
            HIP_TRY(hipEventRecord(eng->ev_fork, eng->stream));
            HIP_TRY(hipStreamWaitEvent(eng->aux, eng->ev_fork, 0));
        }
        auto lane = [&](int l0, int l1, hipStream_t st) -> int {
            for (int it = 0; l1 > l0 && it < owner[l0]->S; it++) {
                int m = l0;
                while (m < l1 && owner[m]->S > it) m++;
                MTR_TRY(fp_batch(eng, B, l0, m - l0, st));
                int s0 = l0;
                while (s0 < m) {
                    int e0 = s0 + 1;
                    while (e0 < m && owner[e0]->lambda == owner[s0]->lambda) e0++;
                    MTR_TRY(bp_batch(eng, B, s0, e0 - s0, false, owner[s0]->lambda, st));
                    s0 = e0;
                }
            }
            return MTR_OK;
        };
        if (two) {
            MTR_TRY(lane(nb, n, eng->stream));
            MTR_TRY(lane(0, nb, eng->aux));
            HIP_TRY(hipEventRecord(eng->ev_join, eng->aux));
            HIP_TRY(hipStreamWaitEvent(eng->stream, eng->ev_join, 0));
        } else {
            MTR_TRY(lane(0, nb, eng->stream));
            MTR_TRY(lane(nb, n, eng->stream));
        }
    }
    // consensus of every active request (MultiAxis.cpp:40-71)
    for (ServerReq *r : srv->active) {
        if (r->step >= r->M) continue;
        mtr_ctx *c = r->ctx;
        std::vector<void *> w, x;
        for (auto &ag : c->agents) {
            MTR_TRY(to_canon(c, ag));
            w.push_back(ag.d_w);
            x.push_back(work_canon(ag));
        }
        {
            Timed t(srv->eng, MTR_K_CONSENSUS);
            HIP_TRY(launch_consensus(c->precision, (int)c->agents.size(), c->N, r->rho, 1, w.data(), x.data(),
                                     c->d_wbar, c->d_cons_partial, consensus_blocks(c->N), c->stream));
        }
        for (auto &ag : c->agents) MTR_TRY(from_canon(c, ag));
        r->step++;
    }
    // the next admissions' host work overlaps this tick (and so does the
    // caller's, until the next step completes it)
    plan_prep(srv);
    if (completed) *completed = done;
    return MTR_OK;
}

int mtr_server_status(const mtr_server *srv, int *pending, int *active, int *done, size_t *hbm_in_use)
{
    if (!srv) return fail(MTR_ERR_ARG, "null server");
    if (pending) *pending = (int)srv->pending.size();
    if (active) *active = (int)srv->active.size();
    if (done) *done = srv->n_done;
    if (hbm_in_use) {
        std::lock_guard<std::mutex> lk(srv->mu);
        *hbm_in_use = srv->used;
    }
    return MTR_OK;
}

int mtr_server_result(mtr_server *srv, int64_t request_id, float *wbar_host, double *times)
{
    if (!srv) return fail(MTR_ERR_ARG, "null server");
    auto it = srv->reqs.find(request_id);
    if (it == srv->reqs.end()) return fail(MTR_ERR_RANGE, "unknown request %lld", (long long)request_id);
    ServerReq *r = it->second.get();
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        srv->cv_retired.wait(lk, [&] { return r->state != REQ_RETIRING; });
    }
    if (times) {
        times[0] = r->t_submit;
        times[1] = r->t_admit;
        times[2] = r->t_done;
    }
    if (r->state == REQ_FAILED) return fail(r->code, "request %lld failed: %s", (long long)request_id, r->error.c_str());
    if (r->state != REQ_DONE) return fail(MTR_ERR_STATE, "request %lld not finished", (long long)request_id);
    if (wbar_host) std::memcpy(wbar_host, r->result.get(), sizeof(float) * r->result_n);
    return MTR_OK;
}

int mtr_server_now(const mtr_server *srv, double *seconds)
{
    if (!srv || !seconds) return fail(MTR_ERR_ARG, "null argument");
    *seconds = srv->now();
    return MTR_OK;
}

int mtr_server_timing(mtr_server *srv, int on) { return srv ? mtr_timing_enable(srv->eng, on) : fail(MTR_ERR_ARG, "null server"); }

int mtr_server_timing_get(mtr_server *srv, int kernel, int64_t *launches, double *ms)
{
    return srv ? mtr_timing_get(srv->eng, kernel, launches, ms) : fail(MTR_ERR_ARG, "null server");
}

}  // extern "C"
