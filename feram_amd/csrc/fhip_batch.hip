// fhip_batch.hip -- replica batches: many independent contexts (replicas of a
// temperature / field sweep) stepped together.
//
// The replicas of one size class (edges, precision, acoustic degrees of
// freedom) share every kernel instantiation and grid, so each stage of a step
// is one launch covering all of them (blockIdx.z = replica): a 12^3 replica
// alone fills a few dozen workgroups of the 256 CUs, twelve of them together
// fill a few hundred.  Each class runs on its own stream, forked from and
// joined to the batch stream, so the classes overlap on the device.  The
// steps between the first and the last of a call are identical launch
// sequences; they are captured once into a HIP graph and replayed.
#include "fhip_host.hpp"

#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

using namespace fhip;
using namespace fhip_host;

namespace {

bool batch_size_ok(int n)
{
    switch (n) {
#define FHIP_CASE(n) case n:
        FHIP_BATCH_SIZES(FHIP_CASE)
#undef FHIP_CASE
        return true;
    default:
        return false;
    }
}

struct BClass {
    int prec = FHIP_FP64;
    int Lx = 0, Ly = 0, Lz = 0, acoustic = 0, nf = 9;
    int zsp = 0;                        // species with different Z*: split dipole-dipole z group
    int xblocks = 0, xfblocks = 0, yblocks = 0, zblocks = 0, kblocks = 0;
    size_t zshm_g = 0;                  // dynamic LDS of the paired z pass
    int zsplit = 0;                     // the paired z pass runs a tile's groups in separate workgroups
    bool film = false;                  // a member mirrors its upper layers (film/epit)
    int mblocks = 0;                    // blocks of the mirrored-layer kick
    // tile-shape plan: rows per xfwd workgroup and lines per y tile, each the
    // full or the half tile, picked by timing (tune_class)
    int xrows = 0, ytile = 0;
    bool tuned = false;
    double tx[2] = {0, 0}, ty[2] = {0, 0};   // measured us per launch: full, half
    std::vector<fhip_ctx *> members;
    void *dreps = nullptr;              // Rep<T>[members.size()] on the device
    std::vector<unsigned char> hreps;   // host image of the same
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool avg = false, hist = false;
    bool heat = false;                  // a leapfrog member has heat-source layers
};

}  // namespace

struct fhip_batch {
    bool tune = true;
    std::vector<fhip_ctx *> ctxs;
    std::vector<BClass> classes;
    int method = FHIP_MD;
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr;
    bool use_graph = true;
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};     // 16, 4 and 1 middle steps
    hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
    std::vector<int> gkey;              // averaging flags and tiles of the captured graphs
    std::string err;
};

namespace {

int bfail(fhip_batch *b, int code, const std::string &msg)
{
    if (b) b->err = msg;
    return code;
}

#define BHIP(expr)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return bfail(b, FHIP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define BLAUNCH(what)                                                                    \
    do {                                                                                 \
        hipError_t e_ = hipGetLastError();                                               \
        if (e_ != hipSuccess)                                                            \
            return bfail(b, FHIP_ERR_HIP, std::string("launch ") + what + ": " + hipGetErrorString(e_)); \
        return FHIP_OK;                                                                  \
    } while (0)

int xf_mode(int method) { return method == FHIP_MD ? XF_NP : (method == FHIP_LF ? XF_LF : XF_VS); }
int xi_mode(int method) { return method == FHIP_MD ? XI_NP : (method == FHIP_LF ? XI_LF : XI_VS); }
int fin_mode(int method) { return method == FHIP_MD ? FIN_NP : (method == FHIP_LF ? FIN_LF : FIN_VS); }

// Host image of a replica's arguments, read from the context as it is now.
template <typename T>
void fill_rep(Rep<T> &r, fhip_ctx *ctx, int method)
{
    std::memset(&r, 0, sizeof(r));
    r.c = ctx->c;
    r.k = ctx->sk;
    r.f = fields<T>(ctx);
    r.sp = tset<T>(ctx).sp;
    r.za = zargs<T>(ctx, !ctx->c.acoustic);
    r.xa.mode = xf_mode(method);
    r.xa.nf = 9 + (ctx->c.acoustic ? 3 : 0) + (ctx->c.zsp ? 3 : 0);
    r.ia.mode = xi_mode(method);
    r.ia.materialize = 0;
    r.ia.partials = ctx->partials;
    r.ia.maxblocks = ctx->maxblocks;
    r.sc = ctx->dsc;
    r.partials = ctx->partials;
    r.twx = tset<T>(ctx).twx;
    r.twy = tset<T>(ctx).twy;
    r.twz = tset<T>(ctx).twz;
    r.acc = ctx->davg;
    r.dist1 = ctx->dist1;
    r.dist3 = ctx->dist3;
    r.maxblocks = ctx->maxblocks;
    r.xblocks = ctx->xblocks;
    r.zblocks = ctx->zblocks[tix<T>()];
    r.kblocks = ctx->kblocks;
    r.ravg = ctx->avg_ravg ? 1 : 0;
    if (method == FHIP_LF && (ctx->delta_e != 0.0 || ctx->delta_e_acou != 0.0)) {
        r.el = ctx->el;
        r.el_sum = ctx->el_sum;
        r.nlayer = ctx->nlayer;
        r.delta_e = ctx->delta_e;
        r.delta_e_acou = ctx->delta_e_acou;
    }
}

template <typename T>
int upload_reps(fhip_batch *b, BClass &k)
{
    const size_t n = k.members.size();
    k.hreps.resize(n * sizeof(Rep<T>));
    Rep<T> *h = reinterpret_cast<Rep<T> *>(k.hreps.data());
    for (size_t i = 0; i < n; ++i) {
        fill_rep<T>(h[i], k.members[i], b->method);
        h[i].za.split = k.zsplit;
    }
    if (!k.dreps) BHIP(hipMalloc(&k.dreps, k.hreps.size()));
    BHIP(hipMemcpy(k.dreps, k.hreps.data(), k.hreps.size(), hipMemcpyHostToDevice));
    return FHIP_OK;
}

// ---- tile-shape plan -----------------------------------------------------------
void set_plan(BClass &k, int xrows, int ytile)
{
    k.xrows = xrows;
    k.ytile = ytile;
    k.xfblocks = (k.Ly * k.Lz + xrows - 1) / xrows;
    k.yblocks = ((k.Lx / 2) * k.Lz + ytile - 1) / ytile;
}

// Plans found by timing, shared by every batch of the process: (Lx, Ly, Lz,
// precision, acoustic, species Z*, replicas) -> (xrows, ytile).
std::mutex g_plan_mu;
std::map<std::array<int, 7>, std::pair<int, int>> g_plans;

std::array<int, 7> plan_key(const BClass &k)
{
    return {k.Lx, k.Ly, k.Lz, k.prec, k.acoustic, k.zsp, (int)k.members.size()};
}

// ---- per-stage launches of one class -----------------------------------------
// Each case calls the launcher instantiated for that edge (fhip_stages.hip).
template <typename T>
int l_xfwd(fhip_batch *b, BClass &k, hipStream_t s, int phase, int mode = -1)
{
    const Rep<T> *r = (const Rep<T> *)k.dreps;
    const dim3 g(k.xfblocks, 1, (unsigned)k.members.size());
    const bool half = k.xrows != tile_rows(k.Lx);
    switch (k.Lx) {
#define FHIP_CASE(n) \
    case n: fhip_stage::b_xfwd<T, n>(r, g, s, half, phase, mode); BLAUNCH("xfwd");
        FHIP_BATCH_SIZES(FHIP_CASE)
#undef FHIP_CASE
    }
    return bfail(b, FHIP_ERR_UNSUPPORTED, "Lx not batched");
}

template <typename T>
int l_ypass(fhip_batch *b, BClass &k, hipStream_t s, int sign)
{
    const Rep<T> *r = (const Rep<T> *)k.dreps;
    const dim3 g(k.yblocks, k.nf, (unsigned)k.members.size());
    const bool half = k.ytile != line_tile_rows<T>(k.Ly);
    switch (k.Ly) {
#define FHIP_CASE(n) \
    case n: fhip_stage::b_ypass<T, n>(r, g, s, half, sign); BLAUNCH("ypass");
        FHIP_BATCH_SIZES(FHIP_CASE)
#undef FHIP_CASE
    }
    return bfail(b, FHIP_ERR_UNSUPPORTED, "Ly not batched");
}

template <typename T, int G2, int DDG>
const void *zpair_fn(int Lz)
{
    switch (Lz) {
#define FHIP_CASE(n) \
    case n: return fhip_stage::b_zpair_fn<T, n, G2, DDG>();
        FHIP_BATCH_SIZES(FHIP_CASE)
#undef FHIP_CASE
    }
    return nullptr;
}

template <typename T, int G2, int DDG>
int l_zpair(fhip_batch *b, BClass &k, hipStream_t s, int want_acouK)
{
    const Rep<T> *r = (const Rep<T> *)k.dreps;
    const dim3 g(k.zblocks * (k.zsplit ? 2 : 1), 1, (unsigned)k.members.size());
    switch (k.Lz) {
#define FHIP_CASE(n) \
    case n: fhip_stage::b_zpair<T, n, G2, DDG>(r, g, s, k.zshm_g, want_acouK); BLAUNCH("zpair");
        FHIP_BATCH_SIZES(FHIP_CASE)
#undef FHIP_CASE
    }
    return bfail(b, FHIP_ERR_UNSUPPORTED, "Lz not batched");
}

template <typename T>
int l_xinv(fhip_batch *b, BClass &k, hipStream_t s, int materialize)
{
    const Rep<T> *r = (const Rep<T> *)k.dreps;
    const dim3 g(k.xblocks, 1, (unsigned)k.members.size());
    switch (k.Lx) {
#define FHIP_CASE(n) \
    case n: fhip_stage::b_xinv<T, n>(r, g, s, k.nf, materialize); BLAUNCH("xinv");
        FHIP_BATCH_SIZES(FHIP_CASE)
#undef FHIP_CASE
    }
    return bfail(b, FHIP_ERR_UNSUPPORTED, "Lx not batched");
}

// One step of every replica of the class: the force pipeline of the
// single-context path (xfwd, y, z, homo, y^-1, xinv, finish), then the
// averages when they are on.
template <typename T>
int class_step(fhip_batch *b, BClass &k, hipStream_t s, bool materialize, int next_step)
{
    const Rep<T> *r = (const Rep<T> *)k.dreps;
    const unsigned n = (unsigned)k.members.size();
    if (k.heat) {                       // leapfrog.F:38-83, before the drift
        kb_layer_ke<T><<<dim3(k.Lz, 1, n), NT, 0, s>>>(r);
        kb_layer_scale<T><<<dim3((k.Lx * k.Ly + NT - 1) / NT, 2, n), NT, 0, s>>>(r);
    }
    // film/epit members: source rows first, then the mirrored rows from them
    // (bulk members do everything in the first launch and skip the second)
    RC(l_xfwd<T>(b, k, s, k.film ? XP_SOURCE : XP_ALL));
    if (k.film) RC(l_xfwd<T>(b, k, s, XP_MIRROR));
    RC(l_ypass<T>(b, k, s, -1));
    const int want = (materialize && !k.acoustic) ? 1 : 0;
    if (k.zsp) {
        if (k.acoustic) RC((l_zpair<T, ZG_ACOU, ZG_DDZ>(b, k, s, want)));
        else RC((l_zpair<T, ZG_STRAIN, ZG_DDZ>(b, k, s, want)));
    } else {
        if (k.acoustic) RC((l_zpair<T, ZG_ACOU, ZG_DD>(b, k, s, want)));
        else RC((l_zpair<T, ZG_STRAIN, ZG_DD>(b, k, s, want)));
    }
    kb_homo<T><<<dim3(1, 1, n), NT, 0, s>>>(r);
    {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return bfail(b, FHIP_ERR_HIP, std::string("launch homo: ") + hipGetErrorString(e));
    }
    RC(l_ypass<T>(b, k, s, +1));
    RC(l_xinv<T>(b, k, s, materialize ? 1 : 0));
    if (k.film) kb_mirror_kick<T><<<dim3(k.mblocks, 1, n), NT, 0, s>>>(r, xi_mode(b->method), materialize ? 1 : 0);
    kb_finish<T><<<dim3(1, 1, n), NT, 0, s>>>(r, fin_mode(b->method), next_step);
    if (k.avg) {
        if (k.hist) kb_avg_cells<T, 1><<<dim3(k.kblocks, 1, n), NT, 0, s>>>(r);
        else kb_avg_cells<T, 0><<<dim3(k.kblocks, 1, n), NT, 0, s>>>(r);
        kb_avg_fold<T><<<dim3(1, 1, n), NT, 0, s>>>(r);
        if (k.heat) kb_layer_acc<T><<<dim3(1, 1, n), NT, 0, s>>>(r);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, FHIP_ERR_HIP, std::string("launch finish: ") + hipGetErrorString(e));
    return FHIP_OK;
}

// First half kick and eq. 38 of the first Nose-Poincare step of a call.
template <typename T>
int class_np_prelude(fhip_batch *b, BClass &k, hipStream_t s)
{
    const Rep<T> *r = (const Rep<T> *)k.dreps;
    const unsigned n = (unsigned)k.members.size();
    kb_kick<T, 1><<<dim3(k.kblocks, 1, n), NT, 0, s>>>(r);
    kb_np_begin<T><<<dim3(1, 1, n), NT, 0, s>>>(r);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, FHIP_ERR_HIP, std::string("launch np prelude: ") + hipGetErrorString(e));
    return FHIP_OK;
}

// Pending velocity-scaling factors applied to the momenta.
template <typename T>
int class_vs_epilogue(fhip_batch *b, BClass &k, hipStream_t s)
{
    const Rep<T> *r = (const Rep<T> *)k.dreps;
    const unsigned n = (unsigned)k.members.size();
    kb_apply_scale<T><<<dim3(k.kblocks, 1, n), NT, 0, s>>>(r);
    kb_reset_scale<T><<<dim3(1, 1, n), 1, 0, s>>>(r);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, FHIP_ERR_HIP, std::string("launch vs epilogue: ") + hipGetErrorString(e));
    return FHIP_OK;
}

enum Phase { PH_PRELUDE = 0, PH_STEP = 1, PH_LAST = 2, PH_EPILOGUE = 3 };

int class_phase(fhip_batch *b, BClass &k, int phase)
{
    const bool f32 = k.prec == FHIP_FP32;
    hipStream_t s = k.stream;
    switch (phase) {
    case PH_PRELUDE:
        return f32 ? class_np_prelude<float>(b, k, s) : class_np_prelude<double>(b, k, s);
    case PH_STEP:
    case PH_LAST: {
        const bool last = phase == PH_LAST;
        const int next = (b->method == FHIP_MD && !last) ? 1 : 0;
        return f32 ? class_step<float>(b, k, s, last, next) : class_step<double>(b, k, s, last, next);
    }
    default:
        return f32 ? class_vs_epilogue<float>(b, k, s) : class_vs_epilogue<double>(b, k, s);
    }
}

// One phase of every class: the class streams fork from the batch stream and
// join back into it.
int batch_phase(fhip_batch *b, int phase)
{
    BHIP(hipEventRecord(b->fork, b->stream));
    for (auto &k : b->classes) {
        BHIP(hipStreamWaitEvent(k.stream, b->fork, 0));
        RC(class_phase(b, k, phase));
        BHIP(hipEventRecord(k.done, k.stream));
        BHIP(hipStreamWaitEvent(b->stream, k.done, 0));
    }
    return FHIP_OK;
}

std::vector<int> graph_key(const fhip_batch *b)
{
    std::vector<int> key;
    for (const auto &k : b->classes) {
        key.push_back(k.avg ? 1 : 0);
        key.push_back(k.hist ? 1 : 0);
        key.push_back(k.heat ? 1 : 0);
        key.push_back(k.xrows);
        key.push_back(k.ytile);
    }
    return key;
}

constexpr int kBatchGraphSteps[3] = {16, 4, 1};

void drop_graph(fhip_batch *b)
{
    for (int j = 0; j < 3; ++j) {
        if (b->gexec[j]) hipGraphExecDestroy(b->gexec[j]);
        if (b->graph[j]) hipGraphDestroy(b->graph[j]);
        b->gexec[j] = nullptr;
        b->graph[j] = nullptr;
    }
    b->gkey.clear();
}

// Capture 16, 4 and 1 middle steps (every class) into three graphs, once per
// averaging and tile configuration: a long call replays mostly 16-step
// graphs, whose kernels run back to back (a graph launch leaves the device
// idle for a few microseconds at its boundary).
int ensure_graph(fhip_batch *b)
{
    const std::vector<int> key = graph_key(b);
    if (b->gexec[0] && key == b->gkey) return FHIP_OK;
    drop_graph(b);
    for (int j = 0; j < 3; ++j) {
        BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = FHIP_OK;
        for (int i = 0; i < kBatchGraphSteps[j] && rc == FHIP_OK; ++i) rc = batch_phase(b, PH_STEP);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc != FHIP_OK) {
            if (g) hipGraphDestroy(g);
            drop_graph(b);
            return rc;
        }
        if (e != hipSuccess) {
            drop_graph(b);
            return bfail(b, FHIP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        }
        b->graph[j] = g;
        e = hipGraphInstantiate(&b->gexec[j], g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            drop_graph(b);
            return bfail(b, FHIP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    }
    b->gkey = key;
    return FHIP_OK;
}

// The autotuner: for each of the two tile knobs, time the full and the half
// tile on the class's own replicas (8 launches after 2 of warm-up) and keep
// the faster.  The timing launches have no lasting effect: xfwd runs without
// drift and writes only the spectra, the y pass transforms the spectra in
// place, and both are rewritten by the next step before anything reads them.
template <typename T>
int tune_class(fhip_batch *b, BClass &k)
{
    const int xf = tile_rows(k.Lx), xh = xfwd_tile_rows(k.Lx);
    const int yf = line_tile_rows<T>(k.Ly), yh = yf > 1 ? yf / 2 : 1;
    {
        std::lock_guard<std::mutex> lock(g_plan_mu);
        auto it = g_plans.find(plan_key(k));
        if (it != g_plans.end()) {
            set_plan(k, it->second.first, it->second.second);
            k.tuned = true;
            return FHIP_OK;
        }
    }
    hipStream_t s = b->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BHIP(hipEventCreate(&e0));
    BHIP(hipEventCreate(&e1));
    int rc = FHIP_OK;
    auto timed = [&](auto &&launch, double &us) {
        for (int i = 0; i < 10 && rc == FHIP_OK; ++i) {
            if (i == 2 && hipEventRecord(e0, s) != hipSuccess) rc = bfail(b, FHIP_ERR_HIP, "tuner event");
            rc = launch();
        }
        if (rc != FHIP_OK) return;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = bfail(b, FHIP_ERR_HIP, "tuner event");
            return;
        }
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e0, e1);
        us = 1e3 * ms / 8;
    };
    int bx = xf, by = yf;
    set_plan(k, xf, yf);
    timed([&] { return l_xfwd<T>(b, k, s, XP_ALL, XF_FORCE); }, k.tx[0]);
    k.tx[1] = k.tx[0];
    if (rc == FHIP_OK && xh != xf) {
        set_plan(k, xh, yf);
        timed([&] { return l_xfwd<T>(b, k, s, XP_ALL, XF_FORCE); }, k.tx[1]);
        if (k.tx[1] < k.tx[0]) bx = xh;
    }
    set_plan(k, bx, yf);
    if (rc == FHIP_OK) timed([&] { return l_ypass<T>(b, k, s, -1); }, k.ty[0]);
    k.ty[1] = k.ty[0];
    if (rc == FHIP_OK && yh != yf) {
        set_plan(k, bx, yh);
        timed([&] { return l_ypass<T>(b, k, s, -1); }, k.ty[1]);
        if (k.ty[1] < k.ty[0]) by = yh;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    RC(rc);
    set_plan(k, bx, by);
    k.tuned = true;
    std::lock_guard<std::mutex> lock(g_plan_mu);
    g_plans[plan_key(k)] = {bx, by};
    return FHIP_OK;
}

int refresh(fhip_batch *b)
{
    for (auto &k : b->classes) {
        bool avg = false, hist = false, heat = false;
        for (size_t i = 0; i < k.members.size(); ++i) {
            fhip_ctx *c = k.members[i];
            if (c->avg_auto && !c->davg) return bfail(b, FHIP_ERR_STATE, "averages not reset (fhip_avg_reset)");
            if (b->method == FHIP_LF && (c->delta_e != 0.0 || c->delta_e_acou != 0.0)) heat = true;
            if (i == 0) {
                avg = c->avg_auto;
                hist = c->avg_auto && c->avg_hist;
            } else if (avg != c->avg_auto || (avg && hist != c->avg_hist)) {
                return bfail(b, FHIP_ERR_UNSUPPORTED,
                             "replicas of one size class must share the averaging settings (fhip_avg_auto, histograms)");
            }
        }
        k.avg = avg;
        k.hist = hist;
        k.heat = heat;
        RC(k.prec == FHIP_FP32 ? upload_reps<float>(b, k) : upload_reps<double>(b, k));
    }
    return FHIP_OK;
}

template <typename T>
int prepare_lds(fhip_batch *b, BClass &k)
{
    // side buffer of the kx = 0 pencils: two NF-field values per point of up
    // to (TB - 1) / Mx + 1 pencils per tile, NF = 3 (dipole-dipole), 6
    // (strain) or 9 (acoustic)
    k.zshm_g = zside_bytes<T>(k.Lz, k.Lx / 2, k.acoustic ? 9 : 6);
    if (k.zshm_g > 96 * 1024)
        return bfail(b, FHIP_ERR_UNSUPPORTED, "z pass: too many kx = 0 pencils per tile (Lx too small for Lz)");
    fhip_ctx *ctx = nullptr;
    const void *fn = k.zsp ? (k.acoustic ? zpair_fn<T, ZG_ACOU, ZG_DDZ>(k.Lz) : zpair_fn<T, ZG_STRAIN, ZG_DDZ>(k.Lz))
                           : (k.acoustic ? zpair_fn<T, ZG_ACOU, ZG_DD>(k.Lz) : zpair_fn<T, ZG_STRAIN, ZG_DD>(k.Lz));
    int rc = ensure_lds(ctx, fn, k.zshm_g);
    if (rc != FHIP_OK) return bfail(b, rc, "z pass: LDS attribute");
    // The two groups of a tile run in separate workgroups when the z tiles of
    // every class of the batch, so split, fit in one round of this kernel's
    // workgroup slots (the classes' z passes run at the same time on their
    // streams), and always where the paired kernel has no prefetch of its
    // second group to give up (z edges below 64, species of different Z*);
    // FHIP_ZSPLIT=0 or 1 overrides the choice.
    long tiles = 0;
    for (const BClass &o : b->classes) tiles += (long)o.zblocks * (long)o.members.size();
    int per_cu = 0;
    BHIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NT, k.zshm_g));
    const long slots = (long)std::max(per_cu, 1) * std::max(k.members[0]->ncu, 1);
    const char *zs = std::getenv("FHIP_ZSPLIT");
    k.zsplit = (zs && (zs[0] == '0' || zs[0] == '1')) ? zs[0] == '1' : (2 * tiles <= slots || k.Lz < 64 || k.zsp);
    return FHIP_OK;
}

}  // namespace

extern "C" {

const char *fhip_batch_last_error(const fhip_batch *b) { return b ? b->err.c_str() : "null batch"; }

void fhip_batch_destroy(fhip_batch *b)
{
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (auto &k : b->classes) {
        if (k.stream) hipStreamSynchronize(k.stream);
        if (k.dreps) hipFree(k.dreps);
        if (k.done) hipEventDestroy(k.done);
        if (k.stream) hipStreamDestroy(k.stream);
    }
    if (b->fork) hipEventDestroy(b->fork);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

int fhip_batch_create(fhip_ctx *const *ctxs, int n, fhip_batch **out)
{
    if (!out) return FHIP_ERR_ARG;
    *out = nullptr;
    fhip_batch *b = new fhip_batch();
    *out = b;
    if (!ctxs || n < 1) return bfail(b, FHIP_ERR_ARG, "need at least one context");
    for (int i = 0; i < n; ++i) {
        fhip_ctx *c = ctxs[i];
        if (!c) return bfail(b, FHIP_ERR_ARG, "null context");
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == c) return bfail(b, FHIP_ERR_ARG, "a context appears twice");
        if (i == 0) b->method = c->p.method;
        else if (c->p.method != b->method)
            return bfail(b, FHIP_ERR_UNSUPPORTED, "every replica of a batch uses the same method");
        if (!batch_size_ok(c->Lx) || !batch_size_ok(c->Ly) || !batch_size_ok(c->Lz))
            return bfail(b, FHIP_ERR_UNSUPPORTED, "supercell edge not supported");
        b->ctxs.push_back(c);
        BClass *k = nullptr;
        for (auto &q : b->classes)
            if (q.Lx == c->Lx && q.Ly == c->Ly && q.Lz == c->Lz && q.prec == c->precision &&
                q.acoustic == c->c.acoustic && q.zsp == c->c.zsp)
                k = &q;
        if (!k) {
            b->classes.emplace_back();
            k = &b->classes.back();
            k->prec = c->precision;
            k->Lx = c->Lx;
            k->Ly = c->Ly;
            k->Lz = c->Lz;
            k->acoustic = c->c.acoustic;
            k->zsp = c->c.zsp;
            k->nf = 9 + (c->c.acoustic ? 3 : 0) + (c->c.zsp ? 3 : 0);
            const int t = c->precision == FHIP_FP32 ? 1 : 0;
            k->xblocks = c->xblocks;
            k->xfblocks = c->xfblocks;
            k->yblocks = c->yblocks[t];
            k->xrows = xfwd_tile_rows(c->Lx);            // the single-context plan until tuned
            k->ytile = t ? line_tile_rows<float>(c->Ly) : line_tile_rows<double>(c->Ly);
            k->zblocks = c->zblocks[t];
            k->kblocks = c->kblocks;
        }
        if (c->c.mir_lo < c->Lz) {
            k->film = true;
            const size_t cells = (size_t)(c->Lz - c->c.mir_lo) * c->Lx * c->Ly;
            k->mblocks = std::max(k->mblocks, (int)((cells + NT - 1) / NT));
        }
        k->members.push_back(c);
    }
    BHIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    BHIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    for (auto &k : b->classes) {
        BHIP(hipStreamCreateWithFlags(&k.stream, hipStreamNonBlocking));
        BHIP(hipEventCreateWithFlags(&k.done, hipEventDisableTiming));
        RC(k.prec == FHIP_FP32 ? prepare_lds<float>(b, k) : prepare_lds<double>(b, k));
    }
    return FHIP_OK;
}

int fhip_batch_graph(fhip_batch *b, int enable)
{
    if (!b) return FHIP_ERR_ARG;
    b->use_graph = enable != 0;
    return FHIP_OK;
}

int fhip_batch_step(fhip_batch *b, int nsteps, fhip_energy *e)
{
    if (!b) return FHIP_ERR_ARG;
    if (nsteps < 1) return bfail(b, FHIP_ERR_ARG, "nsteps must be >= 1");
    // the contexts' own streams are idle after their (synchronous) calls;
    // make sure of it before the batch streams touch their memory
    for (fhip_ctx *c : b->ctxs) BHIP(hipStreamSynchronize(c->stream));
    RC(refresh(b));
    if (b->tune)
        for (auto &k : b->classes)
            if (!k.tuned) {
                RC(k.prec == FHIP_FP32 ? tune_class<float>(b, k) : tune_class<double>(b, k));
                BHIP(hipStreamSynchronize(b->stream));
            }
    if (b->method == FHIP_MD) RC(batch_phase(b, PH_PRELUDE));
    const int middle = nsteps - 1;
    if (middle > 0) {
        if (b->use_graph) {
            RC(ensure_graph(b));
            int left = middle;
            for (int j = 0; j < 3; ++j)
                for (; left >= kBatchGraphSteps[j]; left -= kBatchGraphSteps[j])
                    BHIP(hipGraphLaunch(b->gexec[j], b->stream));
        } else {
            for (int i = 0; i < middle; ++i) RC(batch_phase(b, PH_STEP));
        }
    }
    RC(batch_phase(b, PH_LAST));
    if (b->method == FHIP_VS) RC(batch_phase(b, PH_EPILOGUE));
    for (fhip_ctx *c : b->ctxs)
        BHIP(hipMemcpyAsync(c->hsc, c->dsc, sizeof(Scalars), hipMemcpyDeviceToHost, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < b->ctxs.size(); ++i) {
        fhip_ctx *c = b->ctxs[i];
        if (!c->c.acoustic) c->acouK_valid = true;
        energies_out(c, e ? e + i : nullptr);
    }
    return FHIP_OK;
}

int fhip_batch_tune(fhip_batch *b, int enable)
{
    if (!b) return FHIP_ERR_ARG;
    b->tune = enable != 0;
    return FHIP_OK;
}

int fhip_batch_plan(const fhip_batch *b, int cls, int32_t shape[5], double us[4])
{
    if (!b || !shape) return FHIP_ERR_ARG;
    if (cls < 0 || cls >= (int)b->classes.size()) return FHIP_ERR_ARG;
    const BClass &k = b->classes[cls];
    shape[0] = k.Lx; shape[1] = k.Ly; shape[2] = k.Lz; shape[3] = k.xrows; shape[4] = k.ytile;
    if (us) {
        us[0] = k.tx[0]; us[1] = k.tx[1]; us[2] = k.ty[0]; us[3] = k.ty[1];
    }
    return k.tuned ? 1 : 0;
}

int fhip_batch_size(const fhip_batch *b) { return b ? (int)b->ctxs.size() : 0; }
int fhip_batch_classes(const fhip_batch *b) { return b ? (int)b->classes.size() : 0; }

}  // extern "C"
