// fhip_host.hpp -- host-side context record and helpers shared by the
// translation units of the library (single-context steps in feram_hip.hip,
// replica batches in fhip_batch.hip).
#pragma once
#include "../../include/feram_hip.h"
#include "fhip_kernels.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace fhip;

// Supercell edges with a compiled FFT length: products of 2, 3 and 5 (the
// edges of the reference's examples up to 256, including the odd 81 and 243
// of its gap_id = 1 films).
#define FHIP_SIZES(X) X(4) X(6) X(8) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(60) X(64) X(80) X(81) \
    X(90) X(96) X(100) X(128) X(192) X(243) X(256)
#define FHIP_SIZES_TEXT "4 6 8 12 16 20 24 32 40 48 60 64 80 81 90 96 100 128 192 243 256"

namespace fhip_host {

inline int tile_rows(int n, int cap = 1024)   // must agree with tile_of<n, cap>::value
{
    int p = 1;
    while (2 * p * n <= cap) p *= 2;
    return p;
}

// Rows per xinv workgroup (xinv_rows_of on the device).
inline int xinv_tile_rows(int n) { return fhip::xinv_rows_of(n); }

// Rows per xfwd workgroup (xfwd_rows<n> on the device).
inline int xfwd_tile_rows(int n) { const int t = tile_rows(n); return t > 1 ? t / 2 : 1; }

// Lines per y tile of spectral type T (line_tile_cap<T>() on the device) and
// per z tile (ZTile<T, n>::TB).
template <typename T>
inline int line_tile_rows(int n) { return tile_rows(n, 16384 / (int)sizeof(cpx<T>)); }
template <typename T>
inline int ztile_rows(int n) { return ztile_pencils(n, (int)sizeof(cpx<T>)); }

// Dynamic LDS of a z pass (bytes): the side buffer of the kx = 0 pencils, two
// nf-field values per point of up to (TB - 1) / Mx + 1 such pencils per tile.
template <typename T>
inline size_t zside_bytes(int Lz, int Mx, int nf)
{
    return 2 * (size_t)((ztile_rows<T>(Lz) - 1) / Mx + 1) * Lz * nf * sizeof(cpx<T>);
}

enum KernelId {
    K_XFWD = 0, K_YFWD, K_ZDD, K_ZSTRAIN, K_HOMO, K_YINV, K_XINV, K_FINISH, K_NPBEGIN, K_KICK,
    K_KE, K_SCALE, K_SETUP, K_AVG, K_ZPAIR, K_NKERNEL
};

}  // namespace fhip_host

template <typename T>
struct TypedSet {
    Spectra<T> sp{};
    cpx<T> *twx = nullptr, *twy = nullptr, *twz = nullptr;
    T *Lt = nullptr;                  // L(k) in the spectral type, 6 words per point
};

struct fhip_ctx {
    fhip_params p{};
    int Lx = 0, Ly = 0, Lz = 0, Kx = 0, Mx = 0, N = 0, Kh = 0, nslots = 0;
    size_t Mh = 0;
    int xblocks = 0, xfblocks = 0, kblocks = 0, maxblocks = 0;   // xinv, xfwd, cell-wise grids
    int yblocks[2] = {0, 0}, zblocks[2] = {0, 0};   // per spectral type: [0] FP64, [1] FP32
    int zsplit[2] = {-1, -1};         // per spectral type: the z pair runs its groups in separate workgroups (-1: not chosen yet)
    int ncu = 0;                      // compute units of the context's device
    hipStream_t stream = nullptr;
    std::string err;
    fhip::Consts c{};
    fhip::ScalarConsts sk{};
    fhip::Fields f{};                 // FP64 state (plain mode), local field, dipoRavg, cell codes
    fhip::FieldsT<float> f32{};       // FP32 state (FP32 mode); its other pointers are taken from f
    int precision = FHIP_FP64;
    TypedSet<double> dset;            // FP64 spectra (plain mode) and the FP64 set-up
    TypedSet<float> fset;             // FP32 spectra (FP32 mode)
    Scalars *dsc = nullptr;
    Scalars *hsc = nullptr;           // pinned mirror
    double *partials = nullptr;
    double *Ltab = nullptr;           // FP64 L(k), 6 words per point (master copy)
    cpx<double> *acouK = nullptr;
    cpx<double> *modk = nullptr;      // FFT of the modulation field on the k array
    AvgAcc *davg = nullptr;           // averages (add_to_Average sums)
    unsigned long long *dist1 = nullptr;
    int *dist3 = nullptr;
    bool avg_auto = false, avg_hist = false, avg_ravg = false;
    cpx<double> *csx = nullptr, *csy = nullptr, *csz = nullptr;
    double *kvx = nullptr, *kvy = nullptr, *kvz = nullptr;
    size_t Ks = 0;                    // spectral slot stride: Mh packed points
    unsigned char *code = nullptr;
    std::vector<void *> allocs;
    int64_t bytes = 0;
    bool acouK_valid = false;
    // heat-source layers (leapfrog only)
    double delta_e = 0.0, delta_e_acou = 0.0;
    double *el = nullptr;             // [2][Lz] e_layer_dipo, e_layer_acou of the last step
    double *el_sum = nullptr;         // [2][Lz] their sums over the averaging steps
    int *nlayer = nullptr;            // counted (sum_p) cells per z layer
    std::vector<int> nlayer_host;
    // measurement
    bool prof = false;
    // graphs of 16, 4 and 1 middle steps of a multi-step call
    // (kGraphSteps), and the by-value launch state they were captured with
    bool use_graph = true;
    hipGraph_t sgraph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t sgexec[3] = {nullptr, nullptr, nullptr};
    std::vector<unsigned char> skey;
    std::vector<hipEvent_t> ev_pool;
    struct Rec { int id; hipEvent_t a, b; };
    std::vector<Rec> recs;
    double kms[fhip_host::K_NKERNEL] = {0};
    int64_t kcount[fhip_host::K_NKERNEL] = {0};
};

namespace fhip_host {

template <typename T> constexpr int tix() { return sizeof(T) == 4 ? 1 : 0; }
template <typename T> TypedSet<T> &tset(fhip_ctx *ctx);
template <> inline TypedSet<double> &tset<double>(fhip_ctx *ctx) { return ctx->dset; }
template <> inline TypedSet<float> &tset<float>(fhip_ctx *ctx) { return ctx->fset; }

// Per-cell fields with state of type S: the state pointers are those of the
// context's precision (null in the other type), the local field, dipoRavg
// and cell-code pointers are shared.
template <typename S> fhip::FieldsT<S> fields(const fhip_ctx *ctx);
template <> inline fhip::Fields fields<double>(const fhip_ctx *ctx) { return ctx->f; }
template <> inline fhip::FieldsT<float> fields<float>(const fhip_ctx *ctx)
{
    fhip::FieldsT<float> g = ctx->f32;
    for (int q = 0; q < 3; ++q) {
        g.eloc[q] = ctx->f.eloc[q];
        g.uavg[q] = ctx->f.uavg[q];
    }
    g.code = ctx->f.code;
    return g;
}

inline int fail(fhip_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(ctx, FHIP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define RC(expr)                       \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != FHIP_OK) return rc_; \
    } while (0)

// Static plus dynamic LDS of a kernel above 64 KB needs the kernel's dynamic
// limit raised (gfx950 has 160 KB per CU); done once per kernel and size.
inline int ensure_lds(fhip_ctx *ctx, const void *fn, size_t dyn)
{
    static std::mutex mu;
    static std::map<const void *, size_t> done;
    std::lock_guard<std::mutex> lock(mu);
    auto it = done.find(fn);
    if (it != done.end() && it->second >= dyn) return FHIP_OK;
    hipFuncAttributes fa{};
    HIPCHK(hipFuncGetAttributes(&fa, fn));
    // every launch passes NT threads per workgroup, the block size the
    // kernel's index maps and LDS arrays are sized for (__launch_bounds__(NT))
    if (fa.maxThreadsPerBlock > 0 && fa.maxThreadsPerBlock < NT)
        return fail(ctx, FHIP_ERR_STATE, "kernel cannot run the launch block size");
    if (fa.sharedSizeBytes + dyn > 160 * 1024)
        return fail(ctx, FHIP_ERR_UNSUPPORTED, "kernel needs more LDS than a CU holds");
    if (fa.sharedSizeBytes + dyn > 64 * 1024)
        HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    done[fn] = dyn;
    return FHIP_OK;
}

inline const char *const kKernelNames[K_NKERNEL] = {
    "xfwd_drift_r2c", "y_forward", "z_dipole_dipole", "z_strain", "homo_strain", "y_inverse",
    "xinv_c2r_local_kick", "finish_scalars", "np_begin", "np_kick", "kinetic", "apply_scale", "setup",
    "average", "z_pair"};

inline hipEvent_t take_event(fhip_ctx *ctx)
{
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Launch helper: brackets the launch with events when measuring.
template <typename F>
int launch(fhip_ctx *ctx, int id, F &&fn, hipStream_t st = nullptr)
{
    if (!st) st = ctx->stream;
    hipEvent_t a = nullptr, b = nullptr;
    if (ctx->prof) {
        a = take_event(ctx);
        b = take_event(ctx);
        if (a) (void)hipEventRecord(a, st);
    }
    fn();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(ctx, FHIP_ERR_HIP, std::string("launch ") + kKernelNames[id] + ": " + hipGetErrorString(e));
    if (ctx->prof && a && b) {
        (void)hipEventRecord(b, st);
        ctx->recs.push_back({id, a, b});
    }
    return FHIP_OK;
}

inline int collect_events(fhip_ctx *ctx)
{
    if (ctx->recs.empty()) return FHIP_OK;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (auto &r : ctx->recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            ctx->kms[r.id] += ms;
            ctx->kcount[r.id] += 1;
        }
        ctx->ev_pool.push_back(r.a);
        ctx->ev_pool.push_back(r.b);
    }
    ctx->recs.clear();
    return FHIP_OK;
}

inline void energies_out(const fhip_ctx *ctx, fhip_energy *e)
{
    if (!e) return;
    const double *en = ctx->hsc->en;
    e->acou_kinetic = en[E_ACOU_KIN];
    e->dipo_kinetic = en[E_DIPO_KIN];
    e->short_range = en[E_SR];
    e->long_range = en[E_LR];
    e->dipole_E_field = en[E_DE];
    e->unharmonic = en[E_UNH];
    e->homo_strain = en[E_HS];
    e->homo_coupling = en[E_HC];
    e->inho_strain = en[E_IS];
    e->inho_coupling = en[E_IC];
    e->inho_modulation = en[E_IM];
    e->total = en[E_TOTAL];
    e->H_Nose_Poincare = en[E_HNP];
}

// Arguments of a z pass of the context's spectral type T.
template <typename T>
inline ZArgs<T> zargs(fhip_ctx *ctx, bool want_acouK)
{
    ZArgs<T> a{};
    a.Ltab = tset<T>(ctx).Lt;
    a.csx = ctx->csx;
    a.csy = ctx->csy;
    a.csz = ctx->csz;
    a.kvx = ctx->kvx;
    a.kvy = ctx->kvy;
    a.kvz = ctx->kvz;
    a.modk = ctx->modk;
    a.acouK = want_acouK ? ctx->acouK : nullptr;
    a.partials = ctx->partials;
    a.maxblocks = ctx->maxblocks;
    a.nplain = 0;
    return a;
}

}  // namespace fhip_host

// ---- per-edge stage launchers ----------------------------------------------
// The kernels whose FFT length is a template parameter are launched through
// these functions, one explicit instantiation per edge; fhip_stages.hip
// defines them and is compiled once per group of edges, so the edges build in
// parallel translation units.
namespace fhip_stage {

// single context (arguments from the context, launched on its stream)
template <typename T, int n> int xfwd(fhip_ctx *ctx, const XfwdArgs &a);
template <typename T, int n> int xinv(fhip_ctx *ctx, const XinvArgs &a, int nf);
template <typename T, int n> int ypass(fhip_ctx *ctx, int sign, int slot0, int nf, hipStream_t s);
template <typename T, int n> int zplain(fhip_ctx *ctx, const ZArgs<T> &a, hipStream_t s, size_t shm);
template <typename T, int n, int G2, int DDG> int zpair(fhip_ctx *ctx, const ZArgs<T> &a, size_t shm);

// replica batches: one launch over every replica of a size class
template <typename T, int n> void b_xfwd(const Rep<T> *r, dim3 g, hipStream_t s, bool half, int phase, int mode);
template <typename T, int n> void b_ypass(const Rep<T> *r, dim3 g, hipStream_t s, bool half, int sign);
template <typename T, int n, int G2, int DDG> void b_zpair(const Rep<T> *r, dim3 g, hipStream_t s, size_t shm, int want_acouK);
template <typename T, int n, int G2, int DDG> const void *b_zpair_fn();
template <typename T, int n> void b_xinv(const Rep<T> *r, dim3 g, hipStream_t s, int nf, int materialize);

}  // namespace fhip_stage

// Edges of the batched instantiations: every supported edge.
#define FHIP_BATCH_SIZES(X) FHIP_SIZES(X)
