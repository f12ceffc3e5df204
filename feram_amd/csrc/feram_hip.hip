// feram_hip.hip -- context, set-up and step orchestration behind the C ABI of
// include/feram_hip.h.  One HIP stream per context; every kernel of a step is
// launched on it back to back, and the host only synchronises when it reads
// energies or fields.
#include "fhip_host.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <map>
#include <mutex>

using namespace fhip;
using namespace fhip_host;

namespace {

constexpr double M_PI_D = 3.14159265358979323846;
constexpr double HARTREE_BOHR = 14.399644;      // define.h:32
constexpr double KELVIN_EV = 11604.45;          // define.h:33
constexpr double ELEMENTARY_CHARGE = 1.60217653;

bool size_supported(int n)
{
    switch (n) {
#define FHIP_CASE(n) case n:
        FHIP_SIZES(FHIP_CASE)
#undef FHIP_CASE
        return true;
    default:
        return false;
    }
}

template <typename P>
int dalloc(fhip_ctx *ctx, P **ptr, size_t count)
{
    void *v = nullptr;
    size_t nb = count * sizeof(P);
    if (nb == 0) nb = 16;
    hipError_t e = hipMalloc(&v, nb);
    if (e != hipSuccess) return fail(ctx, FHIP_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemsetAsync(v, 0, nb, ctx->stream);
    if (e != hipSuccess) return fail(ctx, FHIP_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    ctx->allocs.push_back(v);
    ctx->bytes += (int64_t)nb;
    *ptr = (P *)v;
    return FHIP_OK;
}


// ---- size dispatch ---------------------------------------------------------
// Each case calls the launcher instantiated for that edge (fhip_stages.hip).
template <typename T>
int run_xfwd(fhip_ctx *ctx, const XfwdArgs &a)
{
    switch (ctx->Lx) {
#define FHIP_CASE(n) \
    case n: return fhip_stage::xfwd<T, n>(ctx, a);
        FHIP_SIZES(FHIP_CASE)
#undef FHIP_CASE
    }
    return fail(ctx, FHIP_ERR_UNSUPPORTED, "Lx not supported");
}

template <typename T>
int run_xinv(fhip_ctx *ctx, const XinvArgs &a, int nf)
{
    switch (ctx->Lx) {
#define FHIP_CASE(n) \
    case n: return fhip_stage::xinv<T, n>(ctx, a, nf);
        FHIP_SIZES(FHIP_CASE)
#undef FHIP_CASE
    }
    return fail(ctx, FHIP_ERR_UNSUPPORTED, "Lx not supported");
}

template <typename T>
int run_ypass(fhip_ctx *ctx, int sign, int slot0, int nf, hipStream_t st = nullptr)
{
    hipStream_t s = st ? st : ctx->stream;
    switch (ctx->Ly) {
#define FHIP_CASE(n) \
    case n: return fhip_stage::ypass<T, n>(ctx, sign, slot0, nf, s);
        FHIP_SIZES(FHIP_CASE)
#undef FHIP_CASE
    }
    return fail(ctx, FHIP_ERR_UNSUPPORTED, "Ly not supported");
}

// Plain z transforms of nplain slots (set-up work).
template <typename T>
int run_zplain(fhip_ctx *ctx, const ZArgs<T> &a, hipStream_t st = nullptr)
{
    hipStream_t s = st ? st : ctx->stream;
    switch (ctx->Lz) {
#define FHIP_CASE(n) \
    case n: return fhip_stage::zplain<T, n>(ctx, a, s, 0);
        FHIP_SIZES(FHIP_CASE)
#undef FHIP_CASE
    }
    return fail(ctx, FHIP_ERR_UNSUPPORTED, "Lz not supported");
}

// Both z groups of a step in one launch (k_zpair).
template <typename T, int G2>
int run_zpair(fhip_ctx *ctx, const ZArgs<T> &a)
{
    // side buffer of the kx = 0 pencils: two NF-field values per point of up
    // to (TB - 1) / Mx + 1 pencils per tile, NF the larger of the two groups
    const int nf = (G2 == ZG_STRAIN) ? 6 : 9;
    const size_t shm = zside_bytes<T>(ctx->Lz, ctx->Mx, nf);
    if (shm > 96 * 1024)
        return fail(ctx, FHIP_ERR_UNSUPPORTED, "z pass: too many kx = 0 pencils per tile (Lx too small for Lz)");
    switch (ctx->Lz) {
#define FHIP_CASE(n)                                                                       \
    case n:                                                                                \
        return ctx->c.zsp ? fhip_stage::zpair<T, n, G2, ZG_DDZ>(ctx, a, shm)               \
                          : fhip_stage::zpair<T, n, G2, ZG_DD>(ctx, a, shm);
        FHIP_SIZES(FHIP_CASE)
#undef FHIP_CASE
    }
    return fail(ctx, FHIP_ERR_UNSUPPORTED, "Lz not supported");
}


// ---- averages (add_to_Average) ----------------------------------------------
template <typename S>
int avg_cells(fhip_ctx *ctx)
{
    Consts c = ctx->c;
    FieldsT<S> f = fields<S>(ctx);
    double *part = ctx->partials;
    int mb = ctx->maxblocks, kb = ctx->kblocks;
    unsigned long long *d1 = ctx->dist1;
    int *d3 = ctx->dist3;
    const int ravg = ctx->avg_ravg ? 1 : 0;
    hipStream_t s = ctx->stream;
    if (ctx->avg_hist)
        return launch(ctx, K_AVG, [&] { k_avg_cells<1, S><<<kb, NT, 0, s>>>(c, f, part, mb, d1, d3, ravg); });
    return launch(ctx, K_AVG, [&] { k_avg_cells<0, S><<<kb, NT, 0, s>>>(c, f, part, mb, d1, d3, ravg); });
}

int avg_add(fhip_ctx *ctx)
{
    if (!ctx->davg) return fail(ctx, FHIP_ERR_STATE, "averages not reset (fhip_avg_reset)");
    if (ctx->precision == FHIP_FP32) RC(avg_cells<float>(ctx));
    else RC(avg_cells<double>(ctx));
    double *part = ctx->partials;
    int mb = ctx->maxblocks, kb = ctx->kblocks;
    hipStream_t s = ctx->stream;
    Scalars *sc = ctx->dsc;
    AvgAcc *acc = ctx->davg;
    RC(launch(ctx, K_AVG, [&] { k_avg_fold<<<1, NT, 0, s>>>(sc, acc, part, mb, kb); }));
    if (ctx->p.method == FHIP_LF && ctx->delta_e != 0.0) {   // average_module.F:82-85
        const double *el = ctx->el;
        double *sum = ctx->el_sum;
        const int n = 2 * ctx->Lz;
        RC(launch(ctx, K_AVG, [&] { k_layer_acc<<<1, NT, 0, s>>>(el, sum, n); }));
    }
    return FHIP_OK;
}

// ---- one force evaluation (mirror_force_mirror) inside a step ---------------
int nforward(const fhip_ctx *ctx) { return 9 + (ctx->c.acoustic ? 3 : 0) + (ctx->c.zsp ? 3 : 0); }

enum StepKind { SK_FORCE = 0, SK_NP = 1, SK_LF = 2, SK_VS = 3 };

// Heat-source and sink layers at the start of a leapfrog step (leapfrog.F:38-83).
template <typename S>
int heat_source_layers_t(fhip_ctx *ctx)
{
    Consts c = ctx->c;
    FieldsT<S> f = fields<S>(ctx);
    const int *nl = ctx->nlayer;
    double *el = ctx->el;
    const double de = ctx->delta_e, dea = ctx->delta_e_acou;
    hipStream_t s = ctx->stream;
    RC(launch(ctx, K_SCALE, [&] { k_layer_ke<S><<<ctx->Lz, NT, 0, s>>>(c, f, nl, el); }));
    const dim3 g((ctx->Lx * ctx->Ly + NT - 1) / NT, 2);
    return launch(ctx, K_SCALE, [&] { k_layer_scale<S><<<g, NT, 0, s>>>(c, f, el, de, dea); });
}

int heat_source_layers(fhip_ctx *ctx)
{
    return ctx->precision == FHIP_FP32 ? heat_source_layers_t<float>(ctx) : heat_source_layers_t<double>(ctx);
}

template <typename T>
int force_pipeline_t(fhip_ctx *ctx, int kind, bool materialize, int next_step)
{
    const int nf = nforward(ctx);
    if (kind == SK_LF && (ctx->delta_e != 0.0 || ctx->delta_e_acou != 0.0)) RC(heat_source_layers(ctx));
    XfwdArgs xa{};
    xa.mode = (kind == SK_NP) ? XF_NP : (kind == SK_LF ? XF_LF : (kind == SK_VS ? XF_VS : XF_FORCE));
    xa.nf = nf;
    const bool film = ctx->c.mir_lo < ctx->Lz;
    xa.phase = film ? XP_SOURCE : XP_ALL;
    RC(run_xfwd<T>(ctx, xa));
    if (film) {                                   // mirror_force_mirror.F:20-27
        xa.phase = XP_MIRROR;
        RC(run_xfwd<T>(ctx, xa));
    }
    RC(run_ypass<T>(ctx, -1, 0, nf));
    ZArgs<T> za = zargs<T>(ctx, materialize && !ctx->c.acoustic);
    if (ctx->c.acoustic) RC((run_zpair<T, ZG_ACOU>(ctx, za)));
    else RC((run_zpair<T, ZG_STRAIN>(ctx, za)));
    if (materialize && !ctx->c.acoustic) ctx->acouK_valid = true;
    {
        ScalarConsts k = ctx->sk;
        Scalars *sc = ctx->dsc;
        double *part = ctx->partials;
        int mb = ctx->maxblocks, zb = ctx->zblocks[tix<T>()];
        hipStream_t s = ctx->stream;
        RC(launch(ctx, K_HOMO, [&] { k_homo<<<1, NT, 0, s>>>(k, sc, part, mb, zb); }));
    }
    RC(run_ypass<T>(ctx, +1, 0, nf));
    XinvArgs ia{};
    ia.mode = (kind == SK_NP) ? XI_NP : (kind == SK_LF ? XI_LF : (kind == SK_VS ? XI_VS : XI_FORCE));
    ia.materialize = materialize ? 1 : 0;
    ia.partials = ctx->partials;
    ia.maxblocks = ctx->maxblocks;
    RC(run_xinv<T>(ctx, ia, nf));
    if (film) {                                   // mirror_force_mirror.F:80-84
        Consts c = ctx->c;
        FieldsT<T> f = fields<T>(ctx);
        const Scalars *sc = ctx->dsc;
        const int mode = ia.mode, mat = ia.materialize;
        const int nb = (int)(((size_t)(ctx->Lz - c.mir_lo) * ctx->Lx * ctx->Ly + NT - 1) / NT);
        hipStream_t s = ctx->stream;
        RC(launch(ctx, K_XINV, [&] { k_mirror_kick<T><<<nb, NT, 0, s>>>(c, f, sc, mode, mat); }));
    }
    {
        ScalarConsts k = ctx->sk;
        Scalars *sc = ctx->dsc;
        double *part = ctx->partials;
        int mb = ctx->maxblocks, xb = ctx->xblocks, zb = ctx->zblocks[tix<T>()];
        int mode = (kind == SK_NP) ? FIN_NP : (kind == SK_LF ? FIN_LF : (kind == SK_VS ? FIN_VS : FIN_FORCE));
        int ac = ctx->c.acoustic;
        hipStream_t s = ctx->stream;
        RC(launch(ctx, K_FINISH, [&] { k_finish<<<1, NT, 0, s>>>(k, sc, part, mb, xb, zb, mode, ac, next_step); }));
    }
    return FHIP_OK;
}

int force_pipeline(fhip_ctx *ctx, int kind, bool materialize, int next_step = 0)
{
    if (ctx->precision == FHIP_FP32) return force_pipeline_t<float>(ctx, kind, materialize, next_step);
    return force_pipeline_t<double>(ctx, kind, materialize, next_step);
}

// Middle steps are replayed from graphs of 16, 4 and 1 steps: the kernels of
// one graph run back to back, and a graph launch costs ~9 us of idle device
// time at its boundary (128^3), so a long call is mostly 16-step launches.
constexpr int kGraphSteps[3] = {16, 4, 1};

void drop_step_graph(fhip_ctx *ctx)
{
    for (int j = 0; j < 3; ++j) {
        if (ctx->sgexec[j]) hipGraphExecDestroy(ctx->sgexec[j]);
        if (ctx->sgraph[j]) hipGraphDestroy(ctx->sgraph[j]);
        ctx->sgexec[j] = nullptr;
        ctx->sgraph[j] = nullptr;
    }
    ctx->skey.clear();
}

// Everything a middle step passes to its kernels by value or decides on the
// host: a graph captured with one key replays exactly that step.
std::vector<unsigned char> step_key(const fhip_ctx *ctx, int kind, int next_step)
{
    std::vector<unsigned char> k;
    auto put = [&](const void *v, size_t n) {
        const unsigned char *b = (const unsigned char *)v;
        k.insert(k.end(), b, b + n);
    };
    const int flags[4] = {kind, next_step, ctx->avg_auto ? 1 + (ctx->avg_hist ? 2 : 0) + (ctx->avg_ravg ? 4 : 0) : 0,
                          ctx->precision};
    put(flags, sizeof(flags));
    put(&ctx->c, sizeof(Consts));
    put(&ctx->sk, sizeof(ScalarConsts));
    put(&ctx->f, sizeof(Fields));
    put(&ctx->f32, sizeof(ctx->f32));
    const double de[2] = {ctx->delta_e, ctx->delta_e_acou};
    put(de, sizeof(de));
    const void *ptrs[7] = {ctx->davg, ctx->dist1, ctx->dist3, ctx->el, ctx->el_sum, ctx->nlayer, ctx->modk};
    put(ptrs, sizeof(ptrs));
    return k;
}

int one_step(fhip_ctx *ctx, int kind, bool materialize, int next_step)
{
    RC(force_pipeline(ctx, kind, materialize, next_step));
    if (ctx->avg_auto) RC(avg_add(ctx));
    return FHIP_OK;
}

// count steps that are not the last of a call (nothing materialised): with two
// or more, graphs of 16, 4 and 1 such steps (kernel nodes only) are captured
// together and replayed, largest first; the captures are kept while the
// launch state stays the same.
int middle_steps(fhip_ctx *ctx, int kind, int next_step, int count)
{
    if (count <= 0) return FHIP_OK;
    if (!ctx->use_graph || ctx->prof || count < 2) {
        for (int i = 0; i < count; ++i) RC(one_step(ctx, kind, false, next_step));
        return FHIP_OK;
    }
    std::vector<unsigned char> key = step_key(ctx, kind, next_step);
    if (!ctx->sgexec[0] || key != ctx->skey) {
        drop_step_graph(ctx);
        for (int j = 0; j < 3; ++j) {
            HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            int rc = FHIP_OK;
            for (int i = 0; i < kGraphSteps[j] && rc == FHIP_OK; ++i) rc = one_step(ctx, kind, false, next_step);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            if (rc != FHIP_OK) {
                if (g) hipGraphDestroy(g);
                drop_step_graph(ctx);
                return rc;
            }
            if (e != hipSuccess) {
                drop_step_graph(ctx);
                return fail(ctx, FHIP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            }
            ctx->sgraph[j] = g;
            const hipError_t ei = hipGraphInstantiate(&ctx->sgexec[j], g, nullptr, nullptr, 0);
            if (ei != hipSuccess) {
                drop_step_graph(ctx);
                return fail(ctx, FHIP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
            }
        }
        ctx->skey = key;
    }
    for (int j = 0; j < 3; ++j)
        for (; count >= kGraphSteps[j]; count -= kGraphSteps[j]) HIPCHK(hipGraphLaunch(ctx->sgexec[j], ctx->stream));
    return FHIP_OK;
}

int pull_scalars(fhip_ctx *ctx)
{
    HIPCHK(hipMemcpyAsync(ctx->hsc, ctx->dsc, sizeof(Scalars), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return collect_events(ctx);
}

int push_scalars(fhip_ctx *ctx)
{
    HIPCHK(hipMemcpyAsync(ctx->dsc, ctx->hsc, sizeof(Scalars), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FHIP_OK;
}


template <typename S>
int launch_kick_t(fhip_ctx *ctx, bool kick)
{
    Consts c = ctx->c;
    FieldsT<S> f = fields<S>(ctx);
    const Scalars *sc = ctx->dsc;
    double *part = ctx->partials;
    int mb = ctx->maxblocks, kb = ctx->kblocks;
    hipStream_t s = ctx->stream;
    if (kick) return launch(ctx, K_KICK, [&] { k_kick<1, S><<<kb, NT, 0, s>>>(c, f, sc, part, mb); });
    return launch(ctx, K_KE, [&] { k_kick<0, S><<<kb, NT, 0, s>>>(c, f, sc, part, mb); });
}

int launch_kick(fhip_ctx *ctx, bool kick)
{
    return ctx->precision == FHIP_FP32 ? launch_kick_t<float>(ctx, kick) : launch_kick_t<double>(ctx, kick);
}

// ---- host-side layout conversion -------------------------------------------
size_t padded_index(const fhip_ctx *ctx, int comp, int ncomp, int x, int y, int z)
{
    const size_t dx = 2 * (size_t)ctx->Kx, dy = (size_t)ctx->Ly + ctx->p.padding_y;
    return (size_t)comp + ncomp * ((size_t)x + dx * ((size_t)y + dy * (size_t)z));
}

int alloc_uavg(fhip_ctx *ctx)
{
    double *blk = nullptr;
    RC(dalloc(ctx, &blk, 3 * (size_t)ctx->N));
    for (int q = 0; q < 3; ++q) ctx->f.uavg[q] = blk + q * (size_t)ctx->N;
    return FHIP_OK;
}

// State fields are stored in FP32 in the FP32 mode (the host side always
// sees FP64): their device pointers of that type.
bool state_field(int field)
{
    return field == FHIP_DIPO_R || field == FHIP_DIPO_P || field == FHIP_DVDDI || field == FHIP_ACOU_R ||
           field == FHIP_ACOU_P || field == FHIP_DVDAC;
}

float **field_ptrs32(fhip_ctx *ctx, int field)
{
    switch (field) {
    case FHIP_DIPO_R: return ctx->f32.u;
    case FHIP_DIPO_P: return ctx->f32.p;
    case FHIP_DVDDI: return ctx->f32.F;
    case FHIP_ACOU_R: return ctx->f32.w;
    case FHIP_ACOU_P: return ctx->f32.pw;
    case FHIP_DVDAC: return ctx->f32.dVdw;
    default: return nullptr;
    }
}

double **field_ptrs(fhip_ctx *ctx, int field)
{
    switch (field) {
    case FHIP_DIPO_R: return ctx->f.u;
    case FHIP_DIPO_P: return ctx->f.p;
    case FHIP_DVDDI: return ctx->f.F;
    case FHIP_ACOU_R: return ctx->f.w;
    case FHIP_ACOU_P: return ctx->f.pw;
    case FHIP_DVDAC: return ctx->f.dVdw;
    case FHIP_E_LOCAL: return ctx->f.eloc;
    case FHIP_DIPO_R_AVG: return ctx->f.uavg;
    default: return nullptr;
    }
}

void fill_consts(fhip_ctx *ctx)
{
    const fhip_params &p = ctx->p;
    Consts &c = ctx->c;
    c.Lx = p.Lx; c.Ly = p.Ly; c.Lz = p.Lz;
    c.Mx = ctx->Mx; c.Mh = ctx->Mh; c.N = ctx->N; c.Kh = ctx->Kh;
    c.acoustic = p.mass_acou[0] > 0.0;
    c.efield = !p.E_field_always_zero;
    c.has_eloc = 0;
    c.modulation = 0;
    // film/epit mirror (mirror.F:18-33): without dead layers iz = Lz/2 .. Lz-1
    // copy Lz-1-iz, with them iz = Lz-(Lz+1)/2+1 .. Lz-1 copy Lz-iz
    c.mir_lo = p.Lz;
    c.mir_src = 0;
    if (p.bulk_or_film == FHIP_FILM || p.bulk_or_film == FHIP_EPIT) {
        c.mir_lo = p.gap_id == 0 ? p.Lz - p.Lz / 2 : p.Lz - (p.Lz + 1) / 2 + 1;
        c.mir_src = p.gap_id == 0 ? p.Lz - 1 : p.Lz;
    }
    c.Ninv = 1.0 / ctx->N;
    c.dt = p.dt;
    c.dt_2 = p.dt / 2;
    c.a0 = p.a0;
    c.Zstar = p.Z_star[0];
    c.Z2 = p.Z_star[0] * p.Z_star[0];
    for (int s = 0; s < 3; ++s) c.Zs[s] = p.Z_star[s];
    // species with different Z*: Z*(R) u gets its own transforms
    c.zsp = (p.Z_star[1] != p.Z_star[0] || p.Z_star[2] != p.Z_star[0]) ? 1 : 0;
    c.zu_slot = c.acoustic ? 12 : 9;
    for (int s = 0; s < 3; ++s) {
        c.kappa2[s] = p.P_kappa2[s];
        c.fa2g[s] = 4 * p.P_alpha[s] + 2 * p.P_gamma[s];
        c.tg[s] = 2 * p.P_gamma[s];
        c.sk[s] = 6 * p.P_k1[s] + 2 * p.P_k2[s];
        c.tk2[s] = 2 * p.P_k2[s];
        c.fk[s] = 4 * p.P_k2[s] - 2 * p.P_k3[s];
        c.k4x8[s] = 8 * p.P_k4[s];
        c.alpha[s] = p.P_alpha[s];
        c.gamma[s] = p.P_gamma[s];
        c.k1[s] = p.P_k1[s];
        c.k2[s] = p.P_k2[s];
        c.k3[s] = p.P_k3[s];
        c.k4[s] = p.P_k4[s];
        c.minv[s] = 1.0 / p.mass_dipo[s];
        c.mwinv[s] = 1.0 / p.mass_acou[s];
    }
    c.B11 = p.B11; c.B12 = p.B12; c.B44 = p.B44;
    c.B1xx = p.B1xx; c.B1yy = p.B1yy; c.B4yz = p.B4yz;
    c.mod_c = p.modulation_constant;
    for (int i = 0; i < 7; ++i) c.j[i] = p.j[i];
    // homogeneous_coupling_2 (elastic.F:55-71)
    for (int i = 0; i < 36; ++i) c.hc2[i] = 0.0;
    for (int i = 0; i < 3; ++i) c.hc2[i + 6 * i] = p.B1xx / 2;
    for (int i = 3; i < 6; ++i) c.hc2[i + 6 * i] = p.B4yz;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (i != j) c.hc2[i + 6 * j] = p.B1yy / 2;

    ScalarConsts &k = ctx->sk;
    k.geometry = p.bulk_or_film;
    k.method = p.method;
    k.modulation_energy = p.modulation_constant != 0.0;
    k.n_sum_p = ctx->N;
    k.N = ctx->N;
    k.Ninv = 1.0 / ctx->N;
    k.dt = p.dt;
    k.dt_2 = p.dt / 2;
    k.Q = p.Q_Nose;
    k.T_eV = p.kelvin / KELVIN_EV;
    k.n_freedom = (double)(3 * ctx->N);
    k.three_N = (double)(3 * ctx->N);
    k.target_ke = p.kelvin * 1.5 * ctx->N / KELVIN_EV;
    k.epi_strain = p.epi_strain;
    const double V = p.a0 * p.a0 * p.a0;
    for (int i = 0; i < 3; ++i) k.pNa03[i] = p.GPa[i] * 0.01 / ELEMENTARY_CHARGE * ctx->N * V;
    // homogeneous and its Cholesky factor (elastic.F:26-52)
    double *H = k.H, *G = k.G;
    for (int i = 0; i < 36; ++i) { H[i] = 0.0; G[i] = 0.0; }
    for (int i = 0; i < 3; ++i) H[i + 6 * i] = p.B11 * ctx->N;
    for (int i = 3; i < 6; ++i) H[i + 6 * i] = p.B44 * ctx->N;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (i != j) H[i + 6 * j] = p.B12 * ctx->N;
    for (int j = 0; j < 6; ++j) {
        double s = 0.0;
        for (int l = 0; l < j; ++l) s += G[j + 6 * l] * G[j + 6 * l];
        G[j + 6 * j] = std::sqrt(H[j + 6 * j] - s);
        k.Ginv[j] = 1.0 / G[j + 6 * j];
        for (int i = j + 1; i < 6; ++i) {
            double s2 = 0.0;
            for (int l = 0; l < j; ++l) s2 += G[i + 6 * l] * G[j + 6 * l];
            G[i + 6 * j] = (H[i + 6 * j] - s2) * k.Ginv[j];
        }
    }
    for (int i = 0; i < 36; ++i) k.hc2[i] = c.hc2[i];
}

template <typename T>
int upload_table(fhip_ctx *ctx, cpx<T> **dst, int n, bool twiddle)
{
    std::vector<cpx<T>> h(n);
    for (int j = 0; j < n; ++j) {
        const double a = 2.0 * M_PI_D * (double)j / (double)n;
        const double cs = std::cos(a), sn = twiddle ? -std::sin(a) : std::sin(a);
        h[j] = cpx<T>{(T)cs, (T)sn};
    }
    RC(dalloc(ctx, dst, n));
    HIPCHK(hipMemcpyAsync(*dst, h.data(), sizeof(cpx<T>) * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FHIP_OK;
}

// k components kcomp(i, L) for i < n along one axis.
int upload_kcomp(fhip_ctx *ctx, double **dst, int n, int L, bool fold)
{
    std::vector<double> h(n);
    for (int i = 0; i < n; ++i) h[i] = kcomp(i, L, ctx->p.a0, fold);
    RC(dalloc(ctx, dst, n));
    HIPCHK(hipMemcpyAsync(*dst, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FHIP_OK;
}

// Forward 3-D r2c, in FP64, of nf real device fields into the FP64 spectral
// slots 0..nf-1 (set-up transforms run in FP64 in every precision mode).
int forward_generic(fhip_ctx *ctx, const double *const *fields, int nf)
{
    XfwdArgs xa{};
    xa.mode = XF_GENERIC;
    xa.nf = nf;
    for (int i = 0; i < nf; ++i) xa.gen[i] = fields[i];
    RC(run_xfwd<double>(ctx, xa));
    RC(run_ypass<double>(ctx, -1, 0, nf));
    ZArgs<double> za = zargs<double>(ctx, false);
    za.nplain = nf;
    za.plain_sign = -1;
    return run_zplain<double>(ctx, za);
}

// FP64 spectral slots for set-up work in FP32 mode, released afterwards.
struct SetupSlots {
    fhip_ctx *ctx;
    void *raw = nullptr;
    int open(int n)
    {
        if (ctx->precision == FHIP_FP64) return FHIP_OK;
        hipError_t e = hipMalloc(&raw, sizeof(cpx<double>) * (size_t)n * ctx->Ks);
        if (e != hipSuccess) return fail(ctx, FHIP_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
        for (int q = 0; q < MAX_SLOTS; ++q)
            ctx->dset.sp.s[q] = q < n ? (cpx<double> *)raw + (size_t)q * ctx->Ks : nullptr;
        return FHIP_OK;
    }
    ~SetupSlots()
    {
        if (raw) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(raw);
            for (int q = 0; q < MAX_SLOTS; ++q) ctx->dset.sp.s[q] = nullptr;
        }
    }
};

int build_long_range(fhip_ctx *ctx)
{
    const fhip_params &p = ctx->p;
    double *tmp = nullptr;
    void *raw = nullptr;
    HIPCHK(hipMalloc(&raw, sizeof(double) * 6 * (size_t)ctx->N));
    tmp = (double *)raw;
    EwaldArgs e{};
    e.max_x = 24 / p.Lx; e.max_y = 24 / p.Ly; e.max_z = 24 / p.Lz;
    e.kappa = p.kappa > 0.0 ? p.kappa : std::sqrt(M_PI_D) / p.a0 / std::pow((double)ctx->N, 1.0 / 6);
    e.a0 = p.a0;
    e.Lx_a0 = p.Lx * p.a0; e.Ly_a0 = p.Ly * p.a0; e.Lz_a0 = p.Lz * p.a0;
    e.four_pi_over_v = 4 * M_PI_D / (p.a0 * p.a0 * p.a0);
    e.two_pi_over_a = 2 * M_PI_D / p.a0;
    e.four_kappa2 = 4 * (e.kappa * e.kappa);
    e.scale = HARTREE_BOHR / p.epsilon_inf;
    Consts c = ctx->c;
    hipStream_t s = ctx->stream;
    SetupSlots slots{ctx};
    RC(slots.open(6));
    int rc = launch(ctx, K_SETUP, [&] { k_ewald_real<<<ctx->kblocks, NT, 0, s>>>(c, e, tmp); });
    if (rc == FHIP_OK) {
        const double *flds[6];
        for (int m = 0; m < 6; ++m) flds[m] = tmp + (size_t)m * ctx->N;
        rc = forward_generic(ctx, flds, 6);
    }
    if (rc == FHIP_OK) {
        Spectra<double> sp = ctx->dset.sp;
        double *L = ctx->Ltab;
        float *Lf = ctx->fset.Lt;
        const int kb = (ctx->Kh + NT - 1) / NT;
        rc = launch(ctx, K_SETUP, [&] { k_ewald_recip<float><<<kb, NT, 0, s>>>(c, e, sp, L, Lf); });
    }
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    hipFree(raw);
    if (rc != FHIP_OK) return rc;
    if (e2 != hipSuccess) return fail(ctx, FHIP_ERR_HIP, std::string("ewald set-up: ") + hipGetErrorString(e2));
    return FHIP_OK;
}

int check_params(fhip_ctx *ctx, const fhip_params *p)
{
    if (!p) return fail(ctx, FHIP_ERR_ARG, "null params");
    if (p->Lx <= 0 || p->Ly <= 0 || p->Lz <= 0 || p->padding_y < 0)
        return fail(ctx, FHIP_ERR_ARG, "bad supercell size");
    if (p->Lx % 2 != 0) return fail(ctx, FHIP_ERR_UNSUPPORTED, "Lx must be even");
    if (!size_supported(p->Lx) || !size_supported(p->Ly) || !size_supported(p->Lz))
        return fail(ctx, FHIP_ERR_UNSUPPORTED,
                    "supercell edges must be one of " FHIP_SIZES_TEXT);
    if (p->method < 0 || p->method > 2) return fail(ctx, FHIP_ERR_ARG, "method must be md, vs or lf");
    if (p->bulk_or_film < 0 || p->bulk_or_film > 3) return fail(ctx, FHIP_ERR_ARG, "bad geometry");
    if (p->gap_id < 0 || p->gap_id > 2) return fail(ctx, FHIP_ERR_ARG, "gap_id must be 0, 1 or 2");
    if (p->bulk_or_film == FHIP_FILM || p->bulk_or_film == FHIP_EPIT) {
        if (p->Lz % 2 != p->gap_id % 2)       // param_module.F:475-477
            return fail(ctx, FHIP_ERR_ARG, "film or epit needs mod(Lz,2) == mod(gap_id,2)");
        if (p->Lz < 4) return fail(ctx, FHIP_ERR_ARG, "film or epit needs Lz >= 4");
    } else if (p->gap_id != 0) {
        return fail(ctx, FHIP_ERR_ARG, "gap_id must be 0 unless bulk_or_film is film or epit (param_module.F:470-474)");
    }
    if (p->precision != FHIP_FP64 && p->precision != FHIP_FP32)
        return fail(ctx, FHIP_ERR_ARG, "precision must be FHIP_FP64 or FHIP_FP32");
    if (p->method == FHIP_MD && p->mass_acou[0] > 0.0)
        return fail(ctx, FHIP_ERR_ARG, "method 'md' cannot be used with acoustic_mass_amu (param_module.F:238-242)");
    for (int s = 0; s < 3; ++s)
        if (!(p->mass_dipo[s] > 0.0)) return fail(ctx, FHIP_ERR_ARG, "mass_dipo must be positive");
    if (!(p->epsilon_inf > 0.0) || !(p->a0 > 0.0) || !(p->dt > 0.0))
        return fail(ctx, FHIP_ERR_ARG, "a0, dt and epsilon_inf must be positive");
    return FHIP_OK;
}

// Per-cell code byte from the species, sum_p and 1/m (read_relaxor.F,
// read_defects_u.F).  film/epit: the upper half of z is never counted and the
// dead layers (gap_id 1: iz = 0; 2: iz = 0 and Lz/2) are frozen and not
// counted, as read_defects_u.F:52-68 imposes after reading the .defects file.
int set_cells_impl(fhip_ctx *ctx, const int32_t *relaxor, const int32_t *sum_p, const double *minv)
{
    const fhip_params &p = ctx->p;
    const bool film = p.bulk_or_film == FHIP_FILM || p.bulk_or_film == FHIP_EPIT;
    std::vector<unsigned char> h(ctx->N);
    std::vector<int> layer(ctx->Lz, 0);
    int nsum = 0;
    for (int z = 0; z < ctx->Lz; ++z)
        for (int y = 0; y < ctx->Ly; ++y)
            for (int x = 0; x < ctx->Lx; ++x) {
                const size_t hi = padded_index(ctx, 0, 1, x, y, z);
                int s = relaxor ? relaxor[hi] - 1 : 0;
                if (s < 0 || s > 2) return fail(ctx, FHIP_ERR_ARG, "relaxor species must be 1..3");
                bool counted = sum_p ? sum_p[hi] != 0 : true;
                bool frozen = false;
                if (minv) {
                    const double m = minv[hi];
                    if (m == 0.0) frozen = true;
                    else if (m != 1.0 / p.mass_dipo[s])
                        return fail(ctx, FHIP_ERR_UNSUPPORTED, "mass_dipo_inv must be 1/mass_dipo(species) or 0");
                }
                if (film) {
                    if (z >= (p.Lz + 1) / 2) counted = false;
                    if ((p.gap_id > 0 && z == 0) || (p.gap_id == 2 && z == p.Lz / 2)) {
                        counted = false;
                        frozen = true;
                    }
                }
                h[x + (size_t)ctx->Lx * (y + (size_t)ctx->Ly * z)] =
                    (unsigned char)(s | (frozen ? 4 : 0) | (counted ? 8 : 0));
                nsum += counted ? 1 : 0;
                layer[z] += counted ? 1 : 0;
            }
    if (nsum == 0) return fail(ctx, FHIP_ERR_ARG, "no counted cell (sum_p all false)");
    HIPCHK(hipMemcpyAsync(ctx->code, h.data(), ctx->N, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));

    ctx->sk.n_sum_p = nsum;
    ctx->nlayer_host = layer;                      // read_defects_u.F:73-75
    if (ctx->nlayer) {
        HIPCHK(hipMemcpyAsync(ctx->nlayer, layer.data(), sizeof(int) * ctx->Lz, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return FHIP_OK;
}

// ---- the allocation-free force evaluation (fhip_force_static) ----------------
// Every intermediate of the call lives in one fixed device array: the packed
// spectra, the L(k) table in the library's k order, the FFT and k-vector
// tables, the per-workgroup partial sums, the scalars and the cell codes.
// Each device holds its own copy (loaded with the code object) until the
// process ends; calls on one device share it, so overlapping calls race.
constexpr size_t kStaticScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array: hipGetSymbolAddress once per
// device, with that device current.  The array's name is a host-side stand-in
// and is never passed to a kernel.
int scratch_address(void **out, std::string &err)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        err = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return FHIP_ERR_HIP;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return FHIP_ERR_HIP;
        }
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return FHIP_OK;
}

// Carves 256-byte-aligned pieces out of the scratch array (or only counts
// their size when base is null).
struct Carve {
    unsigned char *base;
    size_t used = 0;
    template <typename P>
    P *take(size_t count)
    {
        used = (used + 255) & ~(size_t)255;
        P *r = base ? reinterpret_cast<P *>(base + used) : nullptr;
        used += count * sizeof(P);
        return r;
    }
};

struct StaticTables {
    cpx<double> *twx, *twy, *twz, *csx, *csy, *csz;
    double *kvx, *kvy, *kvz;
    int Lx, Ly, Lz, Kx;
    double a0;
};

// FFT twiddles e^{-2 pi i j / n}, cos/sin tables e^{+2 pi i j / n} and the k
// components of each axis (the tables fhip_create uploads from the host).
static __global__ __launch_bounds__(NT) void k_static_tables(StaticTables t)
{
    const int j = blockIdx.x * NT + threadIdx.x;
    const double M2PI = 6.28318530717958647693;
    auto put = [&](int n, cpx<double> *tw, cpx<double> *cs) {
        if (j >= n) return;
        double sn, cs_;
        sincos(M2PI * (double)j / (double)n, &sn, &cs_);
        tw[j] = cpx<double>{cs_, -sn};
        cs[j] = cpx<double>{cs_, sn};
    };
    put(t.Lx, t.twx, t.csx);
    put(t.Ly, t.twy, t.csy);
    put(t.Lz, t.twz, t.csz);
    if (j < t.Kx) t.kvx[j] = kcomp(j, t.Lx, t.a0, false);
    if (j < t.Ly) t.kvy[j] = kcomp(j, t.Ly, t.a0, true);
    if (j < t.Lz) t.kvz[j] = kcomp(j, t.Lz, t.a0, true);
}

// Cell codes (species 1, counted), the device scalars of a fresh context
// (s = 1, E_add = the static external field), and L(k) from the caller's
// c%long_range layout (3,3,0:Lx/2,0:Ly+padding_y-1,0:Lz-1) into the library's
// k order, 6 upper-triangle words per point.
static __global__ __launch_bounds__(NT) void k_static_init(Consts c, unsigned char *code, Scalars *sc,
                                                           double Ex, double Ey, double Ez,
                                                           const double *__restrict__ lr, int padding_y,
                                                           double *__restrict__ Ltab)
{
    const size_t i = (size_t)blockIdx.x * NT + threadIdx.x;
    if (i < (size_t)c.N) code[i] = 8;
    if (i == 0) {
        Scalars z{};
        z.s = 1.0;
        z.i_step = 1.0;
        z.scale_p = 1.0;
        z.scale_pw = 1.0;
        z.E_add[0] = Ex;
        z.E_add[1] = Ey;
        z.E_add[2] = Ez;
        *sc = z;
    }
    const int Kx = c.Mx + 1;
    if (i < (size_t)Kx * c.Ly * c.Lz) {
        const int x = (int)(i % Kx);
        const int y = (int)((i / Kx) % c.Ly);
        const int z = (int)(i / ((size_t)Kx * c.Ly));
        const size_t k = kindex(c.Mx, c.Ly, c.Lz, c.Mh, x, y, z);
        const size_t base = 9 * (x + (size_t)Kx * (y + (size_t)(c.Ly + padding_y) * z));
        const int ia[6] = {0, 1, 2, 1, 2, 0}, ib[6] = {0, 1, 2, 2, 0, 1};
        for (int m = 0; m < 6; ++m) Ltab[k * 6 + m] = lr[base + ia[m] + 3 * ib[m]];
    }
}

static __global__ void k_static_out(const Scalars *__restrict__ sc, double *energy, double *strain)
{
    const int t = threadIdx.x;
    if (energy && t < 13) energy[t] = sc->en[t];
    if (strain && t < 6) strain[t] = sc->strain[t];
}

// ---- measurement: a plain device copy ---------------------------------------
// 16-byte words from src to dst, 4 per thread, each load and store
// instruction of a workgroup covering NT consecutive words.
static __global__ __launch_bounds__(NT) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                      size_t n)
{
    const size_t base = (size_t)blockIdx.x * 4 * NT + threadIdx.x;
    uint4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const size_t i = base + (size_t)k * NT;
        if (i < n) v[k] = src[i];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const size_t i = base + (size_t)k * NT;
        if (i < n) dst[i] = v[k];
    }
}

int static_fail(char *err, int errlen, int code, const std::string &msg)
{
    if (err && errlen > 0) {
        const size_t n = std::min((size_t)errlen - 1, msg.size());
        std::memcpy(err, msg.data(), n);
        err[n] = 0;
    }
    return code;
}

// ---- binary checkpoint -------------------------------------------------------
// Header, then the sections in this order: the device scalars; u, p, dV/du
// (FP64, 3N each); w, p_w, dV/dw with acoustic MD; the optimised w(k) when it
// is materialised; the averaging sums, histograms and dipoRavg when present;
// the heat-source layer energies when present.
constexpr char kCkptMagic[8] = {'F', 'H', 'I', 'P', 'C', 'K', 'P', 'T'};
constexpr uint32_t kCkptVersion = 1;
enum { CK_AVG = 1, CK_HIST = 2, CK_RAVG = 4, CK_ACOUK = 8, CK_LAYERS = 16, CK_AVG_AUTO = 32 };
struct CkptHeader {
    char magic[8];
    uint32_t version, header_bytes;
    int32_t Lx, Ly, Lz, method, precision, acoustic, geometry, flags;
    int64_t scalars_bytes, total_bytes;
};

struct CkptSection {
    void *dev;
    size_t bytes;
};

std::vector<CkptSection> ckpt_sections(fhip_ctx *ctx, int flags)
{
    std::vector<CkptSection> v;
    const size_t N = ctx->N;
    v.push_back({ctx->dsc, sizeof(Scalars)});
    // the state in the context's precision
    const bool f32 = ctx->precision == FHIP_FP32;
    const size_t sb = 3 * N * (f32 ? sizeof(float) : sizeof(double));
    v.push_back({f32 ? (void *)ctx->f32.u[0] : (void *)ctx->f.u[0], sb});
    v.push_back({f32 ? (void *)ctx->f32.p[0] : (void *)ctx->f.p[0], sb});
    v.push_back({f32 ? (void *)ctx->f32.F[0] : (void *)ctx->f.F[0], sb});
    if (ctx->c.acoustic) {
        v.push_back({f32 ? (void *)ctx->f32.w[0] : (void *)ctx->f.w[0], sb});
        v.push_back({f32 ? (void *)ctx->f32.pw[0] : (void *)ctx->f.pw[0], sb});
        v.push_back({f32 ? (void *)ctx->f32.dVdw[0] : (void *)ctx->f.dVdw[0], sb});
    }
    if (flags & CK_ACOUK) v.push_back({ctx->acouK, 3 * (size_t)ctx->Kh * sizeof(cpx<double>)});
    if (flags & CK_AVG) v.push_back({ctx->davg, sizeof(AvgAcc)});
    if (flags & CK_HIST) {
        v.push_back({ctx->dist1, 3 * (size_t)DIST_W * sizeof(unsigned long long)});
        v.push_back({ctx->dist3, (size_t)DIST_W * DIST_W * DIST_W * sizeof(int)});
    }
    if (flags & CK_RAVG) v.push_back({ctx->f.uavg[0], 3 * N * sizeof(double)});
    if (flags & CK_LAYERS) {
        v.push_back({ctx->el, 2 * (size_t)ctx->Lz * sizeof(double)});
        v.push_back({ctx->el_sum, 2 * (size_t)ctx->Lz * sizeof(double)});
    }
    return v;
}

int ckpt_flags(const fhip_ctx *ctx)
{
    int fl = 0;
    if (ctx->acouK_valid && !ctx->c.acoustic) fl |= CK_ACOUK;
    if (ctx->davg) fl |= CK_AVG;
    if (ctx->davg && ctx->avg_hist && ctx->dist1) fl |= CK_HIST;
    if (ctx->davg && ctx->avg_ravg && ctx->f.uavg[0]) fl |= CK_RAVG;
    if (ctx->el && ctx->el_sum) fl |= CK_LAYERS;
    if (ctx->avg_auto) fl |= CK_AVG_AUTO;
    return fl;
}

size_t ckpt_total(fhip_ctx *ctx, int flags)
{
    size_t t = sizeof(CkptHeader);
    for (const auto &sct : ckpt_sections(ctx, flags)) t += sct.bytes;
    return t;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

const char *fhip_last_error(const fhip_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int64_t fhip_device_bytes(const fhip_ctx *ctx) { return ctx ? ctx->bytes : 0; }

void fhip_destroy(fhip_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (void *v : ctx->allocs) hipFree(v);
    if (ctx->hsc) hipHostFree(ctx->hsc);
    for (auto &r : ctx->recs) {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    for (auto e : ctx->ev_pool) hipEventDestroy(e);
    drop_step_graph(ctx);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

int fhip_create(const fhip_params *p, fhip_ctx **out)
{
    if (!out) return FHIP_ERR_ARG;
    *out = nullptr;
    fhip_ctx *ctx = new fhip_ctx();
    int rc = check_params(ctx, p);
    if (rc != FHIP_OK) {
        // keep the context so the caller can read the message
        *out = ctx;
        return rc;
    }
    *out = ctx;
    ctx->p = *p;
    ctx->Lx = p->Lx; ctx->Ly = p->Ly; ctx->Lz = p->Lz;
    ctx->Kx = p->Lx / 2 + 1;
    ctx->Mx = p->Lx / 2;
    ctx->Mh = (size_t)ctx->Mx * p->Ly * p->Lz;
    ctx->N = p->Lx * p->Ly * p->Lz;
    ctx->Kh = ctx->Kx * p->Ly * p->Lz;
    ctx->Ks = ctx->Mh;
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    fill_consts(ctx);
    ctx->nslots = nforward(ctx);
    ctx->xblocks = (ctx->Ly * ctx->Lz + xinv_tile_rows(ctx->Lx) - 1) / xinv_tile_rows(ctx->Lx);
    ctx->xfblocks = (ctx->Ly * ctx->Lz + xfwd_tile_rows(ctx->Lx) - 1) / xfwd_tile_rows(ctx->Lx);
    for (int t = 0; t < 2; ++t) {
        const int ty = t ? line_tile_rows<float>(ctx->Ly) : line_tile_rows<double>(ctx->Ly);
        const int tz = t ? ztile_rows<float>(ctx->Lz) : ztile_rows<double>(ctx->Lz);
        ctx->yblocks[t] = (ctx->Mx * ctx->Lz + ty - 1) / ty;
        ctx->zblocks[t] = (ctx->Mx * ctx->Ly + tz - 1) / tz;
    }
    ctx->kblocks = (ctx->N + NT - 1) / NT;
    ctx->maxblocks = std::max(std::max(ctx->xblocks, std::max(ctx->zblocks[0], ctx->zblocks[1])), ctx->kblocks);
    {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, dev));
    }

    const size_t N = ctx->N, Kh = ctx->Kh;
    // u, p, pk, F (3 each); w, p_w, dV/dw only with acoustic MD; FP32 in
    // the FP32 mode
    const int nreal = ctx->c.acoustic ? 21 : 12;
    auto place = [&](auto &fs, auto *blk) {
        for (int q = 0; q < 3; ++q) {
            fs.u[q] = blk + (0 + q) * N;
            fs.p[q] = blk + (3 + q) * N;
            fs.pk[q] = blk + (6 + q) * N;
            fs.F[q] = blk + (9 + q) * N;
            if (ctx->c.acoustic) {
                fs.w[q] = blk + (12 + q) * N;
                fs.pw[q] = blk + (15 + q) * N;
                fs.dVdw[q] = blk + (18 + q) * N;
            }
        }
    };
    if (p->precision == FHIP_FP32) {
        float *blk = nullptr;
        RC(dalloc(ctx, &blk, nreal * N));
        place(ctx->f32, blk);
    } else {
        double *blk = nullptr;
        RC(dalloc(ctx, &blk, nreal * N));
        place(ctx->f, blk);
    }
    RC(dalloc(ctx, &ctx->code, N));
    RC(set_cells_impl(ctx, nullptr, nullptr, nullptr));
    ctx->f.code = ctx->code;
    ctx->precision = p->precision;
    if (ctx->precision == FHIP_FP32) {
        cpx<float> *spec = nullptr;
        RC(dalloc(ctx, &spec, (size_t)ctx->nslots * ctx->Ks));
        for (int s = 0; s < MAX_SLOTS; ++s) ctx->fset.sp.s[s] = (s < ctx->nslots) ? spec + (size_t)s * ctx->Ks : nullptr;
        RC(dalloc(ctx, &ctx->fset.Lt, 6 * Kh));
    } else {
        cpx<double> *spec = nullptr;
        RC(dalloc(ctx, &spec, (size_t)ctx->nslots * ctx->Ks));
        for (int s = 0; s < MAX_SLOTS; ++s) ctx->dset.sp.s[s] = (s < ctx->nslots) ? spec + (size_t)s * ctx->Ks : nullptr;
    }
    RC(dalloc(ctx, &ctx->Ltab, 6 * Kh));
    ctx->dset.Lt = ctx->Ltab;
    RC(dalloc(ctx, &ctx->acouK, 3 * Kh));
    RC(dalloc(ctx, &ctx->partials, (size_t)P_NSLOT * ctx->maxblocks));
    RC(dalloc(ctx, &ctx->dsc, 1));
    {
        void *h = nullptr;
        HIPCHK(hipHostMalloc(&h, sizeof(Scalars), 0));
        ctx->hsc = (Scalars *)h;
        std::memset(ctx->hsc, 0, sizeof(Scalars));
        ctx->hsc->s = 1.0;
        ctx->hsc->i_step = 1.0;         // molecular_dynamics counts steps from 1
        ctx->hsc->scale_p = 1.0;
        ctx->hsc->scale_pw = 1.0;
        for (int q = 0; q < 3; ++q) ctx->hsc->E_add[q] = p->external_E_field[q];
        ctx->hsc->e_mir = ctx->c.mir_lo < ctx->Lz ? 1.0 : 0.0;   // initialize_E_field.F:68-71
        RC(push_scalars(ctx));
    }
    RC(upload_table(ctx, &ctx->dset.twx, ctx->Lx, true));
    RC(upload_table(ctx, &ctx->dset.twy, ctx->Ly, true));
    RC(upload_table(ctx, &ctx->dset.twz, ctx->Lz, true));
    if (ctx->precision == FHIP_FP32) {
        RC(upload_table(ctx, &ctx->fset.twx, ctx->Lx, true));
        RC(upload_table(ctx, &ctx->fset.twy, ctx->Ly, true));
        RC(upload_table(ctx, &ctx->fset.twz, ctx->Lz, true));
    }
    RC(upload_table(ctx, &ctx->csx, ctx->Lx, false));
    RC(upload_table(ctx, &ctx->csy, ctx->Ly, false));
    RC(upload_table(ctx, &ctx->csz, ctx->Lz, false));
    RC(upload_kcomp(ctx, &ctx->kvx, ctx->Kx, ctx->Lx, false));
    RC(upload_kcomp(ctx, &ctx->kvy, ctx->Ly, ctx->Ly, true));
    RC(upload_kcomp(ctx, &ctx->kvz, ctx->Lz, ctx->Lz, true));
    RC(build_long_range(ctx));
    if (p->modulation_constant != 0.0) {
        // the reference stops when modulation_constant is set without a .modulation
        // file (elastic.F:226-233); the table is filled by fhip_set_modulation.
        RC(dalloc(ctx, &ctx->modk, Kh));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FHIP_OK;
}

int fhip_set_long_range(fhip_ctx *ctx, const double *lr)
{
    if (!ctx || !lr) return fail(ctx, FHIP_ERR_ARG, "null argument");
    const int Kx = ctx->Kx, Ly = ctx->Ly, Lz = ctx->Lz;
    const size_t dy = (size_t)Ly + ctx->p.padding_y;
    std::vector<double> h(6 * (size_t)ctx->Kh);
    const int ia[6] = {0, 1, 2, 1, 2, 0}, ib[6] = {0, 1, 2, 2, 0, 1};
    for (int z = 0; z < Lz; ++z)
        for (int y = 0; y < Ly; ++y)
            for (int x = 0; x < Kx; ++x) {
                const size_t k = kindex(ctx->Mx, Ly, Lz, ctx->Mh, x, y, z);
                const size_t base = 9 * (x + (size_t)Kx * (y + dy * z));
                for (int m = 0; m < 6; ++m) h[k * 6 + m] = lr[base + ia[m] + 3 * ib[m]];
            }
    HIPCHK(hipMemcpyAsync(ctx->Ltab, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, ctx->stream));
    std::vector<float> hf;
    if (ctx->precision == FHIP_FP32) {
        hf.resize(h.size());
        for (size_t q = 0; q < h.size(); ++q) hf[q] = (float)h[q];
        HIPCHK(hipMemcpyAsync(ctx->fset.Lt, hf.data(), sizeof(float) * hf.size(), hipMemcpyHostToDevice, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FHIP_OK;
}

int fhip_get_long_range(fhip_ctx *ctx, double *lr)
{
    if (!ctx || !lr) return fail(ctx, FHIP_ERR_ARG, "null argument");
    const int Kx = ctx->Kx, Ly = ctx->Ly, Lz = ctx->Lz;
    const size_t dy = (size_t)Ly + ctx->p.padding_y;
    std::vector<double> h(6 * (size_t)ctx->Kh);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->Ltab, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int ia[6] = {0, 1, 2, 1, 2, 0}, ib[6] = {0, 1, 2, 2, 0, 1};
    for (int z = 0; z < Lz; ++z)
        for (int y = 0; y < Ly; ++y)
            for (int x = 0; x < Kx; ++x) {
                const size_t k = kindex(ctx->Mx, Ly, Lz, ctx->Mh, x, y, z);
                const size_t base = 9 * (x + (size_t)Kx * (y + dy * z));
                for (int m = 0; m < 6; ++m) {
                    const double v = h[k * 6 + m];
                    lr[base + ia[m] + 3 * ib[m]] = v;
                    lr[base + ib[m] + 3 * ia[m]] = v;
                }
            }
    return FHIP_OK;
}

int fhip_set_modulation(fhip_ctx *ctx, const double *mod)
{
    if (!ctx || !mod) return fail(ctx, FHIP_ERR_ARG, "null argument");
    if (ctx->p.modulation_constant == 0.0)
        return fail(ctx, FHIP_ERR_ARG, "There is a .modulation file, but modulation_constant is zero (elastic.F:218-222)");
    std::vector<double> h(ctx->N);
    for (int z = 0; z < ctx->Lz; ++z)
        for (int y = 0; y < ctx->Ly; ++y)
            for (int x = 0; x < ctx->Lx; ++x)
                h[x + (size_t)ctx->Lx * (y + (size_t)ctx->Ly * z)] = mod[padded_index(ctx, 0, 1, x, y, z)];
    // the field goes through a temporary device buffer, so the state (dV/du,
    // which the next call's first kick reads, included) is left as it is
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(double) * ctx->N));
    struct FreeOnExit {
        double *p;
        ~FreeOnExit() { (void)hipFree(p); }
    } scratch_guard{d};
    HIPCHK(hipMemcpyAsync(d, h.data(), sizeof(double) * ctx->N, hipMemcpyHostToDevice, ctx->stream));
    const double *flds[1] = {d};
    {
        SetupSlots slots{ctx};
        RC(slots.open(1));
        RC(forward_generic(ctx, flds, 1));
        Consts c = ctx->c;
        const cpx<double> *s0 = ctx->dset.sp.s[0];
        cpx<double> *mk = ctx->modk;
        hipStream_t st = ctx->stream;
        const int kb = (ctx->Kh + NT - 1) / NT;
        RC(launch(ctx, K_SETUP, [&] { k_unpack<double><<<kb, NT, 0, st>>>(c, s0, mk); }));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->c.modulation = 1;
    return collect_events(ctx);
}

int fhip_set_cells(fhip_ctx *ctx, const int32_t *relaxor, const int32_t *sum_p, const double *minv)
{
    if (!ctx) return FHIP_ERR_ARG;
    return set_cells_impl(ctx, relaxor, sum_p, minv);
}


static int put_impl(fhip_ctx *ctx, int field, const double *host, bool soa)
{
    if (!ctx || !host) return fail(ctx, FHIP_ERR_ARG, "null argument");
    if (field == FHIP_DVDDI || field == FHIP_DVDAC || field == FHIP_ACOU_R_OPT)
        return fail(ctx, FHIP_ERR_ARG, "field is read only");
    if (field == FHIP_DIPO_R_AVG && !ctx->f.uavg[0]) RC(alloc_uavg(ctx));
    if (field == FHIP_E_LOCAL && !ctx->f.eloc[0]) {
        double *blk = nullptr;
        RC(dalloc(ctx, &blk, 3 * (size_t)ctx->N));
        for (int q = 0; q < 3; ++q) ctx->f.eloc[q] = blk + q * (size_t)ctx->N;
    }
    const bool f32 = ctx->precision == FHIP_FP32 && state_field(field);
    double **d = field_ptrs(ctx, field);
    float **d32 = f32 ? field_ptrs32(ctx, field) : nullptr;
    if (f32 ? (!d32 || !d32[0]) : (!d || !d[0])) return fail(ctx, FHIP_ERR_ARG, "field not present in this configuration");
    const size_t N = ctx->N;
    std::vector<double> h(3 * N);
    if (soa) {
        std::memcpy(h.data(), host, sizeof(double) * 3 * N);
    } else {
        for (int z = 0; z < ctx->Lz; ++z)
            for (int y = 0; y < ctx->Ly; ++y)
                for (int x = 0; x < ctx->Lx; ++x)
                    for (int q = 0; q < 3; ++q)
                        h[q * N + x + (size_t)ctx->Lx * (y + (size_t)ctx->Ly * z)] = host[padded_index(ctx, q, 3, x, y, z)];
    }
    if (f32) {
        const std::vector<float> hf(h.begin(), h.end());          // rounded to nearest
        for (int q = 0; q < 3; ++q)
            HIPCHK(hipMemcpyAsync(d32[q], hf.data() + q * N, sizeof(float) * N, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    } else {
        for (int q = 0; q < 3; ++q)
            HIPCHK(hipMemcpyAsync(d[q], h.data() + q * N, sizeof(double) * N, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    if (field == FHIP_E_LOCAL) ctx->c.has_eloc = 1;
    return FHIP_OK;
}

static int reconstruct_w(fhip_ctx *ctx, std::vector<double> &h);

static int get_impl(fhip_ctx *ctx, int field, double *host, bool soa)
{
    if (!ctx || !host) return fail(ctx, FHIP_ERR_ARG, "null argument");
    const size_t N = ctx->N;
    std::vector<double> h(3 * N);
    if (field == FHIP_ACOU_R_OPT) {
        RC(reconstruct_w(ctx, h));
    } else if (ctx->precision == FHIP_FP32 && state_field(field)) {
        float **d = field_ptrs32(ctx, field);
        if (!d || !d[0]) return fail(ctx, FHIP_ERR_ARG, "field not present in this configuration");
        std::vector<float> hf(3 * N);
        for (int q = 0; q < 3; ++q)
            HIPCHK(hipMemcpyAsync(hf.data() + q * N, d[q], sizeof(float) * N, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i < 3 * N; ++i) h[i] = hf[i];
    } else {
        double **d = field_ptrs(ctx, field);
        if (!d || !d[0]) return fail(ctx, FHIP_ERR_ARG, "field not present in this configuration");
        for (int q = 0; q < 3; ++q)
            HIPCHK(hipMemcpyAsync(h.data() + q * N, d[q], sizeof(double) * N, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    if (soa) {
        std::memcpy(host, h.data(), sizeof(double) * 3 * N);
    } else {
        for (int z = 0; z < ctx->Lz; ++z)
            for (int y = 0; y < ctx->Ly; ++y)
                for (int x = 0; x < ctx->Lx; ++x)
                    for (int q = 0; q < 3; ++q)
                        host[padded_index(ctx, q, 3, x, y, z)] = h[q * N + x + (size_t)ctx->Lx * (y + (size_t)ctx->Ly * z)];
    }
    return FHIP_OK;
}

int fhip_put(fhip_ctx *ctx, int field, const double *host) { return put_impl(ctx, field, host, false); }
int fhip_get(fhip_ctx *ctx, int field, double *host) { return get_impl(ctx, field, host, false); }
int fhip_put_soa(fhip_ctx *ctx, int field, const double *host) { return put_impl(ctx, field, host, true); }
int fhip_get_soa(fhip_ctx *ctx, int field, double *host) { return get_impl(ctx, field, host, true); }

int fhip_get_nose(fhip_ctx *ctx, double *s, double *pi, double *H0)
{
    if (!ctx) return FHIP_ERR_ARG;
    RC(pull_scalars(ctx));
    if (s) *s = ctx->hsc->s;
    if (pi) *pi = ctx->hsc->pi;
    if (H0) *H0 = ctx->hsc->H0;
    return FHIP_OK;
}

int fhip_set_nose(fhip_ctx *ctx, double s, double pi, double H0)
{
    if (!ctx) return FHIP_ERR_ARG;
    if (!(s > 0.0)) return fail(ctx, FHIP_ERR_ARG, "s_Nose must be positive");
    RC(pull_scalars(ctx));
    ctx->hsc->s = s;
    ctx->hsc->pi = pi;
    ctx->hsc->H0 = H0;
    return push_scalars(ctx);
}

int fhip_get_strain(fhip_ctx *ctx, double strain[6])
{
    if (!ctx || !strain) return FHIP_ERR_ARG;
    RC(pull_scalars(ctx));
    for (int i = 0; i < 6; ++i) strain[i] = ctx->hsc->strain[i];
    return FHIP_OK;
}

int fhip_set_E_field_add(fhip_ctx *ctx, const double E[3])
{
    if (!ctx || !E) return FHIP_ERR_ARG;
    RC(pull_scalars(ctx));
    for (int q = 0; q < 3; ++q) ctx->hsc->E_add[q] = E[q];
    ctx->hsc->e_mir = 0.0;       // per-step rewrite of E_field, not mirrored (molecular-dynamics.F:66-76)
    return push_scalars(ctx);
}

int fhip_set_E_schedule(fhip_ctx *ctx, int type, const double amplitude[3], int period, int n_thermalize,
                        int n_average, int64_t i_step)
{
    if (!ctx || !amplitude) return FHIP_ERR_ARG;
    if (type < FHIP_E_CONSTANT || type > FHIP_E_HYSTERESIS) return fail(ctx, FHIP_ERR_ARG, "unknown E schedule");
    if ((type == FHIP_E_TRIANGULAR_COS || type == FHIP_E_TRIANGULAR_SIN) && period < 1)
        return fail(ctx, FHIP_ERR_ARG, "triangular waves need n_E_wave_period >= 1");
    if ((type == FHIP_E_RAMPING_ON || type == FHIP_E_RAMPING_OFF) && (period < 1 || period > n_thermalize))
        return fail(ctx, FHIP_ERR_ARG, "period must be smaller than n_thermalize (wave_module.F:41-44)");
    if (type == FHIP_E_HYSTERESIS && n_average < 1) return fail(ctx, FHIP_ERR_ARG, "n_average must be >= 1");
    if (i_step < 0) return fail(ctx, FHIP_ERR_ARG, "i_step must be >= 0");
    if (!ctx->c.efield && (amplitude[0] != 0.0 || amplitude[1] != 0.0 || amplitude[2] != 0.0))
        return fail(ctx, FHIP_ERR_STATE, "the context was created with E_field_always_zero");
    RC(pull_scalars(ctx));
    Scalars &h = *ctx->hsc;
    for (int q = 0; q < 3; ++q) h.e_amp[q] = amplitude[q];
    h.e_type = type;
    h.e_period = period;
    h.e_ntherm = n_thermalize;
    h.e_navg = n_average;
    h.i_step = (double)i_step;
    h.e_mir = ctx->c.mir_lo < ctx->Lz ? 1.0 : 0.0;   // the field of initialize_E_field, mirrored
    if (type == FHIP_E_CONSTANT)
        for (int q = 0; q < 3; ++q) h.E_add[q] = amplitude[q];
    else
        e_field_next(&h);
    return push_scalars(ctx);
}

int fhip_get_step(fhip_ctx *ctx, int64_t *i_step, double E_add[3])
{
    if (!ctx) return FHIP_ERR_ARG;
    RC(pull_scalars(ctx));
    if (i_step) *i_step = (int64_t)ctx->hsc->i_step;
    if (E_add)
        for (int q = 0; q < 3; ++q) E_add[q] = ctx->hsc->E_add[q];
    return FHIP_OK;
}

int fhip_set_heat_source(fhip_ctx *ctx, double delta_e, double delta_e_acou)
{
    if (!ctx) return FHIP_ERR_ARG;
    if (ctx->p.method != FHIP_LF)
        return fail(ctx, FHIP_ERR_STATE, "heat-source layers act in leapfrog steps (leapfrog.F:38-83)");
    if (delta_e_acou != 0.0 && !ctx->c.acoustic)
        return fail(ctx, FHIP_ERR_ARG, "delta_e_acou needs acoustic MD (mass_acou > 0)");
    if (!ctx->el) {
        RC(dalloc(ctx, &ctx->el, 2 * (size_t)ctx->Lz));
        RC(dalloc(ctx, &ctx->el_sum, 2 * (size_t)ctx->Lz));
        RC(dalloc(ctx, &ctx->nlayer, (size_t)ctx->Lz));
        if ((int)ctx->nlayer_host.size() != ctx->Lz) ctx->nlayer_host.assign(ctx->Lz, ctx->Lx * ctx->Ly);
        HIPCHK(hipMemcpyAsync(ctx->nlayer, ctx->nlayer_host.data(), sizeof(int) * ctx->Lz, hipMemcpyHostToDevice,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    for (int z = 0; z < ctx->Lz; ++z)
        if (ctx->nlayer_host[z] == 0 && delta_e != 0.0)
            return fail(ctx, FHIP_ERR_ARG, "a z layer without counted cells (sum_p)");
    ctx->delta_e = delta_e;
    ctx->delta_e_acou = delta_e_acou;
    return FHIP_OK;
}

int fhip_get_layer_energies(fhip_ctx *ctx, double *e_dipo, double *e_acou, int averaged)
{
    if (!ctx) return FHIP_ERR_ARG;
    if (!ctx->el) return fail(ctx, FHIP_ERR_STATE, "no heat-source layers (fhip_set_heat_source)");
    std::vector<double> h(2 * (size_t)ctx->Lz);
    HIPCHK(hipMemcpyAsync(h.data(), averaged ? ctx->el_sum : ctx->el, sizeof(double) * h.size(),
                          hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (e_dipo) std::memcpy(e_dipo, h.data(), sizeof(double) * ctx->Lz);
    if (e_acou) std::memcpy(e_acou, h.data() + ctx->Lz, sizeof(double) * ctx->Lz);
    return FHIP_OK;
}

int fhip_set_kelvin(fhip_ctx *ctx, double kelvin)
{
    if (!ctx) return FHIP_ERR_ARG;
    if (!(kelvin >= 0.0)) return fail(ctx, FHIP_ERR_ARG, "kelvin must be >= 0");
    ctx->p.kelvin = kelvin;
    ctx->sk.T_eV = kelvin / KELVIN_EV;                       // param_module.F:226
    ctx->sk.target_ke = kelvin * 1.5 * ctx->N / KELVIN_EV;   // param_module.F:468
    return FHIP_OK;
}

int fhip_force(fhip_ctx *ctx, fhip_energy *e)
{
    if (!ctx) return FHIP_ERR_ARG;
    RC(force_pipeline(ctx, SK_FORCE, true));
    RC(pull_scalars(ctx));
    energies_out(ctx, e);
    return FHIP_OK;
}

int fhip_force_static(const fhip_params *p, const double *long_range, const double *u, double *dVddi,
                      fhip_energy *energy, double *strain, void *stream, char *err, int errlen)
{
    if (!p || !long_range || !u || !dVddi || !energy)
        return static_fail(err, errlen, FHIP_ERR_ARG, "null argument");
    // a context record on the stack: its device pointers all point into the
    // scratch array, and nothing on it allocates device memory
    fhip_ctx ctx;
    int rc = check_params(&ctx, p);
    if (rc != FHIP_OK) return static_fail(err, errlen, rc, ctx.err);
    if (p->precision != FHIP_FP64 || p->mass_acou[0] > 0.0 || p->modulation_constant != 0.0 ||
        (p->bulk_or_film != FHIP_BULK && p->bulk_or_film != FHIP_E110) || p->Z_star[1] != p->Z_star[0] ||
        p->Z_star[2] != p->Z_star[0])
        return static_fail(err, errlen, FHIP_ERR_UNSUPPORTED,
                           "fhip_force_static: FP64, bulk or e110, optimised acoustic strain, no modulation, "
                           "one Z* only");
    ctx.p = *p;
    ctx.Lx = p->Lx; ctx.Ly = p->Ly; ctx.Lz = p->Lz;
    ctx.Kx = p->Lx / 2 + 1;
    ctx.Mx = p->Lx / 2;
    ctx.Mh = (size_t)ctx.Mx * p->Ly * p->Lz;
    ctx.N = p->Lx * p->Ly * p->Lz;
    ctx.Kh = ctx.Kx * p->Ly * p->Lz;
    ctx.Ks = ctx.Mh;
    ctx.stream = (hipStream_t)stream;
    fill_consts(&ctx);
    ctx.nslots = 9;
    ctx.xblocks = (ctx.Ly * ctx.Lz + xinv_tile_rows(ctx.Lx) - 1) / xinv_tile_rows(ctx.Lx);
    ctx.xfblocks = (ctx.Ly * ctx.Lz + xfwd_tile_rows(ctx.Lx) - 1) / xfwd_tile_rows(ctx.Lx);
    ctx.yblocks[0] = (ctx.Mx * ctx.Lz + line_tile_rows<double>(ctx.Ly) - 1) / line_tile_rows<double>(ctx.Ly);
    ctx.zblocks[0] = (ctx.Mx * ctx.Ly + ztile_rows<double>(ctx.Lz) - 1) / ztile_rows<double>(ctx.Lz);
    ctx.kblocks = (ctx.N + NT - 1) / NT;
    ctx.maxblocks = std::max(std::max(ctx.xblocks, ctx.zblocks[0]), ctx.kblocks);
    ctx.precision = FHIP_FP64;

    // the pieces of the scratch array, counted first
    StaticTables t{};
    cpx<double> *spec = nullptr;
    auto carve = [&](Carve &cv) {
        spec = cv.take<cpx<double>>(9 * ctx.Ks);
        ctx.Ltab = cv.take<double>(6 * (size_t)ctx.Kh);
        ctx.partials = cv.take<double>((size_t)P_NSLOT * ctx.maxblocks);
        ctx.dsc = cv.take<Scalars>(1);
        ctx.code = cv.take<unsigned char>(ctx.N);
        t.twx = cv.take<cpx<double>>(ctx.Lx);
        t.twy = cv.take<cpx<double>>(ctx.Ly);
        t.twz = cv.take<cpx<double>>(ctx.Lz);
        t.csx = cv.take<cpx<double>>(ctx.Lx);
        t.csy = cv.take<cpx<double>>(ctx.Ly);
        t.csz = cv.take<cpx<double>>(ctx.Lz);
        t.kvx = cv.take<double>(ctx.Kx);
        t.kvy = cv.take<double>(ctx.Ly);
        t.kvz = cv.take<double>(ctx.Lz);
    };
    Carve count{nullptr};
    carve(count);
    if (count.used > kStaticScratchBytes)
        return static_fail(err, errlen, FHIP_ERR_UNSUPPORTED,
                           "fhip_force_static: the supercell needs " + std::to_string(count.used) +
                               " bytes of scratch, more than the " + std::to_string(kStaticScratchBytes) +
                               " the library holds per device (N = Lx*Ly*Lz up to about 110592)");
    void *base = nullptr;
    std::string msg;
    rc = scratch_address(&base, msg);
    if (rc != FHIP_OK) return static_fail(err, errlen, rc, msg);
    Carve cv{(unsigned char *)base};
    carve(cv);
    for (int q = 0; q < MAX_SLOTS; ++q) ctx.dset.sp.s[q] = q < 9 ? spec + (size_t)q * ctx.Ks : nullptr;
    ctx.dset.Lt = ctx.Ltab;
    ctx.dset.twx = t.twx; ctx.dset.twy = t.twy; ctx.dset.twz = t.twz;
    ctx.csx = t.csx; ctx.csy = t.csy; ctx.csz = t.csz;
    ctx.kvx = t.kvx; ctx.kvy = t.kvy; ctx.kvz = t.kvz;
    ctx.f.code = ctx.code;
    const size_t N = ctx.N;
    for (int q = 0; q < 3; ++q) {
        ctx.f.u[q] = const_cast<double *>(u) + q * N;
        ctx.f.F[q] = dVddi + q * N;
    }
    t.Lx = ctx.Lx; t.Ly = ctx.Ly; t.Lz = ctx.Lz; t.Kx = ctx.Kx; t.a0 = p->a0;
    hipStream_t s = ctx.stream;
    const int lmax = std::max(std::max(ctx.Lx, ctx.Ly), ctx.Lz);
    k_static_tables<<<(lmax + NT - 1) / NT, NT, 0, s>>>(t);
    const size_t ninit = std::max(N, (size_t)ctx.Kh);
    const double *E = p->E_field_always_zero ? nullptr : p->external_E_field;
    k_static_init<<<(unsigned)((ninit + NT - 1) / NT), NT, 0, s>>>(ctx.c, ctx.code, ctx.dsc, E ? E[0] : 0.0,
                                                                 E ? E[1] : 0.0, E ? E[2] : 0.0, long_range,
                                                                 p->padding_y, ctx.Ltab);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return static_fail(err, errlen, FHIP_ERR_HIP, std::string("set-up launch: ") + hipGetErrorString(e));
    rc = force_pipeline(&ctx, SK_FORCE, true);
    if (rc != FHIP_OK) return static_fail(err, errlen, rc, ctx.err);
    k_static_out<<<1, 64, 0, s>>>(ctx.dsc, (double *)energy, strain);
    e = hipGetLastError();
    if (e != hipSuccess) return static_fail(err, errlen, FHIP_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
    return FHIP_OK;
}

int fhip_copy_rate(int64_t bytes, int reps, double *gbs, double *us_per_copy)
{
    if (bytes < 16 || reps < 1 || !gbs) return FHIP_ERR_ARG;
    const size_t n = (size_t)bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FHIP_ERR_HIP;
    if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
        rc = FHIP_ERR_ALLOC;
    } else if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
               hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
               hipMemsetAsync(a, 0x3c, n * 16, st) == hipSuccess) {
        const unsigned blocks = (unsigned)((n + 4 * (size_t)NT - 1) / (4 * (size_t)NT));
        bool ok = true;
        for (int i = 0; i < 2 + reps && ok; ++i) {
            if (i == 2) ok = hipEventRecord(e0, st) == hipSuccess;
            k_copy16<<<blocks, NT, 0, st>>>((const uint4 *)a, (uint4 *)b, n);
            ok = ok && hipGetLastError() == hipSuccess;
        }
        float ms = 0.0f;
        if (ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
            hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0.0f) {
            const double t = 1e-3 * ms / reps;
            *gbs = 2.0 * 16.0 * (double)n / t / 1e9;
            if (us_per_copy) *us_per_copy = 1e6 * t;
            rc = FHIP_OK;
        }
    }
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int64_t fhip_checkpoint_bytes(fhip_ctx *ctx)
{
    if (!ctx) return FHIP_ERR_ARG;
    return (int64_t)ckpt_total(ctx, ckpt_flags(ctx));
}

int fhip_checkpoint_save(fhip_ctx *ctx, void *buf, int64_t bytes)
{
    if (!ctx || !buf) return fail(ctx, FHIP_ERR_ARG, "null argument");
    const int flags = ckpt_flags(ctx);
    const size_t total = ckpt_total(ctx, flags);
    if (bytes < (int64_t)total) return fail(ctx, FHIP_ERR_ARG, "checkpoint buffer too small (fhip_checkpoint_bytes)");
    CkptHeader h{};
    std::memcpy(h.magic, kCkptMagic, sizeof(h.magic));
    h.version = kCkptVersion;
    h.header_bytes = sizeof(CkptHeader);
    h.Lx = ctx->Lx; h.Ly = ctx->Ly; h.Lz = ctx->Lz;
    h.method = ctx->p.method; h.precision = ctx->precision; h.acoustic = ctx->c.acoustic;
    h.geometry = ctx->p.bulk_or_film; h.flags = flags;
    h.scalars_bytes = sizeof(Scalars);
    h.total_bytes = (int64_t)total;
    unsigned char *out = (unsigned char *)buf;
    std::memcpy(out, &h, sizeof(h));
    size_t off = sizeof(h);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (const auto &sct : ckpt_sections(ctx, flags)) {
        HIPCHK(hipMemcpy(out + off, sct.dev, sct.bytes, hipMemcpyDeviceToHost));
        off += sct.bytes;
    }
    return FHIP_OK;
}

int fhip_checkpoint_load(fhip_ctx *ctx, const void *buf, int64_t bytes)
{
    if (!ctx || !buf) return fail(ctx, FHIP_ERR_ARG, "null argument");
    if (bytes < (int64_t)sizeof(CkptHeader)) return fail(ctx, FHIP_ERR_ARG, "checkpoint too short");
    CkptHeader h;
    std::memcpy(&h, buf, sizeof(h));
    if (std::memcmp(h.magic, kCkptMagic, sizeof(h.magic)) != 0 || h.version != kCkptVersion ||
        h.header_bytes != sizeof(CkptHeader) || h.scalars_bytes != (int64_t)sizeof(Scalars))
        return fail(ctx, FHIP_ERR_ARG, "not a checkpoint of this library version");
    if (h.Lx != ctx->Lx || h.Ly != ctx->Ly || h.Lz != ctx->Lz || h.method != ctx->p.method ||
        h.precision != ctx->precision || h.acoustic != ctx->c.acoustic || h.geometry != ctx->p.bulk_or_film)
        return fail(ctx, FHIP_ERR_ARG, "checkpoint of another supercell, method, precision or geometry");
    if ((h.flags & CK_LAYERS) && !(ctx->el && ctx->el_sum))
        return fail(ctx, FHIP_ERR_STATE, "checkpoint holds heat-source layers: call fhip_set_heat_source first");
    // averaging buffers as the saving context had them
    if (h.flags & CK_AVG) {
        const int opts = ((h.flags & CK_HIST) ? FHIP_AVG_HISTOGRAMS : 0) | ((h.flags & CK_RAVG) ? FHIP_AVG_DIPO_R_AVG : 0);
        RC(fhip_avg_reset(ctx, opts));
    }
    const size_t total = ckpt_total(ctx, h.flags & ~CK_AVG_AUTO);
    if (h.total_bytes != (int64_t)total || bytes < (int64_t)total)
        return fail(ctx, FHIP_ERR_ARG, "checkpoint size does not match its sections");
    const unsigned char *in = (const unsigned char *)buf;
    size_t off = sizeof(h);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (const auto &sct : ckpt_sections(ctx, h.flags)) {
        HIPCHK(hipMemcpy(sct.dev, in + off, sct.bytes, hipMemcpyHostToDevice));
        off += sct.bytes;
    }
    std::memcpy(ctx->hsc, in + sizeof(h), sizeof(Scalars));
    ctx->acouK_valid = (h.flags & CK_ACOUK) != 0;
    ctx->avg_auto = (h.flags & CK_AVG_AUTO) != 0 && ctx->davg;
    return FHIP_OK;
}

int fhip_md_prologue(fhip_ctx *ctx, fhip_energy *e)
{
    if (!ctx) return FHIP_ERR_ARG;
    RC(force_pipeline(ctx, SK_FORCE, true));
    RC(launch_kick(ctx, false));
    {
        ScalarConsts k = ctx->sk;
        Scalars *sc = ctx->dsc;
        double *part = ctx->partials;
        int mb = ctx->maxblocks, kb = ctx->kblocks;
        hipStream_t s = ctx->stream;
        RC(launch(ctx, K_KE, [&] { k_ke_report<<<1, NT, 0, s>>>(k, sc, part, mb, kb); }));
    }
    RC(pull_scalars(ctx));
    // initialize-dipoR.F:64-68 and molecular-dynamics.F:44-55, in FP64 on the host
    Scalars &h = *ctx->hsc;
    h.s = 1.0;
    h.pi = 0.0;
    double *en = h.en;
    en[E_HNP] = 0.0;
    const double V = en[E_SR] + en[E_LR] + en[E_DE] + en[E_UNH] + en[E_HS] + en[E_HC] + en[E_IS] + en[E_IC] + en[E_IM];
    en[E_TOTAL] = en[E_DIPO_KIN] / (h.s * h.s) + en[E_ACOU_KIN] + V;
    auto np_h = [&]() {
        const double H = en[E_DIPO_KIN] / (h.s * h.s) + V + h.pi * h.pi / 2 / ctx->sk.Q
                       + ctx->sk.three_N * ctx->sk.T_eV * std::log(h.s) - h.H0;
        return H * h.s;
    };
    h.H0 = 0.0;
    en[E_HNP] = np_h();
    h.H0 = en[E_HNP] / h.s;
    en[E_HNP] = np_h();
    RC(push_scalars(ctx));
    energies_out(ctx, e);
    return FHIP_OK;
}

int fhip_nose_poincare(fhip_ctx *ctx, int nsteps, fhip_energy *e)
{
    if (!ctx) return FHIP_ERR_ARG;
    if (nsteps < 1) return fail(ctx, FHIP_ERR_ARG, "nsteps must be >= 1");
    if (ctx->p.method != FHIP_MD) return fail(ctx, FHIP_ERR_STATE, "context was created for another method");
    // first half kick and eq. 38 of the first step; later steps get theirs from
    // the fused kick in xinv and the scalar tail of k_finish
    RC(launch_kick(ctx, true));
    {
        ScalarConsts k = ctx->sk;
        Scalars *sc = ctx->dsc;
        double *part = ctx->partials;
        int mb = ctx->maxblocks, kb = ctx->kblocks;
        hipStream_t s = ctx->stream;
        RC(launch(ctx, K_NPBEGIN, [&] { k_np_begin<<<1, NT, 0, s>>>(k, sc, part, mb, kb); }));
    }
    RC(middle_steps(ctx, SK_NP, 1, nsteps - 1));
    RC(one_step(ctx, SK_NP, true, 0));
    RC(pull_scalars(ctx));
    energies_out(ctx, e);
    return FHIP_OK;
}

int fhip_leapfrog(fhip_ctx *ctx, int i_step, int nsteps, fhip_energy *e)
{
    if (!ctx) return FHIP_ERR_ARG;
    if (nsteps < 1) return fail(ctx, FHIP_ERR_ARG, "nsteps must be >= 1");
    if (ctx->p.method != FHIP_LF) return fail(ctx, FHIP_ERR_STATE, "context was created for another method");
    if (i_step == 0)
        return fail(ctx, FHIP_ERR_UNSUPPORTED,
                    "the i_step == 0 half kick is not reachable from molecular_dynamics (molecular-dynamics.F:64)");
    RC(middle_steps(ctx, SK_LF, 0, nsteps - 1));
    RC(one_step(ctx, SK_LF, true, 0));
    RC(pull_scalars(ctx));
    energies_out(ctx, e);
    return FHIP_OK;
}

int fhip_velocity_scaling(fhip_ctx *ctx, int nsteps, fhip_energy *e)
{
    if (!ctx) return FHIP_ERR_ARG;
    if (nsteps < 1) return fail(ctx, FHIP_ERR_ARG, "nsteps must be >= 1");
    if (ctx->p.method != FHIP_VS) return fail(ctx, FHIP_ERR_STATE, "context was created for another method");
    RC(middle_steps(ctx, SK_VS, 0, nsteps - 1));
    RC(one_step(ctx, SK_VS, true, 0));
    {
        Consts c = ctx->c;
        Scalars *sc = ctx->dsc;
        int kb = ctx->kblocks;
        hipStream_t s = ctx->stream;
        if (ctx->precision == FHIP_FP32) {
            FieldsT<float> f = fields<float>(ctx);
            RC(launch(ctx, K_SCALE, [&] { k_apply_scale<float><<<kb, NT, 0, s>>>(c, f, sc); }));
        } else {
            Fields f = ctx->f;
            RC(launch(ctx, K_SCALE, [&] { k_apply_scale<double><<<kb, NT, 0, s>>>(c, f, sc); }));
        }
        RC(launch(ctx, K_SCALE, [&] { k_reset_scale<<<1, 1, 0, s>>>(sc); }));
    }
    RC(pull_scalars(ctx));
    energies_out(ctx, e);
    return FHIP_OK;
}

int fhip_avg_reset(fhip_ctx *ctx, int options)
{
    if (!ctx) return FHIP_ERR_ARG;
    const bool histograms = (options & FHIP_AVG_HISTOGRAMS) != 0;
    const bool ravg = (options & FHIP_AVG_DIPO_R_AVG) != 0;
    if (!ctx->davg) RC(dalloc(ctx, &ctx->davg, 1));
    if (ravg && !ctx->f.uavg[0]) RC(alloc_uavg(ctx));
    ctx->avg_ravg = ravg;
    if (ravg)
        for (int q = 0; q < 3; ++q) HIPCHK(hipMemsetAsync(ctx->f.uavg[q], 0, sizeof(double) * ctx->N, ctx->stream));
    if (histograms && !ctx->dist1) {
        RC(dalloc(ctx, &ctx->dist1, 3 * (size_t)DIST_W));
        RC(dalloc(ctx, &ctx->dist3, (size_t)DIST_W * DIST_W * DIST_W));
    }
    ctx->avg_hist = histograms;
    HIPCHK(hipMemsetAsync(ctx->davg, 0, sizeof(AvgAcc), ctx->stream));
    if (ctx->el_sum) HIPCHK(hipMemsetAsync(ctx->el_sum, 0, sizeof(double) * 2 * ctx->Lz, ctx->stream));
    if (ctx->dist1) {
        HIPCHK(hipMemsetAsync(ctx->dist1, 0, sizeof(unsigned long long) * 3 * DIST_W, ctx->stream));
        HIPCHK(hipMemsetAsync(ctx->dist3, 0, sizeof(int) * (size_t)DIST_W * DIST_W * DIST_W, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FHIP_OK;
}

int fhip_avg_add(fhip_ctx *ctx)
{
    if (!ctx) return FHIP_ERR_ARG;
    return avg_add(ctx);
}

int fhip_avg_auto(fhip_ctx *ctx, int enable)
{
    if (!ctx) return FHIP_ERR_ARG;
    if (enable && !ctx->davg) return fail(ctx, FHIP_ERR_STATE, "averages not reset (fhip_avg_reset)");
    ctx->avg_auto = enable != 0;
    return FHIP_OK;
}

int fhip_avg_get(fhip_ctx *ctx, fhip_average *out, int64_t *distribution, int32_t *distribution3d)
{
    if (!ctx || !out) return fail(ctx, FHIP_ERR_ARG, "null argument");
    if (!ctx->davg) return fail(ctx, FHIP_ERR_STATE, "averages not reset (fhip_avg_reset)");
    if ((distribution || distribution3d) && !ctx->avg_hist)
        return fail(ctx, FHIP_ERR_STATE, "histograms were not requested in fhip_avg_reset");
    AvgAcc a{};
    HIPCHK(hipMemcpyAsync(&a, ctx->davg, sizeof(AvgAcc), hipMemcpyDeviceToHost, ctx->stream));
    std::vector<unsigned long long> d1;
    if (distribution) {
        d1.resize(3 * (size_t)DIST_W);
        HIPCHK(hipMemcpyAsync(d1.data(), ctx->dist1, sizeof(unsigned long long) * d1.size(), hipMemcpyDeviceToHost,
                              ctx->stream));
    }
    if (distribution3d)
        HIPCHK(hipMemcpyAsync(distribution3d, ctx->dist3, sizeof(int) * (size_t)DIST_W * DIST_W * DIST_W,
                              hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    static_assert(sizeof(AvgAcc) == sizeof(fhip_average), "fhip_average mirrors AvgAcc");
    std::memcpy(out, &a, sizeof(AvgAcc) - sizeof(double));
    out->n_add = (int64_t)a.n_add;
    // distribution(3, -max_k:max_k) with the component fastest (average_module.F:28)
    if (distribution)
        for (int k = 0; k < DIST_W; ++k)
            for (int q = 0; q < 3; ++q) distribution[q + 3 * k] = (int64_t)d1[q * DIST_W + k];
    return collect_events(ctx);
}

int fhip_use_graph(fhip_ctx *ctx, int enable)
{
    if (!ctx) return FHIP_ERR_ARG;
    ctx->use_graph = enable != 0;
    if (!ctx->use_graph) drop_step_graph(ctx);
    return FHIP_OK;
}

int fhip_profile(fhip_ctx *ctx, int enable)
{
    if (!ctx) return FHIP_ERR_ARG;
    RC(collect_events(ctx));
    ctx->prof = enable != 0;
    for (int i = 0; i < K_NKERNEL; ++i) {
        ctx->kms[i] = 0.0;
        ctx->kcount[i] = 0;
    }
    return FHIP_OK;
}

int fhip_kernel_stats(fhip_ctx *ctx, int n, const char **names, double *total_ms, int64_t *launches)
{
    if (!ctx) return FHIP_ERR_ARG;
    RC(collect_events(ctx));
    for (int i = 0; i < K_NKERNEL && i < n; ++i) {
        if (names) names[i] = kKernelNames[i];
        if (total_ms) total_ms[i] = ctx->kms[i];
        if (launches) launches[i] = ctx->kcount[i];
    }
    return K_NKERNEL;
}

int fhip_synchronize(fhip_ctx *ctx)
{
    if (!ctx) return FHIP_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return collect_events(ctx);
}

}  // extern "C"

// w reconstructed from the optimised w(k): acouR = c2r(-i acouK)/N
// (coord_module.F:457-469).
template <typename T>
static int reconstruct_w_t(fhip_ctx *ctx, std::vector<double> &h)
{
    if (ctx->c.acoustic) return fail(ctx, FHIP_ERR_ARG, "w is dynamical here: read FHIP_ACOU_R");
    if (!ctx->acouK_valid) return fail(ctx, FHIP_ERR_STATE, "no force evaluation yet");
    const size_t N = ctx->N;
    double *tmp = nullptr;
    void *raw = nullptr;
    HIPCHK(hipMalloc(&raw, sizeof(double) * 3 * N));
    tmp = (double *)raw;
    Consts c = ctx->c;
    Spectra<T> sp = tset<T>(ctx).sp;
    const cpx<double> *aK = ctx->acouK;
    hipStream_t s = ctx->stream;
    const int kb = (int)((ctx->Mh + NT - 1) / NT);
    int rc = launch(ctx, K_SETUP, [&] { k_prep_w<T><<<kb, NT, 0, s>>>(c, sp, aK); });
    if (rc == FHIP_OK) {
        ZArgs<T> za = zargs<T>(ctx, false);
        za.nplain = 3;
        za.plain_sign = +1;
        rc = run_zplain<T>(ctx, za);
    }
    if (rc == FHIP_OK) rc = run_ypass<T>(ctx, +1, 0, 3);
    if (rc == FHIP_OK) {
        XinvArgs ia{};
        ia.mode = XI_GENERIC;
        ia.partials = ctx->partials;
        ia.maxblocks = ctx->maxblocks;
        for (int q = 0; q < 3; ++q) ia.gen[q] = tmp + q * N;
        rc = run_xinv<T>(ctx, ia, 3);
    }
    if (rc == FHIP_OK) {
        hipError_t e = hipMemcpyAsync(h.data(), tmp, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = fail(ctx, FHIP_ERR_HIP, std::string("reconstruct_w: ") + hipGetErrorString(e));
    }
    hipFree(raw);
    if (rc != FHIP_OK) return rc;
    for (size_t i = 0; i < 3 * N; ++i) h[i] = h[i] * ctx->c.Ninv;
    // the spectral slots were used as scratch: the next force evaluation refills them
    return collect_events(ctx);
}

static int reconstruct_w(fhip_ctx *ctx, std::vector<double> &h)
{
    if (ctx->precision == FHIP_FP32) return reconstruct_w_t<float>(ctx, h);
    return reconstruct_w_t<double>(ctx, h);
}
