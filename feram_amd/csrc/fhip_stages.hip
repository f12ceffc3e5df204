// fhip_stages.hip -- launchers of the kernels whose FFT length is a template
// parameter, for one group of supercell edges.
//
// The file is compiled once per group (FHIP_STAGE_GROUP = 0 .. 12), each
// object holding the device code of its edges only, so the edges
// compile in parallel translation units; feram_hip.hip and fhip_batch.hip
// dispatch on the edge and call the instantiation of that edge.
#include "fhip_host.hpp"

#include <algorithm>
#include <cstdlib>

using namespace fhip;
using namespace fhip_host;

#ifndef FHIP_STAGE_GROUP
#error "compile with -DFHIP_STAGE_GROUP=<group>"
#endif

// groups of edges (every edge also gets the batch launchers)
#if FHIP_STAGE_GROUP == 0
#define FHIP_GROUP_SIZES(X) X(4) X(6)
#define FHIP_GROUP_BATCH(X) X(4) X(6)
#elif FHIP_STAGE_GROUP == 1
#define FHIP_GROUP_SIZES(X) X(8) X(12)
#define FHIP_GROUP_BATCH(X) X(8) X(12)
#elif FHIP_STAGE_GROUP == 2
#define FHIP_GROUP_SIZES(X) X(16) X(20)
#define FHIP_GROUP_BATCH(X) X(16) X(20)
#elif FHIP_STAGE_GROUP == 3
#define FHIP_GROUP_SIZES(X) X(24) X(32)
#define FHIP_GROUP_BATCH(X) X(24) X(32)
#elif FHIP_STAGE_GROUP == 4
#define FHIP_GROUP_SIZES(X) X(40) X(48)
#define FHIP_GROUP_BATCH(X) X(40) X(48)
#elif FHIP_STAGE_GROUP == 5
#define FHIP_GROUP_SIZES(X) X(60) X(64)
#define FHIP_GROUP_BATCH(X) X(60) X(64)
#elif FHIP_STAGE_GROUP == 6
#define FHIP_GROUP_SIZES(X) X(80) X(81)
#define FHIP_GROUP_BATCH(X) X(80) X(81)
#elif FHIP_STAGE_GROUP == 7
#define FHIP_GROUP_SIZES(X) X(90) X(96)
#define FHIP_GROUP_BATCH(X) X(90) X(96)
#elif FHIP_STAGE_GROUP == 8
#define FHIP_GROUP_SIZES(X) X(100)
#define FHIP_GROUP_BATCH(X) X(100)
#elif FHIP_STAGE_GROUP == 9
#define FHIP_GROUP_SIZES(X) X(128)
#define FHIP_GROUP_BATCH(X) X(128)
#elif FHIP_STAGE_GROUP == 10
#define FHIP_GROUP_SIZES(X) X(192)
#define FHIP_GROUP_BATCH(X) X(192)
#elif FHIP_STAGE_GROUP == 11
#define FHIP_GROUP_SIZES(X) X(243)
#define FHIP_GROUP_BATCH(X) X(243)
#elif FHIP_STAGE_GROUP == 12
#define FHIP_GROUP_SIZES(X) X(256)
#define FHIP_GROUP_BATCH(X) X(256)
#else
#error "unknown FHIP_STAGE_GROUP"
#endif

namespace fhip_stage {

template <typename T, int n>
int xfwd(fhip_ctx *ctx, const XfwdArgs &a)
{
    Consts c = ctx->c;
    FieldsT<T> f = fields<T>(ctx);
    Spectra<T> sp = tset<T>(ctx).sp;
    const Scalars *sc = ctx->dsc;
    const cpx<T> *tw = tset<T>(ctx).twx;
    const int nb = ctx->xfblocks;
    hipStream_t s = ctx->stream;
    return launch(ctx, K_XFWD, [&] { k_xfwd<T, n><<<nb, NT, 0, s>>>(c, f, sp, a, sc, tw); });
}

template <typename T, int n>
int xinv(fhip_ctx *ctx, const XinvArgs &a, int nf)
{
    Consts c = ctx->c;
    FieldsT<T> f = fields<T>(ctx);
    Spectra<T> sp = tset<T>(ctx).sp;
    const Scalars *sc = ctx->dsc;
    const cpx<T> *tw = tset<T>(ctx).twx;
    const int nb = ctx->xblocks;
    hipStream_t s = ctx->stream;
    if (nf == 15) return launch(ctx, K_XINV, [&] { k_xinv<T, n, 15><<<nb, NT, 0, s>>>(c, f, sp, a, sc, tw); });
    if (nf == 12) return launch(ctx, K_XINV, [&] { k_xinv<T, n, 12><<<nb, NT, 0, s>>>(c, f, sp, a, sc, tw); });
    if (nf == 9) return launch(ctx, K_XINV, [&] { k_xinv<T, n, 9><<<nb, NT, 0, s>>>(c, f, sp, a, sc, tw); });
    return launch(ctx, K_XINV, [&] { k_xinv<T, n, 3><<<nb, NT, 0, s>>>(c, f, sp, a, sc, tw); });
}

template <typename T, int n>
int ypass(fhip_ctx *ctx, int sign, int slot0, int nf, hipStream_t s)
{
    Consts c = ctx->c;
    Spectra<T> sp = tset<T>(ctx).sp;
    const cpx<T> *tw = tset<T>(ctx).twy;
    dim3 grid(ctx->yblocks[tix<T>()], nf);
    const int id = sign < 0 ? K_YFWD : K_YINV;
    if (sign < 0) return launch(ctx, id, [&] { k_ypass<T, n, -1><<<grid, NT, 0, s>>>(c, sp, slot0, tw); }, s);
    return launch(ctx, id, [&] { k_ypass<T, n, +1><<<grid, NT, 0, s>>>(c, sp, slot0, tw); }, s);
}

template <typename T, int n>
int zplain(fhip_ctx *ctx, const ZArgs<T> &a, hipStream_t s, size_t shm)
{
    Consts c = ctx->c;
    Spectra<T> sp = tset<T>(ctx).sp;
    Scalars *sc = ctx->dsc;
    const cpx<T> *tw = tset<T>(ctx).twz;
    const int nb = ctx->zblocks[tix<T>()];
    const ZArgs<T> za = a;
    RC(ensure_lds(ctx, (const void *)k_zpass<T, n, ZG_PLAIN>, shm));
    return launch(ctx, K_SETUP, [&] { k_zpass<T, n, ZG_PLAIN><<<nb, NT, shm, s>>>(c, sp, za, sc, tw); }, s);
}

template <typename T, int n, int G2, int DDG>
int zpair(fhip_ctx *ctx, const ZArgs<T> &a, size_t shm)
{
    Consts c = ctx->c;
    Spectra<T> sp = tset<T>(ctx).sp;
    Scalars *sc = ctx->dsc;
    const cpx<T> *tw = tset<T>(ctx).twz;
    hipStream_t s = ctx->stream;
    RC(ensure_lds(ctx, (const void *)k_zpair<T, n, G2, DDG>, shm));
    // The two groups of a tile run one after the other in one workgroup, or
    // side by side in two where the grid then needs fewer rounds of the
    // device's workgroup slots (counted in half-tile durations): grids of
    // fewer tiles than slots, or just over a whole number of rounds.
    // FHIP_ZSPLIT=0 or 1 overrides the choice.
    int &zsplit = ctx->zsplit[tix<T>()];
    if (zsplit < 0) {
        int per_cu = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_zpair<T, n, G2, DDG>, NT, shm));
        const long slots = (long)std::max(per_cu, 1) * std::max(ctx->ncu, 1);
        const long tiles = ctx->zblocks[tix<T>()];
        const long whole = (tiles + slots - 1) / slots, halves = (2 * tiles + slots - 1) / slots;
        const char *zs = std::getenv("FHIP_ZSPLIT");
        zsplit = (zs && (zs[0] == '0' || zs[0] == '1')) ? zs[0] == '1' : halves < 2 * whole;
    }
    ZArgs<T> za = a;
    za.split = zsplit;
    const int nb = ctx->zblocks[tix<T>()] * (zsplit ? 2 : 1);
    return launch(ctx, K_ZPAIR, [&] { k_zpair<T, n, G2, DDG><<<nb, NT, shm, s>>>(c, sp, za, sc, tw); });
}

// ---- replica batches -------------------------------------------------------
template <typename T, int n>
constexpr int ytile_full() { return tile_of<n, line_tile_cap<T>()>::value; }
template <typename T, int n>
constexpr int ytile_half() { return ytile_full<T, n>() > 1 ? ytile_full<T, n>() / 2 : 1; }

template <typename T, int n>
void b_xfwd(const Rep<T> *r, dim3 g, hipStream_t s, bool half, int phase, int mode)
{
    if (half) kb_xfwd<T, n, xfwd_rows<n>::value><<<g, NT, 0, s>>>(r, phase, mode);
    else kb_xfwd<T, n, tile_of<n>::value><<<g, NT, 0, s>>>(r, phase, mode);
}

template <typename T, int n>
void b_ypass(const Rep<T> *r, dim3 g, hipStream_t s, bool half, int sign)
{
    if (half) {
        if (sign < 0) kb_ypass<T, n, -1, ytile_half<T, n>()><<<g, NT, 0, s>>>(r);
        else kb_ypass<T, n, +1, ytile_half<T, n>()><<<g, NT, 0, s>>>(r);
    } else {
        if (sign < 0) kb_ypass<T, n, -1, ytile_full<T, n>()><<<g, NT, 0, s>>>(r);
        else kb_ypass<T, n, +1, ytile_full<T, n>()><<<g, NT, 0, s>>>(r);
    }
}

template <typename T, int n, int G2, int DDG>
void b_zpair(const Rep<T> *r, dim3 g, hipStream_t s, size_t shm, int want_acouK)
{
    kb_zpair<T, n, G2, DDG><<<g, NT, shm, s>>>(r, want_acouK);
}

template <typename T, int n, int G2, int DDG>
const void *b_zpair_fn()
{
    return (const void *)kb_zpair<T, n, G2, DDG>;
}

template <typename T, int n>
void b_xinv(const Rep<T> *r, dim3 g, hipStream_t s, int nf, int materialize)
{
    // 9 forward fields, 3 more with acoustic MD and 3 more for species of
    // different Z* (Z*(R) u)
    if (nf == 15) kb_xinv<T, n, 15><<<g, NT, 0, s>>>(r, materialize);
    else if (nf == 12) kb_xinv<T, n, 12><<<g, NT, 0, s>>>(r, materialize);
    else kb_xinv<T, n, 9><<<g, NT, 0, s>>>(r, materialize);
}

// ---- explicit instantiations of this group's edges ------------------------------
#define FHIP_INST_T(T, n)                                                                   \
    template int xfwd<T, n>(fhip_ctx *, const XfwdArgs &);                                  \
    template int xinv<T, n>(fhip_ctx *, const XinvArgs &, int);                             \
    template int ypass<T, n>(fhip_ctx *, int, int, int, hipStream_t);                       \
    template int zplain<T, n>(fhip_ctx *, const ZArgs<T> &, hipStream_t, size_t);           \
    template int zpair<T, n, ZG_STRAIN, ZG_DD>(fhip_ctx *, const ZArgs<T> &, size_t);       \
    template int zpair<T, n, ZG_ACOU, ZG_DD>(fhip_ctx *, const ZArgs<T> &, size_t);         \
    template int zpair<T, n, ZG_STRAIN, ZG_DDZ>(fhip_ctx *, const ZArgs<T> &, size_t);      \
    template int zpair<T, n, ZG_ACOU, ZG_DDZ>(fhip_ctx *, const ZArgs<T> &, size_t);
#define FHIP_INST_B(T, n)                                                                   \
    template void b_xfwd<T, n>(const Rep<T> *, dim3, hipStream_t, bool, int, int);          \
    template void b_ypass<T, n>(const Rep<T> *, dim3, hipStream_t, bool, int);              \
    template void b_zpair<T, n, ZG_STRAIN, ZG_DD>(const Rep<T> *, dim3, hipStream_t, size_t, int);  \
    template void b_zpair<T, n, ZG_ACOU, ZG_DD>(const Rep<T> *, dim3, hipStream_t, size_t, int);    \
    template void b_zpair<T, n, ZG_STRAIN, ZG_DDZ>(const Rep<T> *, dim3, hipStream_t, size_t, int); \
    template void b_zpair<T, n, ZG_ACOU, ZG_DDZ>(const Rep<T> *, dim3, hipStream_t, size_t, int);   \
    template const void *b_zpair_fn<T, n, ZG_STRAIN, ZG_DD>();                                      \
    template const void *b_zpair_fn<T, n, ZG_ACOU, ZG_DD>();                                        \
    template const void *b_zpair_fn<T, n, ZG_STRAIN, ZG_DDZ>();                                     \
    template const void *b_zpair_fn<T, n, ZG_ACOU, ZG_DDZ>();                                       \
    template void b_xinv<T, n>(const Rep<T> *, dim3, hipStream_t, int, int);

#define FHIP_INST(n)         \
    FHIP_INST_T(double, n)   \
    FHIP_INST_T(float, n)
FHIP_GROUP_SIZES(FHIP_INST)

#ifdef FHIP_GROUP_BATCH
#define FHIP_INSTB(n)        \
    FHIP_INST_B(double, n)   \
    FHIP_INST_B(float, n)
FHIP_GROUP_BATCH(FHIP_INSTB)
#endif

}  // namespace fhip_stage
