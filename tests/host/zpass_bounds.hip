// zpass_bounds.hip -- host-side bounds check of every index the z pass
// (zpass_body / zpair_body in fhip_kernels.hpp) forms, for every compiled
// edge, spectral type and z group, with the thread count NT the kernels are
// built and launched with.
//
// The thread-to-element maps are the kernels' own (fhip::ZMap, host and
// device code); the LDS pool sizes are the kernels' (zbuf_size,
// zpair_buf_size, ZLds) and the dynamic side-buffer size the launchers'
// (fhip_host::zside_bytes).  For every (thread, iteration) pair the check
// asserts that
//   * the HBM-order and line-order maps give t < LZ, b < TB, and each covers
//     every element of the tile exactly once;
//   * every LDS slot of the transforms (t * STB + f * TB + b, f < FB) lies in
//     the pool of the kernel;
//   * every side-buffer index of the kx = 0 pencils -- first evaluation,
//     second half, the R-pencil split and join partners -- lies in the
//     dynamic LDS the launcher requests, for every tile position of every
//     compiled Lx (Mx = Lx / 2) and Ly;
//   * the HBM offset l + t * nlines stays below Mh = Mx Ly Lz, and the k-array
//     index below Kh;
//   * the block-sum scratch (NS * NT/64 doubles) fits red[NT];
//   * the paired kernel's LDS-DMA prefetch of the second group's first ZPF
//     field tiles: every 16-byte chunk's LDS bytes (wave-uniform base plus
//     16 bytes per lane) lie in the staging area, which fits the pool, the
//     chunks cover every staged element of a field exactly once, the HBM
//     bytes a copied chunk reads lie inside the spectrum, and the reads that
//     move the staged tiles into registers stay in the staging area (or, for
//     chunks not copied, inside the spectrum).
// Run on the CPU; prints one line per configuration group and exits 1 on the
// first violation.
//
// With an argument, `zpass_bounds THREADS`, the same maps are enumerated for a
// launch of THREADS threads per workgroup of kernels compiled for NT: with
// THREADS > NT the extra threads' tiles rows t run past LZ, so their HBM
// offsets run past the spectrum (the check reports the first one).
#include "../../feram_amd/csrc/fhip_host.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace fhip;
using namespace fhip_host;

static long long g_checks = 0;
static int g_threads = NT;      // threads per workgroup of the enumerated launch

#define CHECK(cond, ...)                                                  \
    do {                                                                  \
        ++g_checks;                                                       \
        if (!(cond)) {                                                    \
            std::fprintf(stderr, "violation: %s: ", #cond);              \
            std::fprintf(stderr, __VA_ARGS__);                            \
            std::fprintf(stderr, "\n");                                   \
            std::exit(1);                                                 \
        }                                                                 \
    } while (0)

static const int kSizes[] = {
#define FHIP_CASE(n) n,
    FHIP_SIZES(FHIP_CASE)
#undef FHIP_CASE
};

// Ordinal of the kx = 0 pencils of tile w, as zpass_body numbers them.
static void pencils(int w, int TB, int Mx, int nlines, std::vector<int>& lord, int& b0, int& nkx0)
{
    const int l0 = w * TB;
    b0 = (Mx - l0 % Mx) % Mx;
    nkx0 = 0;
    lord.assign(TB, -1);
    for (int b = 0; b < TB; ++b) {
        const int l = l0 + b;
        const int kx = l % Mx;
        if (l < nlines && kx == 0) {
            lord[b] = (b - b0) / Mx;
            CHECK(lord[b] == nkx0, "pencil ordinal %d at b=%d", lord[b], b);
            ++nkx0;
        }
    }
}

template <typename T, int LZ, int GROUP, int NBUF>
static void check_group(const char* what, size_t pool_elems, int side_nf)
{
    constexpr int TB = ZTile<T, LZ>::TB;
    constexpr int NF = ZGroupShape<GROUP>::NF;
    constexpr int FB = ZGroupShape<GROUP>::FB;
    constexpr int STB = FB * TB + 1;
    using Map = ZMap<NT, TB, LZ>;
    constexpr int EL = Map::EL, PPT = Map::PPT;
    static_assert(EL == ZTile<T, LZ>::EL && PPT == ZTile<T, LZ>::PPT, "tile geometry");
    (void)NBUF;
    CHECK((size_t)LZ * STB <= pool_elems, "%s LZ=%d pool %zu < %d", what, LZ, pool_elems, LZ * STB);

    // the two maps: in range, and each a bijection onto the tile
    std::vector<int> seen_h(EL, 0), seen_l(EL, 0);
    for (int tid = 0; tid < g_threads; ++tid)
        for (int i = 0; i < PPT; ++i) {
            if (!Map::live(tid, i)) continue;
            int t, b, te, be;
            Map::hbm(tid, i, t, b);
            Map::hbm_e(tid + i * NT, te, be);
            CHECK(t == te && b == be, "%s LZ=%d hbm map tid=%d i=%d", what, LZ, tid, i);
            CHECK(t >= 0 && t < LZ && b >= 0 && b < TB,
                  "%s %s LZ=%d: thread %d iteration %d maps to row t=%d (pencil %d): its HBM offset "
                  "l + t*nlines passes Mh = nlines*LZ", what, sizeof(T) == 8 ? "fp64" : "fp32", LZ, tid, i, t, b);
            seen_h[t * TB + b] += 1;
            for (int f = 0; f < FB; ++f)
                CHECK((size_t)(t * STB + f * TB + b) < pool_elems, "%s LZ=%d lds slot", what, LZ);
            Map::line(tid, i, t, b);
            CHECK(t >= 0 && t < LZ && b >= 0 && b < TB, "%s LZ=%d line t=%d b=%d", what, LZ, t, b);
            seen_l[t * TB + b] += 1;
            for (int f = 0; f < FB; ++f) {
                CHECK((size_t)(t * STB + f * TB + b) < pool_elems, "%s LZ=%d line lds slot", what, LZ);
                if (t != 0)
                    CHECK((size_t)((LZ - t) * STB + f * TB + b) < pool_elems, "%s LZ=%d split partner", what, LZ);
            }
        }
    for (int e = 0; e < EL; ++e)
        CHECK(seen_h[e] == 1 && seen_l[e] == 1, "%s LZ=%d element %d seen %d/%d", what, LZ, e, seen_h[e], seen_l[e]);
    // PLAIN loops: e = tid, tid + NT, ... < EL
    for (int e = 0; e < EL; ++e) {
        int t, b;
        Map::hbm_e(e, t, b);
        CHECK((size_t)(t * STB + b) < pool_elems, "%s LZ=%d plain slot", what, LZ);
    }
    if (GROUP == ZG_PLAIN) return;

    // register first / last passes (zpass_body's REG path): thread (t0, b)
    // writes its first-pass outputs to rows PPT*t0 + q and reads its last-pass
    // inputs from rows t0 + q*S, with twiddles tws[t0*q]
    constexpr int S = NT / TB;
    constexpr bool REG = NT % TB == 0 && EL % NT == 0 && dft_radix(PPT) && LZ % (PPT * PPT) == 0;
    if constexpr (REG) {
        std::vector<int> seen_w(EL, 0), seen_r(EL, 0);
        for (int tid = 0; tid < g_threads; ++tid) {
            const int t0 = tid / TB, b = tid % TB;
            for (int q = 0; q < PPT; ++q) {
                const int tw_ = PPT * t0 + q, tr = t0 + q * S;
                CHECK(tw_ < LZ && tr < LZ && t0 * q < LZ, "%s LZ=%d register pass rows t0=%d q=%d", what, LZ, t0, q);
                seen_w[tw_ * TB + b] += 1;
                seen_r[tr * TB + b] += 1;
                for (int f = 0; f < FB; ++f) {
                    CHECK((size_t)(tw_ * STB + f * TB + b) < pool_elems, "%s register first-pass slot", what);
                    CHECK((size_t)(tr * STB + f * TB + b) < pool_elems, "%s register last-pass slot", what);
                }
                int th, bh;
                Map::hbm(tid, q, th, bh);
                CHECK(th == tr && bh == b, "%s register points differ from the HBM map", what);
            }
        }
        for (int e = 0; e < EL; ++e) CHECK(seen_w[e] == 1 && seen_r[e] == 1, "%s register pass coverage", what);
    }

    // side buffer and global offsets for every compiled Lx, Ly
    std::vector<int> lord;
    for (int Lx : kSizes) {
        if (Lx % 2) continue;
        const int Mx = Lx / 2;
        const size_t shm = zside_bytes<T>(LZ, Mx, side_nf);
        if (shm > 96 * 1024) continue;                 // rejected by the launchers
        const size_t side_elems = shm / sizeof(cpx<T>);
        for (int Ly : kSizes) {
            const int nlines = Mx * Ly;
            const int ntiles = (nlines + TB - 1) / TB;
            const size_t Mh = (size_t)nlines * LZ;
            const size_t Kh = (size_t)(Mx + 1) * Ly * LZ;
            std::vector<int> tiles;
            for (int w = 0; w < ntiles && w < Mx + 1; ++w) tiles.push_back(w);
            tiles.push_back(ntiles - 1);
            for (int w : tiles) {
                int b0, nkx0;
                pencils(w, TB, Mx, nlines, lord, b0, nkx0);
                const int l0 = w * TB;
                const int nsp = (b0 < TB && b0 < nlines - l0) ? ((std::min(TB, nlines - l0) - 1 - b0) / Mx + 1) * LZ : 0;
                CHECK(nsp == nkx0 * LZ, "%s nsp %d vs %d pencils", what, nsp, nkx0);
                CHECK((size_t)2 * nsp * NF <= side_elems, "%s side buffer LZ=%d Lx=%d Ly=%d: %d > %zu", what, LZ, Lx,
                      Ly, 2 * nsp * NF, side_elems);
                for (int o = 0; o < nkx0; ++o) CHECK(b0 + o * Mx < TB, "%s lky index", what);
                for (int tid = 0; tid < g_threads; ++tid)
                    for (int i = 0; i < PPT; ++i) {
                        if (!Map::live(tid, i)) continue;
                        int t, b;
                        Map::hbm(tid, i, t, b);
                        const int l = l0 + b;
                        if (l < nlines) {
                            const size_t off = (size_t)l + (size_t)t * nlines;
                            CHECK(off < Mh, "%s hbm offset", what);
                        }
                        Map::line(tid, i, t, b);
                        const int ord = lord[b];
                        if (ord >= 0) {
                            for (int fi = 0; fi < NF; ++fi) {
                                CHECK((size_t)((ord * LZ + t) * NF + fi) < side_elems, "%s side slot", what);
                                if (t != 0)
                                    CHECK((size_t)((ord * LZ + LZ - t) * NF + fi) < side_elems, "%s side partner", what);
                            }
                        } else if (l0 + b < nlines && (l0 + b) % Mx > 0) {
                            const size_t kk = (size_t)(l0 + b) + (size_t)t * nlines;
                            CHECK(kk < Mh && kk < Kh, "%s k index", what);
                        }
                    }
                // register path: the R-pencil split pairs (t, LZ - t) in the side buffer
                if constexpr (REG) {
                    constexpr int npair = (LZ - 1) / 2;
                    const int npen = nsp / LZ;
                    for (int e = 0; e < npen * npair * NF; ++e) {
                        const int fi = e % NF, rest = e / NF, t = 1 + rest % npair, ord = rest / npair;
                        CHECK(b0 + ord * Mx < TB, "%s split pencil", what);
                        CHECK((size_t)((ord * LZ + t) * NF + fi) < side_elems, "%s split lo", what);
                        CHECK((size_t)((ord * LZ + LZ - t) * NF + fi) < side_elems && LZ - t > t, "%s split hi", what);
                    }
                }
                // kx = 0 evaluations: j < nsp over the threads, first half at side + nsp*NF
                for (int j = 0; j < nsp; ++j)
                    for (int q = 0; q < NF; ++q) {
                        CHECK((size_t)(j * NF + q) < side_elems, "%s side j", what);
                        CHECK((size_t)(nsp * NF + j * NF + q) < side_elems, "%s first j", what);
                    }
            }
        }
    }
}

// The prefetch of zpair_body (zprefetch_issue / zprefetch_collect) for every
// compiled Lx, Ly at this Lz.
template <typename T, int LZ>
static void check_prefetch(size_t pool_elems)
{
    using P = ZPrefetch<T, LZ>;
    if constexpr (!P::OK) return;
    constexpr int TB = P::TB, EL = P::EL, CH = P::CH, NCH = P::NCH, PPT = P::PPT;
    using Map = ZMap<NT, TB, LZ>;
    CHECK((size_t)ZPF * EL <= pool_elems, "prefetch staging LZ=%d: %d > %zu", LZ, ZPF * EL, pool_elems);
    CHECK(CH * (int)sizeof(cpx<T>) == 16 && TB % CH == 0, "prefetch chunk LZ=%d", LZ);
    // LDS side: wave-uniform base (it*NT + wave*64)*CH, lane l at +l*CH
    std::vector<int> cover((size_t)ZPF * EL, 0);
    for (int f = 0; f < ZPF; ++f)
        for (int it = 0; it < (NCH + NT - 1) / NT; ++it)
            for (int tid = 0; tid < g_threads; ++tid) {
                const int j = it * NT + tid;
                if (!(NCH % NT == 0 || j < NCH)) continue;
                const int wbase = tid & ~63, lane = tid & 63;
                const size_t lds0 = (size_t)f * EL + (size_t)(it * NT + wbase) * CH + (size_t)lane * CH;
                CHECK(lds0 + CH <= (size_t)ZPF * EL, "prefetch LDS LZ=%d f=%d it=%d tid=%d", LZ, f, it, tid);
                const int e = j * CH, t = e / TB, b = e - t * TB;
                CHECK(t < LZ && b + CH <= TB, "prefetch chunk row LZ=%d t=%d b=%d", LZ, t, b);
                for (int q = 0; q < CH; ++q) {
                    CHECK(lds0 + q == (size_t)f * EL + (size_t)t * TB + b + q, "prefetch layout LZ=%d", LZ);
                    cover[lds0 + q] += 1;
                }
            }
    for (size_t e = 0; e < cover.size(); ++e) CHECK(cover[e] == 1, "prefetch coverage LZ=%d e=%zu", LZ, e);
    // HBM side and the collect, for every tile of every compiled Lx, Ly
    for (int Lx : kSizes) {
        if (Lx % 2) continue;
        const int Mx = Lx / 2;
        for (int Ly : kSizes) {
            const int nlines = Mx * Ly;
            const int ntiles = (nlines + TB - 1) / TB;
            const size_t Mh = (size_t)nlines * LZ;
            for (int w : {0, 1, ntiles / 2, ntiles - 2, ntiles - 1}) {
                if (w < 0 || w >= ntiles) continue;
                for (int j = 0; j < NCH; ++j) {
                    const int e = j * CH, t = e / TB, b = e - t * TB;
                    if (zchunk_ok<T, LZ>(w, b, nlines)) {
                        const size_t off = (size_t)(w * TB + b) + (size_t)t * nlines;
                        CHECK(off + CH <= Mh, "prefetch HBM LZ=%d Lx=%d Ly=%d w=%d j=%d", LZ, Lx, Ly, w, j);
                        CHECK((off * sizeof(cpx<T>)) % 16 == 0, "prefetch alignment LZ=%d Lx=%d Ly=%d", LZ, Lx, Ly);
                    }
                }
                for (int tid = 0; tid < g_threads; ++tid)
                    for (int i = 0; i < PPT; ++i) {
                        if (!Map::live(tid, i)) continue;
                        int t, b;
                        Map::hbm(tid, i, t, b);
                        const int l = w * TB + b;
                        if (zchunk_ok<T, LZ>(w, b - b % CH, nlines))
                            CHECK((size_t)(2 * EL + t * TB + b) < (size_t)ZPF * EL, "prefetch collect LZ=%d", LZ);
                        else if (l < nlines)
                            CHECK((size_t)l + (size_t)t * nlines < Mh, "prefetch direct read LZ=%d", LZ);
                    }
            }
        }
    }
}

template <typename T, int LZ>
static void check_edge()
{
    // k_zpass pools (one group per kernel) and the k_zpair pools (both groups)
    check_group<T, LZ, ZG_PLAIN, zbuf_size<T, LZ, ZG_PLAIN>()>("plain", zbuf_size<T, LZ, ZG_PLAIN>(), 1);
    constexpr size_t pool_s = zpair_buf_size<T, LZ, ZG_STRAIN>();
    constexpr size_t pool_a = zpair_buf_size<T, LZ, ZG_ACOU>();
    check_group<T, LZ, ZG_DD, 0>("zpair(strain) dd", pool_s, 6);
    check_group<T, LZ, ZG_STRAIN, 0>("zpair(strain) strain", pool_s, 6);
    check_group<T, LZ, ZG_DD, 0>("zpair(acou) dd", pool_a, 9);
    check_group<T, LZ, ZG_ACOU, 0>("zpair(acou) acou", pool_a, 9);
    check_prefetch<T, LZ>(pool_s);
    check_prefetch<T, LZ>(pool_a);
    // static LDS of the paired kernel fits the 64 KB static limit
    CHECK(sizeof(ZLds<T, LZ, zpair_buf_size<T, LZ, ZG_STRAIN>()>) <= 64 * 1024, "static LDS strain LZ=%d", LZ);
    CHECK(sizeof(ZLds<T, LZ, zpair_buf_size<T, LZ, ZG_ACOU>()>) <= 64 * 1024, "static LDS acou LZ=%d", LZ);
    constexpr int TB = ZTile<T, LZ>::TB, PPT = ZTile<T, LZ>::PPT;
    constexpr bool REG = NT % TB == 0 && (LZ * TB) % NT == 0 && dft_radix(PPT) && LZ % (PPT * PPT) == 0;
    std::printf("z pass %s LZ=%3d TB=%2d PPT=%d%s: in bounds\n", sizeof(T) == 8 ? "fp64" : "fp32", LZ, TB, PPT,
                REG ? " (register first/last passes)" : "");
}

int main(int argc, char** argv)
{
    if (argc > 1) g_threads = std::atoi(argv[1]);
    // block_sums<NS> scratch: NS * NT/64 doubles in red[NT]
    CHECK(3 * (NT / 64) <= NT, "block_sums scratch");
#define FHIP_CASE(n) check_edge<double, n>(); check_edge<float, n>();
    FHIP_SIZES(FHIP_CASE)
#undef FHIP_CASE
    std::printf("%lld checks passed (NT = %d, %d threads per workgroup)\n", g_checks, NT, g_threads);
    return 0;
}
