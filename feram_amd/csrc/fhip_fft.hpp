// fhip_fft.hpp -- LDS-staged Stockham FFT building blocks for gfx950.
//
// A workgroup holds a tile of TB lines of length N in LDS, element t of line b
// at buf[t*TB + b] (lines interleaved, so that TB consecutive lines that are
// consecutive in HBM load and store coalesced).  Each pass of radix R does the
// N/R*TB butterflies of the tile: every thread loads its butterflies into
// registers, the workgroup synchronises, and the results are written back to
// their self-sorting (Stockham) positions.  Radices 16, 8, 4, 2, 3 and 5
// cover the supercell sizes feram runs (powers of two times 3 or 5).
//
// Conventions follow FFTW's: forward is e^{-i}, inverse is e^{+i}, neither is
// normalised.  Twiddles come from a per-size table tw[j] = e^{-2 pi i j / N}
// held in HBM (read through the L1/L2 caches).
#pragma once
#include <hip/hip_runtime.h>

namespace fhip {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// (SIGN * i) * z
template <int SIGN, typename T>
__device__ __forceinline__ cpx<T> mul_si(cpx<T> z)
{
    return SIGN > 0 ? cpx<T>{-z.y, z.x} : cpx<T>{z.y, -z.x};
}

// In-register DFT of length R with kernel e^{SIGN 2 pi i pq / R}.
template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft(cpx<T>* v)
{
    if constexpr (R == 2) {
        cpx<T> a = v[0];
        v[0] = cadd(a, v[1]);
        v[1] = csub(a, v[1]);
    } else if constexpr (R == 4) {
        cpx<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        cpx<T> s13 = cadd(v[1], v[3]), d13 = mul_si<SIGN>(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    } else if constexpr (R == 8) {
        // 2 x 4: even and odd halves by radix 4, then one radix-2 stage with
        // the twiddles w^k, w = e^{SIGN 2 pi i / 8}
        const T r2 = T(0.70710678118654752440);
        cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
        cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
        dft<4, SIGN>(e);
        dft<4, SIGN>(o);
        const cpx<T> o1 = {r2 * (o[1].x - SIGN * o[1].y), r2 * (o[1].y + SIGN * o[1].x)};   // w^1 o1
        const cpx<T> o2 = mul_si<SIGN>(o[2]);                                             // w^2 o2
        const cpx<T> o3 = {r2 * (-o[3].x - SIGN * o[3].y), r2 * (-o[3].y + SIGN * o[3].x)}; // w^3 o3
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    } else if constexpr (R == 16) {
        // 4 x 4: column DFTs over n1 of x[4 n1 + n2], twiddles w^{n2 k1}
        // (w = e^{SIGN 2 pi i / 16}), row DFTs over n2; X[k1 + 4 k2]
        const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173);
        const T r2 = T(0.70710678118654752440);
        cpx<T> y[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
#pragma unroll
            for (int n1 = 0; n1 < 4; ++n1) y[n2][n1] = v[4 * n1 + n2];
            dft<4, SIGN>(y[n2]);
        }
        // w^e = cos(2 pi e / 16) + SIGN i sin(2 pi e / 16) for the exponents n2 * k1
        auto tw = [&](cpx<T> z, T c, T s) { return cpx<T>{z.x * c - SIGN * z.y * s, z.y * c + SIGN * z.x * s}; };
        y[1][1] = tw(y[1][1], c1, s1);      // e = 1
        y[1][2] = tw(y[1][2], r2, r2);      // e = 2
        y[1][3] = tw(y[1][3], s1, c1);      // e = 3
        y[2][1] = tw(y[2][1], r2, r2);      // e = 2
        y[2][2] = mul_si<SIGN>(y[2][2]);    // e = 4
        y[2][3] = tw(y[2][3], -r2, r2);     // e = 6
        y[3][1] = tw(y[3][1], s1, c1);      // e = 3
        y[3][2] = tw(y[3][2], -r2, r2);     // e = 6
        y[3][3] = tw(y[3][3], -c1, -s1);    // e = 9
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            cpx<T> z[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
            dft<4, SIGN>(z);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = z[k2];
        }
    } else if constexpr (R == 3) {
        const T h = T(0.5), s3 = T(0.86602540378443864676);
        cpx<T> t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        cpx<T> m = {v[0].x - h * t1.x, v[0].y - h * t1.y};
        cpx<T> r = mul_si<SIGN>(cpx<T>{s3 * t2.x, s3 * t2.y});
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    } else if constexpr (R == 5) {
        const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
        const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
        cpx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        cpx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        cpx<T> m1 = {v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
        cpx<T> m2 = {v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
        cpx<T> r1 = mul_si<SIGN>(cpx<T>{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y});
        cpx<T> r2 = mul_si<SIGN>(cpx<T>{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y});
        v[0] = cadd(v[0], cadd(a1, a2));
        v[1] = cadd(m1, r1);
        v[4] = csub(m1, r1);
        v[2] = cadd(m2, r2);
        v[3] = csub(m2, r2);
    }
}

// Radix of the pass after NS: the largest of 16, 8, 4 that divides what is
// left and still gives every thread of the workgroup a butterfly (fewer
// passes, fewer trips through LDS), else 4, 2, 3, 5.
template <int N, int TB = 1, int NT = 1>
struct radix_of {
    // radix of a pass that leaves the factor REM of the length to transform
    template <int REM>
    static constexpr int of_rem()
    {
        if constexpr (REM % 16 == 0 && (N / 16) * TB >= NT) return 16;
        if constexpr (REM % 8 == 0 && (N / 8) * TB >= NT) return 8;
        return (REM % 4 == 0) ? 4 : (REM % 2 == 0) ? 2 : (REM % 3 == 0) ? 3 : (REM % 5 == 0) ? 5 : 0;
    }
    template <int NS>
    static constexpr int at() { return of_rem<N / NS>(); }
};

// Radices with an in-register DFT.
constexpr bool dft_radix(int r) { return r == 2 || r == 3 || r == 4 || r == 5 || r == 8 || r == 16; }

template <typename T, int N, int TB, int NT, int R, int NS, int SIGN, int ST>
__device__ __forceinline__ void stockham_pass(cpx<T>* buf, const cpx<T>* __restrict__ tw, int tid)
{
    constexpr int M = N / R;
    constexpr int NB = M * TB;
    constexpr int BPT = (NB + NT - 1) / NT;
    cpx<T> v[BPT][R];
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int bf = tid + i * NT;
        if (NB % NT == 0 || bf < NB) {
            const int j = bf / TB, b = bf - j * TB;
            const int k = j % NS;
#pragma unroll
            for (int q = 0; q < R; ++q) v[i][q] = buf[(j + q * M) * ST + b];
            if constexpr (NS > 1) {
#pragma unroll
                for (int q = 1; q < R; ++q) {
                    cpx<T> w = tw[k * q * (N / (NS * R))];
                    if (SIGN > 0) w.y = -w.y;
                    v[i][q] = cmul(v[i][q], w);
                }
            }
            dft<R, SIGN>(v[i]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
        const int bf = tid + i * NT;
        if (NB % NT == 0 || bf < NB) {
            const int j = bf / TB, b = bf - j * TB;
            const int k = j % NS;
            const int idxD = (j / NS) * NS * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) buf[(idxD + q * NS) * ST + b] = v[i][q];
        }
    }
    __syncthreads();
}

// Full in-LDS transform of TB interleaved lines of length N, element t of
// line b at buf[t*ST + b] (ST >= TB; a stride of TB+1 spreads the row-wise
// accesses of the x passes over all LDS banks).  Callers synchronise before
// (data in LDS) and may read the result right after.
template <typename T, int N, int TB, int NT, int SIGN, int ST = TB, int NS = 1>
__device__ __forceinline__ void fft_lds(cpx<T>* buf, const cpx<T>* __restrict__ tw, int tid)
{
    if constexpr (NS < N) {
        constexpr int R = radix_of<N, TB, NT>::template at<NS>();
        static_assert(R > 1, "FFT length must factor into 2, 3 and 5");
        stockham_pass<T, N, TB, NT, R, NS, SIGN, ST>(buf, tw, tid);
        fft_lds<T, N, TB, NT, SIGN, ST, NS * R>(buf, tw, tid);
    }
}

// The Stockham passes of a transform from stride NS up to (not including)
// stride NSEND: the middle passes of a transform whose first and last passes
// run in registers (see zpass_body).
template <typename T, int N, int TB, int NT, int SIGN, int ST, int NS, int NSEND>
__device__ __forceinline__ void fft_lds_span(cpx<T>* buf, const cpx<T>* __restrict__ tw, int tid)
{
    if constexpr (NS < NSEND) {
        constexpr int R = radix_of<N, TB, NT>::template of_rem<NSEND / NS>();
        static_assert(R > 1 && (NSEND / NS) % R == 0, "span must factor into 2, 3 and 5");
        stockham_pass<T, N, TB, NT, R, NS, SIGN, ST>(buf, tw, tid);
        fft_lds_span<T, N, TB, NT, SIGN, ST, NS * R, NSEND>(buf, tw, tid);
    }
}

// Largest power of two p with p*N <= cap (lines per tile).
template <int N, int CAP = 1024>
struct tile_of {
    static constexpr int tb()
    {
        int p = 1;
        while (2 * p * N <= CAP) p *= 2;
        return p;
    }
    static constexpr int value = tb();
};

}  // namespace fhip
