// fhip_kernels.hpp -- the per-step kernels of the force-and-integrate path.
//
// Device layout (structure of arrays, unpadded, x fastest):
//   real fields   f[c],  c = x + Lx*(y + Ly*z)                 (one array per component)
//   half spectra  Mx = Lx/2 columns kx = 0..Mx-1 in rows of Mx points,
//                 s[(kz*Ly + ky)*Mx + kx], followed by the Nyquist column
//                 kx = Mx at s[Mh + kz*Ly + ky], Mh = Mx*Ly*Lz.  Rows of
//                 Mx points keep every tile of the y and z passes aligned to
//                 whole cache lines.
// Spectral slots of one step:
//   0..2  u_x,u_y,u_z        -> (S + Z*^2 L) u   (dipole-dipole force before 1/N)
//   3..8  uu_xx..uu_xy       -> eta(k) = k (x) w(k)  (inhomogeneous strain)
//   9..11 w (acoustic MD)    -> dV/dw(k)
// One step is: xfwd (drift + r2c along x) -> y forward -> z pass (forward,
// k-space stage, inverse) -> scalar stage -> y inverse -> xinv (c2r along x,
// local forces, kick).  Reductions go through per-block partial sums that a
// single workgroup folds in a fixed order, so every run is bitwise repeatable.
#pragma once
#include "fhip_fft.hpp"

namespace fhip {

constexpr int NT = 256;          // threads per workgroup in every kernel
constexpr int MAX_SLOTS = 12;    // spectral slots
enum PartialSlot {
    P_ESR = 0, P_ELR, P_EIS, P_EIC, P_EIM, P_EUNH, P_EDE, P_KE1, P_KE2, P_KEA, P_KEPK, P_NSLOT
};

// Constants shared by all kernels (by value in the kernel arguments).
struct Consts {
    int Lx, Ly, Lz, Mx, N, Kh;
    size_t Mh;           // Mx*Ly*Lz: start of the Nyquist column
    int ylines_main;     // Mx*Lz rounded up to whole y tiles
    int zlines_main;     // Mx*Ly rounded up to whole z tiles
    int acoustic;        // mass_acou > 0: w is dynamical
    int efield;          // !E_field_always_zero
    int has_eloc;        // per-cell local field present
    int modulation;      // modulation_constant != 0 with a .modulation field
    double Ninv, dt, dt_2, a0, Z2, Zstar;
    double kappa2[3], fa2g[3], tg[3], sk[3], tk2[3], fk[3], k4x8[3];
    double alpha[3], gamma[3], k1[3], k2[3], k3[3], k4[3];
    double minv[3], mwinv[3];
    double B11, B12, B44, B1xx, B1yy, B4yz, mod_c;
    double j[7];
    double hc2[36];      // homogeneous_coupling_2(i,j) at [i + 6 j]
};

// Scalars living on the device between kernels (FP64 in every mode).
struct Scalars {
    double s, pi, H0;                // Nose-Poincare state visible to the host
    double s_old, c_drift;           // drift factor dt/2 (1/s + 1/s_old)
    double scale_p, scale_pw;        // velocity-scaling factors pending on p, p_w
    double dc[6];                    // Re uu^(k=0)
    double strain[6];
    double e_homo[6];                // sum_i eta_i hc2(i, j)
    double en[13];                   // Energy_type order
    double ke_pk;                    // kinetic energy after the first half kick
    double E_add[3];                 // uniform part of the electric field
    double pad[3];
};
enum EnergyIdx {
    E_ACOU_KIN = 0, E_DIPO_KIN, E_SR, E_LR, E_DE, E_UNH, E_HS, E_HC, E_IS, E_IC, E_IM, E_TOTAL, E_HNP
};

struct Fields {
    double* u[3];
    double* p[3];
    double* pk[3];       // Nose-Poincare: momentum after the first half kick of the next step
    double* F[3];        // materialised dV/du
    double* w[3];
    double* pw[3];
    double* dVdw[3];     // materialised dV/dw
    double* eloc[3];
    const unsigned char* code;   // bits 0-1 species, bit 2 frozen (1/m = 0), bit 3 counted (sum_p)
};

template <typename T>
struct Spectra {
    cpx<T>* s[MAX_SLOTS];
};

template <typename T>
__device__ __forceinline__ cpx<double> todbl(cpx<T> a) { return cpx<double>{(double)a.x, (double)a.y}; }

// Index of half-spectrum point (kx, ky, kz) in the split layout.
__host__ __device__ __forceinline__ size_t kindex(int Mx, int Ly, size_t Mh, int kx, int ky, int kz)
{
    return kx < Mx ? ((size_t)kz * Ly + ky) * Mx + kx : Mh + (size_t)kz * Ly + ky;
}

__device__ __forceinline__ double block_sum(double v, double* red)
{
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
#pragma unroll
    for (int o = NT / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] = red[tid] + red[tid + o];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// xfwd: drift, then r2c along x of every forward field, two real fields per
// complex transform.
// ---------------------------------------------------------------------------
enum XfwdMode { XF_FORCE = 0, XF_NP = 1, XF_LF = 2, XF_VS = 3, XF_GENERIC = 4 };

struct XfwdArgs {
    int mode;
    int nf;                       // number of forward fields
    const double* gen[12];        // XF_GENERIC: real input fields
};

template <typename T, int LX>
__global__ __launch_bounds__(NT) void k_xfwd(Consts c, Fields f, Spectra<T> sp, XfwdArgs a,
                                             const Scalars* __restrict__ sc,
                                             const cpx<T>* __restrict__ twx)
{
    constexpr int RB = tile_of<LX>::value;
    constexpr int ST = RB + 1;
    constexpr int CELLS = RB * LX;
    constexpr int CPT = (CELLS + NT - 1) / NT;
    constexpr int MX = LX / 2;
    __shared__ cpx<T> buf[LX * ST];
    __shared__ cpx<T> tws[LX];
    const int tid = threadIdx.x;
    const int nrows = c.Ly * c.Lz;
    const int row0 = blockIdx.x * RB;
    for (int i = tid; i < LX; i += NT) tws[i] = twx[i];

    double u[CPT][3], w[CPT][3];
    bool ok[CPT];
    size_t cell[CPT];
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
        const int e = tid + i * NT;
        const int r = e / LX, x = e - r * LX;
        ok[i] = (CELLS % NT == 0 || e < CELLS) && (row0 + r < nrows);
        cell[i] = (size_t)(row0 + r) * LX + x;
        if (a.mode == XF_GENERIC) continue;
        if (!ok[i]) {
            u[i][0] = u[i][1] = u[i][2] = 0.0;
            w[i][0] = w[i][1] = w[i][2] = 0.0;
            continue;
        }
        const size_t cc = cell[i];
        const unsigned char code = f.code[cc];
        const int spc = code & 3;
        const double minv = (code & 4) ? 0.0 : c.minv[spc];
        if (a.mode == XF_NP) {
            const double cd = sc->c_drift;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                double uq = f.u[q][cc] + cd * f.pk[q][cc] * minv;     // nose-poincare.F:41-53
                u[i][q] = uq;
                f.u[q][cc] = uq;
            }
        } else if (a.mode == XF_LF || a.mode == XF_VS) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                double pq = f.p[q][cc];
                if (a.mode == XF_VS) {                                 // velocity_scaling.F:77-83
                    pq = pq * sc->scale_p;
                    f.p[q][cc] = pq;
                }
                double uq = f.u[q][cc] + c.dt * pq * minv;            // leapfrog.F:104-114
                u[i][q] = uq;
                f.u[q][cc] = uq;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 3; ++q) u[i][q] = f.u[q][cc];
        }
        if (c.acoustic) {
            const double mwinv = c.mwinv[spc];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                double wq = f.w[q][cc];
                if (a.mode == XF_LF || a.mode == XF_VS) {
                    double pq = f.pw[q][cc];
                    if (a.mode == XF_VS) {
                        pq = pq * sc->scale_pw;
                        f.pw[q][cc] = pq;
                    }
                    wq = wq + c.dt * pq * mwinv;                        // leapfrog.F:116-130
                    f.w[q][cc] = wq;
                }
                w[i][q] = wq;
            }
        }
    }

    const int nf = a.nf;
    for (int f0 = 0; f0 < nf; f0 += 2) {
        const bool two = (f0 + 1 < nf);
#pragma unroll
        for (int i = 0; i < CPT; ++i) {
            const int e = tid + i * NT;
            if (CELLS % NT != 0 && e >= CELLS) continue;
            const int r = e / LX, x = e - r * LX;
            double va = 0.0, vb = 0.0;
            if (ok[i]) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int fi = f0 + h;
                    if (h == 1 && !two) break;
                    double v;
                    if (a.mode == XF_GENERIC) {
                        v = a.gen[fi][cell[i]];
                    } else if (fi < 3) {
                        v = fi == 0 ? u[i][0] : (fi == 1 ? u[i][1] : u[i][2]);
                    } else if (fi < 9) {                                 // dipoFFT.F:20-33
                        const int m = fi - 3;
                        const double ua = (m == 0 || m == 5) ? u[i][0] : (m == 1 || m == 3) ? u[i][1] : u[i][2];
                        const double ub = (m == 0 || m == 4) ? u[i][0] : (m == 1 || m == 5) ? u[i][1] : u[i][2];
                        v = ua * ub;
                    } else {
                        const int q = fi - 9;
                        v = q == 0 ? w[i][0] : (q == 1 ? w[i][1] : w[i][2]);
                    }
                    if (h == 0) va = v; else vb = v;
                }
            }
            buf[x * ST + r] = cpx<T>{T(va), T(vb)};
        }
        __syncthreads();
        fft_lds<T, LX, RB, NT, -1, ST>(buf, tws, tid);
        // unpack Z = A + iB: A[k] = (Z[k] + conj Z[-k])/2, B[k] = (Z[k] - conj Z[-k])/(2i)
        for (int e = tid; e < RB * (MX + 1); e += NT) {
            int r, kx;
            if (e < RB * MX) { r = e / MX; kx = e - r * MX; }
            else { r = e - RB * MX; kx = MX; }
            if (row0 + r >= nrows) continue;
            const cpx<T> zk = buf[kx * ST + r];
            const cpx<T> zn = buf[((LX - kx) % LX) * ST + r];
            const size_t o = (kx < MX) ? (size_t)(row0 + r) * MX + kx : c.Mh + (size_t)(row0 + r);
            sp.s[f0][o] = cpx<T>{(zk.x + zn.x) * T(0.5), (zk.y - zn.y) * T(0.5)};
            if (two) sp.s[f0 + 1][o] = cpx<T>{(zk.y + zn.y) * T(0.5), (zn.x - zk.x) * T(0.5)};
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Line transform along y, in place, one field per blockIdx.y.  Lines
// l < ylines_main are (z, kx<Mx) columns of stride Mx; lines from
// ylines_main on are the Nyquist column of plane z (contiguous).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool yline(const Consts& c, int l, size_t& base, int& stride)
{
    if (l < c.ylines_main) {
        if (l >= c.Mx * c.Lz) return false;
        const int z = l / c.Mx, kx = l - z * c.Mx;
        base = (size_t)z * c.Ly * c.Mx + kx;
        stride = c.Mx;
        return true;
    }
    const int z = l - c.ylines_main;
    if (z >= c.Lz) return false;
    base = c.Mh + (size_t)z * c.Ly;
    stride = 1;
    return true;
}

// Tiles of the y and z passes hold 16 KB: 8 FP64 or 16 FP32 complex lines at
// N = 128, so that every tile row is a whole 128-byte cache line.
template <typename T>
constexpr int line_tile_cap() { return 16384 / (int)sizeof(cpx<T>); }

template <typename T, int LY, int SIGN>
__global__ __launch_bounds__(NT) void k_ypass(Consts c, Spectra<T> sp, int slot0,
                                              const cpx<T>* __restrict__ twy)
{
    constexpr int TB = tile_of<LY, line_tile_cap<T>()>::value;
    __shared__ cpx<T> buf[LY * TB];
    __shared__ cpx<T> tws[LY];
    const int tid = threadIdx.x;
    for (int i = tid; i < LY; i += NT) tws[i] = twy[i];
    const int l0 = blockIdx.x * TB;
    cpx<T>* s = sp.s[slot0 + blockIdx.y];
    const bool nyq = l0 >= c.ylines_main;
    for (int e = tid; e < LY * TB; e += NT) {
        // main tiles: consecutive threads walk consecutive kx; Nyquist tiles:
        // consecutive threads walk y, which is contiguous there
        const int t = nyq ? e % LY : e / TB, b = nyq ? e / LY : e - (e / TB) * TB;
        size_t base; int stride;
        if (yline(c, l0 + b, base, stride)) buf[t * TB + b] = s[base + (size_t)t * stride];
    }
    __syncthreads();
    fft_lds<T, LY, TB, NT, SIGN>(buf, tws, tid);
    for (int e = tid; e < LY * TB; e += NT) {
        const int t = nyq ? e % LY : e / TB, b = nyq ? e / LY : e - (e / TB) * TB;
        size_t base; int stride;
        if (yline(c, l0 + b, base, stride)) s[base + (size_t)t * stride] = buf[t * TB + b];
    }
}

// ---------------------------------------------------------------------------
// z pass: forward transform along z, k-space stage, inverse transform.
// Lines l < zlines_main are (ky, kx<Mx) pencils of stride Mx*Ly; lines from
// zlines_main on are the Nyquist pencils (ky) of stride Ly.
// ---------------------------------------------------------------------------
enum ZGroup { ZG_PLAIN = 0, ZG_DD = 1, ZG_STRAIN = 2, ZG_ACOU = 3 };

template <typename T>
struct ZArgs {
    const T* Ltab;                   // [Kh][6] long-range kernel L(k): xx yy zz yz zx xy
    const cpx<double>* csx;          // cos/sin tables (2 pi i / L) per axis
    const cpx<double>* csy;
    const cpx<double>* csz;
    const cpx<T>* modk;              // FFT of the modulation field (or null)
    cpx<double>* acouK;              // materialised optimised w(k) (or null)
    double* partials;                // [P_NSLOT][maxblocks]
    int maxblocks;
    int nplain;                      // ZG_PLAIN: number of slots
    int plain_sign;                  // ZG_PLAIN: -1 forward, +1 inverse
};

struct ZLine {
    size_t base, stride;
    int kx, ky;
    bool ok;
};

__device__ __forceinline__ ZLine zline(const Consts& c, int l)
{
    ZLine z;
    if (l < c.zlines_main) {
        z.ok = l < c.Mx * c.Ly;
        z.ky = l / c.Mx;
        z.kx = l - z.ky * c.Mx;
        z.base = (size_t)l;
        z.stride = (size_t)c.Mx * c.Ly;
        return z;
    }
    z.ky = l - c.zlines_main;
    z.ok = z.ky < c.Ly;
    z.kx = c.Mx;
    z.base = c.Mh + (size_t)z.ky;
    z.stride = (size_t)c.Ly;
    return z;
}

// k vector of the half-spectrum point (elastic.F:83-89)
__device__ __forceinline__ void kvec(const Consts& c, int kx, int ky, int kz, double k[3])
{
    const double M2PI = 6.28318530717958647693;
    double f3 = (double)kz / c.Lz; if (f3 > 0.5) f3 = f3 - 1.0;
    double f2 = (double)ky / c.Ly; if (f2 > 0.5) f2 = f2 - 1.0;
    k[2] = f3 * M2PI / c.a0;
    k[1] = f2 * M2PI / c.a0;
    k[0] = (double)kx / c.Lx * M2PI / c.a0;
}

// coup2(a, i) (elastic.F:148-175)
__device__ __forceinline__ double coup2(const Consts& c, const double k[3], int a, int i)
{
    if (i < 3) return (i == a) ? k[a] * c.B1xx / 2 : k[a] * c.B1yy / 2;
    // ki = ((0,3,2),(3,0,1),(2,1,0)) by column: component of k multiplying w_a in eta_{i}
    const int m = i - 3;
    int q;
    if (m == 0) q = (a == 0) ? -1 : (a == 1 ? 2 : 1);
    else if (m == 1) q = (a == 0) ? 2 : (a == 1 ? -1 : 0);
    else q = (a == 0) ? 1 : (a == 1 ? 0 : -1);
    return (q < 0 ? 0.0 : k[q]) * c.B4yz;
}

template <typename T, int LZ, int GROUP>
__global__ __launch_bounds__(NT) void k_zpass(Consts c, Spectra<T> sp, ZArgs<T> a, Scalars* sc,
                                              const cpx<T>* __restrict__ twz)
{
    constexpr int TB = tile_of<LZ, line_tile_cap<T>()>::value;
    constexpr int PPT = (LZ * TB + NT - 1) / NT;
    constexpr int NF = (GROUP == ZG_DD) ? 3 : (GROUP == ZG_STRAIN ? 6 : (GROUP == ZG_ACOU ? 9 : 1));
    constexpr int F0 = (GROUP == ZG_DD || GROUP == ZG_PLAIN) ? 0 : 3;
    __shared__ cpx<T> buf[LZ * TB];
    __shared__ cpx<T> tws[LZ];
    __shared__ double red[NT];
    const int tid = threadIdx.x;
    const int l0 = blockIdx.x * TB;
    for (int i = tid; i < LZ; i += NT) tws[i] = twz[i];

    // consecutive lines are consecutive in HBM for both kinds of pencils
#define FHIP_LOAD_TILE(SRC)                                                    \
    for (int e = tid; e < LZ * TB; e += NT) {                                  \
        const int t = e / TB, b = e - t * TB;                                  \
        const ZLine zl = zline(c, l0 + b);                                     \
        buf[e] = zl.ok ? (SRC)[zl.base + (size_t)t * zl.stride] : cpx<T>{T(0), T(0)}; \
    }
#define FHIP_STORE_TILE(DST)                                                   \
    for (int e = tid; e < LZ * TB; e += NT) {                                  \
        const int t = e / TB, b = e - t * TB;                                  \
        const ZLine zl = zline(c, l0 + b);                                     \
        if (zl.ok) (DST)[zl.base + (size_t)t * zl.stride] = buf[e];            \
    }

    if constexpr (GROUP == ZG_PLAIN) {
        for (int sl = 0; sl < a.nplain; ++sl) {
            FHIP_LOAD_TILE(sp.s[sl]);
            __syncthreads();
            if (a.plain_sign < 0) fft_lds<T, LZ, TB, NT, -1>(buf, tws, tid);
            else fft_lds<T, LZ, TB, NT, +1>(buf, tws, tid);
            FHIP_STORE_TILE(sp.s[sl]);
            __syncthreads();
        }
        return;
    } else {
        cpx<T> v[NF][PPT];
        // forward transforms, results kept in registers; the tile of the next
        // field is fetched into registers while the current one is transformed
        cpx<T> pf[PPT];
        auto fetch = [&](const cpx<T>* __restrict__ src) {
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                const int e = tid + i * NT;
                pf[i] = cpx<T>{T(0), T(0)};
                if ((LZ * TB) % NT == 0 || e < LZ * TB) {
                    const int t = e / TB, b = e - t * TB;
                    const ZLine zl = zline(c, l0 + b);
                    if (zl.ok) pf[i] = src[zl.base + (size_t)t * zl.stride];
                }
            }
        };
        fetch(sp.s[F0]);
#pragma unroll
        for (int fi = 0; fi < NF; ++fi) {
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                const int e = tid + i * NT;
                if ((LZ * TB) % NT == 0 || e < LZ * TB) buf[e] = pf[i];
            }
            __syncthreads();
            if (fi + 1 < NF) fetch(sp.s[F0 + fi + 1]);
            fft_lds<T, LZ, TB, NT, -1>(buf, tws, tid);
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                const int e = tid + i * NT;
                if ((LZ * TB) % NT == 0 || e < LZ * TB) v[fi][i] = buf[e];
            }
            __syncthreads();
        }

        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            const int e = tid + i * NT;
            if (!((LZ * TB) % NT == 0 || e < LZ * TB)) continue;
            const int kz = e / TB, b = e - kz * TB;
            const ZLine zl = zline(c, l0 + b);
            if (!zl.ok) continue;
            const int kx = zl.kx, ky = zl.ky;
            const size_t kk = zl.base + (size_t)kz * zl.stride;
            const bool edge = (kx == 0) || (kx == c.Mx);
            if constexpr (GROUP == ZG_DD) {
                // S(k): Fourier transform of the short-range couplings
                // (dipole-dipole-short-range.F:35-83) in closed form.
                const cpx<double> ex = a.csx[kx], ey = a.csy[ky], ez = a.csz[kz];
                const double cx = ex.x, cy = ey.x, cz = ez.x, sx = ex.y, sy = ey.y, szz = ez.y;
                const double j1 = c.j[0], j2 = c.j[1], j3 = c.j[2], j4 = c.j[3], j5 = c.j[4],
                             j6 = c.j[5], j7 = c.j[6];
                const double k2t = 2.0 * c.kappa2[0];
                const double cxyz = cx * cy * cz;
                double S[6];
                S[0] = k2t + 2 * j2 * cx + 2 * j1 * (cy + cz) + 4 * j3 * cx * (cy + cz) + 4 * j4 * cy * cz + 8 * j6 * cxyz;
                S[1] = k2t + 2 * j2 * cy + 2 * j1 * (cz + cx) + 4 * j3 * cy * (cz + cx) + 4 * j4 * cz * cx + 8 * j6 * cxyz;
                S[2] = k2t + 2 * j2 * cz + 2 * j1 * (cx + cy) + 4 * j3 * cz * (cx + cy) + 4 * j4 * cx * cy + 8 * j6 * cxyz;
                S[3] = -4 * j5 * sy * szz - 8 * j7 * sy * szz * cx;   // yz
                S[4] = -4 * j5 * szz * sx - 8 * j7 * szz * sx * cy;   // zx
                S[5] = -4 * j5 * sx * sy - 8 * j7 * sx * sy * cz;     // xy
                double L[6];
                const T* lt = a.Ltab + kk * 6;
#pragma unroll
                for (int m = 0; m < 6; ++m) L[m] = (double)lt[m] * c.Z2;
                const cpx<double> ux = todbl(v[0][i]), uy = todbl(v[1][i]), uz = todbl(v[2][i]);
                // F = (S + Z^2 L) u ; energies via Parseval with weight 1 (edge) or 2
                cpx<double> fs[3], fl[3];
                fs[0] = {S[0] * ux.x + S[5] * uy.x + S[4] * uz.x, S[0] * ux.y + S[5] * uy.y + S[4] * uz.y};
                fs[1] = {S[5] * ux.x + S[1] * uy.x + S[3] * uz.x, S[5] * ux.y + S[1] * uy.y + S[3] * uz.y};
                fs[2] = {S[4] * ux.x + S[3] * uy.x + S[2] * uz.x, S[4] * ux.y + S[3] * uy.y + S[2] * uz.y};
                fl[0] = {L[0] * ux.x + L[5] * uy.x + L[4] * uz.x, L[0] * ux.y + L[5] * uy.y + L[4] * uz.y};
                fl[1] = {L[5] * ux.x + L[1] * uy.x + L[3] * uz.x, L[5] * ux.y + L[1] * uy.y + L[3] * uz.y};
                fl[2] = {L[4] * ux.x + L[3] * uy.x + L[2] * uz.x, L[4] * ux.y + L[3] * uy.y + L[2] * uz.y};
                const double wgt = edge ? 1.0 : 2.0;
                double es = 0.0, el = 0.0;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const cpx<double> uq = todbl(v[q][i]);
                    es += uq.x * fs[q].x + uq.y * fs[q].y;
                    el += uq.x * fl[q].x + uq.y * fl[q].y;
                    v[q][i] = cpx<T>{T(fs[q].x + fl[q].x), T(fs[q].y + fl[q].y)};
                }
                acc0 += wgt * es;
                acc1 += wgt * el;
            } else {
                double k[3];
                kvec(c, kx, ky, kz, k);
                const bool k0 = (kx == 0 && ky == 0 && kz == 0);
                if (k0) {
#pragma unroll
                    for (int m = 0; m < 6; ++m) sc->dc[m] = (double)v[m][i].x;
                }
                // coupling rhs: c_mod + C uu  (optimize-inho-strain.F:36-42, dVdac.F:46-52)
                cpx<double> cm[3], Cu[3];
                cpx<double> mk = {0.0, 0.0};
                if (c.modulation) {
                    const cpx<T> m0 = a.modk[kk];
                    mk = {(double)m0.x, (double)m0.y};
                }
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double fq = c.mod_c * k[q];
                    cm[q] = c.modulation ? cpx<double>{fq * mk.x, fq * mk.y} : cpx<double>{0.0, 0.0};
                    double rr = 0.0, ri = 0.0;
#pragma unroll
                    for (int m = 0; m < 6; ++m) {
                        const double cq = coup2(c, k, q, m);
                        rr = rr + cq * (double)v[m][i].x;
                        ri = ri + cq * (double)v[m][i].y;
                    }
                    Cu[q] = {rr, ri};
                }
                // K(k) (elastic.F:90-104)
                double K[3][3];
#pragma unroll
                for (int bb = 0; bb < 3; ++bb) {
                    const int o1 = (bb + 2) % 3, o2 = (bb + 1) % 3;
#pragma unroll
                    for (int aa = 0; aa < 3; ++aa) {
                        if (aa == bb)
                            K[aa][bb] = c.B11 * k[aa] * k[aa] + c.B44 * k[o1] * k[o1] + c.B44 * k[o2] * k[o2];
                        else
                            K[aa][bb] = c.B12 * k[aa] * k[bb] + c.B44 * k[aa] * k[bb];
                    }
                }
                cpx<double> w[3];
                if constexpr (GROUP == ZG_STRAIN) {
                    if (k0) {
                        w[0] = w[1] = w[2] = cpx<double>{0.0, 0.0};
                    } else {
                        // Cholesky K = G G^T (elastic.F:131-141), then the two
                        // triangular solves of optimize-inho-strain.F:44-57
                        double G[3][3], Gi[3];
#pragma unroll
                        for (int bb = 0; bb < 3; ++bb) {
                            double s2 = 0.0;
#pragma unroll
                            for (int l2 = 0; l2 < bb; ++l2) s2 += G[bb][l2] * G[bb][l2];
                            G[bb][bb] = sqrt(K[bb][bb] - s2);
                            Gi[bb] = 1.0 / G[bb][bb];
#pragma unroll
                            for (int aa = bb + 1; aa < 3; ++aa) {
                                double s3 = 0.0;
#pragma unroll
                                for (int l2 = 0; l2 < bb; ++l2) s3 += G[aa][l2] * G[bb][l2];
                                G[aa][bb] = (K[aa][bb] - s3) * Gi[bb];
                            }
                        }
#pragma unroll
                        for (int q = 0; q < 3; ++q) w[q] = {-cm[q].x - Cu[q].x, -cm[q].y - Cu[q].y};
#pragma unroll
                        for (int aa = 0; aa < 3; ++aa) {
                            double tr = 0.0, ti = 0.0;
#pragma unroll
                            for (int bb = 0; bb < aa; ++bb) { tr += G[aa][bb] * w[bb].x; ti += G[aa][bb] * w[bb].y; }
                            w[aa] = {(w[aa].x - tr) * Gi[aa], (w[aa].y - ti) * Gi[aa]};
                        }
#pragma unroll
                        for (int aa = 2; aa >= 0; --aa) {
                            double tr = 0.0, ti = 0.0;
#pragma unroll
                            for (int bb = aa + 1; bb < 3; ++bb) { tr += w[bb].x * G[bb][aa]; ti += w[bb].y * G[bb][aa]; }
                            w[aa] = {(w[aa].x - tr) * Gi[aa], (w[aa].y - ti) * Gi[aa]};
                        }
                    }
                    if (a.acouK) {
#pragma unroll
                        for (int q = 0; q < 3; ++q) a.acouK[(size_t)q * c.Kh + kk] = w[q];
                    }
                } else {
                    // acoustic MD: w(k) = i FFT(w) (acouFFT.F:19-24);
                    // dV/dw = -i [c_mod + C uu + K w] (dVdac.F:44-54)
#pragma unroll
                    for (int q = 0; q < 3; ++q) w[q] = {-(double)v[6 + q][i].y, (double)v[6 + q][i].x};
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        double kr = 0.0, ki = 0.0;
#pragma unroll
                        for (int bb = 0; bb < 3; ++bb) { kr += K[q][bb] * w[bb].x; ki += K[q][bb] * w[bb].y; }
                        const double vr = ((0.0 + cm[q].x) + Cu[q].x) + kr;
                        const double vi = ((0.0 + cm[q].y) + Cu[q].y) + ki;
                        v[6 + q][i] = cpx<T>{T(vi), T(-vr)};
                    }
                }
                // energies with weight 1/2 on the kx = 0 and kx = Lx/2 planes
                // (energy_module.F:162-239)
                const double h = edge ? 0.5 : 1.0;
                double eis = 0.0, eic = 0.0, eim = 0.0;
#pragma unroll
                for (int aa = 0; aa < 3; ++aa) {
#pragma unroll
                    for (int bb = 0; bb < 3; ++bb)
                        eis += K[aa][bb] * (w[aa].x * w[bb].x + w[aa].y * w[bb].y);
                    eic += w[aa].x * Cu[aa].x + w[aa].y * Cu[aa].y;
                    eim += w[aa].x * cm[aa].x + w[aa].y * cm[aa].y;
                }
                acc0 += h * eis;
                acc1 += h * eic;
                acc2 += h * eim;
                // eta(k) = k (x) w (dVddi_inho_coupling.F:41-47)
                v[0][i] = cpx<T>{T(k[0] * w[0].x), T(k[0] * w[0].y)};
                v[1][i] = cpx<T>{T(k[1] * w[1].x), T(k[1] * w[1].y)};
                v[2][i] = cpx<T>{T(k[2] * w[2].x), T(k[2] * w[2].y)};
                v[3][i] = cpx<T>{T(k[2] * w[1].x + k[1] * w[2].x), T(k[2] * w[1].y + k[1] * w[2].y)};
                v[4][i] = cpx<T>{T(k[0] * w[2].x + k[2] * w[0].x), T(k[0] * w[2].y + k[2] * w[0].y)};
                v[5][i] = cpx<T>{T(k[1] * w[0].x + k[0] * w[1].x), T(k[1] * w[0].y + k[0] * w[1].y)};
            }
        }
        // inverse transforms
#pragma unroll
        for (int fi = 0; fi < NF; ++fi) {
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                const int e = tid + i * NT;
                if ((LZ * TB) % NT == 0 || e < LZ * TB) buf[e] = v[fi][i];
            }
            __syncthreads();
            fft_lds<T, LZ, TB, NT, +1>(buf, tws, tid);
            FHIP_STORE_TILE(sp.s[F0 + fi]);
            __syncthreads();
        }
#undef FHIP_LOAD_TILE
#undef FHIP_STORE_TILE
        const int nb = a.maxblocks;
        if constexpr (GROUP == ZG_DD) {
            double r0 = block_sum(acc0, red), r1 = block_sum(acc1, red);
            if (tid == 0) {
                a.partials[P_ESR * nb + blockIdx.x] = r0;
                a.partials[P_ELR * nb + blockIdx.x] = r1;
            }
        } else {
            double r0 = block_sum(acc0, red), r1 = block_sum(acc1, red), r2 = block_sum(acc2, red);
            if (tid == 0) {
                a.partials[P_EIS * nb + blockIdx.x] = r0;
                a.partials[P_EIC * nb + blockIdx.x] = r1;
                a.partials[P_EIM * nb + blockIdx.x] = r2;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// xinv: c2r along x of the NF inverse fields, then the local forces, energies
// and the kick.
// ---------------------------------------------------------------------------
enum XinvMode { XI_FORCE = 0, XI_NP = 1, XI_LF = 2, XI_VS = 3, XI_GENERIC = 4 };

struct XinvArgs {
    int mode;
    int materialize;             // write F (and p for NP, dV/dw) for the host
    double* partials;
    int maxblocks;
    double* gen[12];             // XI_GENERIC: real outputs (unnormalised c2r)
};

template <typename T, int LX, int NF>
__global__ __launch_bounds__(NT) void k_xinv(Consts c, Fields f, Spectra<T> sp, XinvArgs a,
                                             const Scalars* __restrict__ sc,
                                             const cpx<T>* __restrict__ twx)
{
    constexpr int RB = tile_of<LX>::value;
    constexpr int ST = RB + 1;
    constexpr int CELLS = RB * LX;
    constexpr int CPT = (CELLS + NT - 1) / NT;
    constexpr int MX = LX / 2;
    __shared__ cpx<T> buf[LX * ST];
    __shared__ cpx<T> tws[LX];
    __shared__ double red[NT];
    const int tid = threadIdx.x;
    const int nrows = c.Ly * c.Lz;
    const int row0 = blockIdx.x * RB;
    for (int i = tid; i < LX; i += NT) tws[i] = twx[i];

    // half-spectrum points of this block's rows, read once each; the pair of
    // fields after the current one is fetched while the current one transforms
    constexpr int HP = RB * (MX + 1);
    constexpr int HPT = (HP + NT - 1) / NT;
    cpx<T> pa[HPT], pb[HPT];
    auto fetch = [&](int f0) {
        const bool two = (f0 + 1 < NF);
#pragma unroll
        for (int i = 0; i < HPT; ++i) {
            const int e = tid + i * NT;
            pa[i] = cpx<T>{T(0), T(0)};
            pb[i] = cpx<T>{T(0), T(0)};
            if (HP % NT != 0 && e >= HP) continue;
            int r, kx;
            if (e < RB * MX) { r = e / MX; kx = e - r * MX; }
            else { r = e - RB * MX; kx = MX; }
            if (row0 + r < nrows) {
                const size_t o = (kx < MX) ? (size_t)(row0 + r) * MX + kx : c.Mh + (size_t)(row0 + r);
                pa[i] = sp.s[f0][o];
                if (two) pb[i] = sp.s[f0 + 1][o];
            }
        }
    };

    double acc[CPT][NF];
    fetch(0);
#pragma unroll
    for (int f0 = 0; f0 < NF; f0 += 2) {
        const bool two = (f0 + 1 < NF);
        // Z = A + i B with both half spectra completed Hermitian; Im of
        // kx = 0 and kx = Lx/2 is dropped, as a half-complex c2r does.
#pragma unroll
        for (int i = 0; i < HPT; ++i) {
            const int e = tid + i * NT;
            if (HP % NT != 0 && e >= HP) continue;
            int r, kx;
            if (e < RB * MX) { r = e / MX; kx = e - r * MX; }
            else { r = e - RB * MX; kx = MX; }
            cpx<T> A = pa[i], B = pb[i];
            if (kx == 0 || kx == MX) { A.y = T(0); B.y = T(0); }
            buf[kx * ST + r] = cpx<T>{A.x - B.y, A.y + B.x};
            if (kx != 0 && kx != MX) buf[(LX - kx) * ST + r] = cpx<T>{A.x + B.y, B.x - A.y};
        }
        __syncthreads();
        if (f0 + 2 < NF) fetch(f0 + 2);
        fft_lds<T, LX, RB, NT, +1, ST>(buf, tws, tid);
#pragma unroll
        for (int i = 0; i < CPT; ++i) {
            const int e = tid + i * NT;
            if (CELLS % NT != 0 && e >= CELLS) continue;
            const int r = e / LX, x = e - r * LX;
            const cpx<T> z = buf[x * ST + r];
            acc[i][f0] = (double)z.x;
            if (two) acc[i][f0 + 1] = (double)z.y;
        }
        __syncthreads();
    }

    double eunh = 0.0, ede = 0.0, ke1 = 0.0, ke2 = 0.0, kea = 0.0;
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
        const int e = tid + i * NT;
        if (CELLS % NT != 0 && e >= CELLS) continue;
        const int r = e / LX, x = e - r * LX;
        if (row0 + r >= nrows) continue;
        const size_t cc = (size_t)(row0 + r) * LX + x;
        if (a.mode == XI_GENERIC) {
#pragma unroll
            for (int q = 0; q < NF; ++q) a.gen[q][cc] = acc[i][q];
            continue;
        }
        if constexpr (NF >= 9) {
            const unsigned char code = f.code[cc];
            const int s = code & 3;
            const bool counted = (code & 8) != 0;
            const double minv = (code & 4) ? 0.0 : c.minv[s];
            double u[3], F[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                u[q] = f.u[q][cc];
                F[q] = acc[i][q] * c.Ninv;                            // dVddi-dipole-dipole.F:34-38
            }
            // on-site anharmonic terms (dVddi-unharmonic.F:34-65)
            const double d0 = u[0] * u[0], d1 = u[1] * u[1], d2 = u[2] * u[2];
            const double d[3] = {d0, d1, d2};
            const double ss = d0 + d1 + d2;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int q1 = (q + 1) % 3, q2 = (q + 2) % 3;
                F[q] = F[q] + u[q] * 2 * (c.kappa2[s] - c.kappa2[0]);
                F[q] = F[q] + u[q] * (c.fa2g[s] * ss - c.tg[s] * d[q]);
                F[q] = F[q] + u[q] * (c.sk[s] * (ss * ss) + c.k4x8[s] * (ss * ss * ss)
                                      - c.tk2[s] * (d[q] * d[q]) - c.fk[s] * d[q1] * d[q2]);
            }
            eunh += (c.kappa2[s] - c.kappa2[0]) * ss + c.alpha[s] * (ss * ss)
                  + c.gamma[s] * (d0 * d1 + d1 * d2 + d2 * d0) + c.k1[s] * (ss * ss * ss)
                  + c.k2[s] * ((d0 * d0) * (d1 + d2) + (d1 * d1) * (d2 + d0) + (d2 * d2) * (d0 + d1))
                  + c.k3[s] * d0 * d1 * d2 + c.k4[s] * ((ss * ss) * (ss * ss));
            // external and local field (mirror_force_mirror.F:39-65)
            if (c.efield) {
                double E[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) E[q] = c.has_eloc ? f.eloc[q][cc] + sc->E_add[q] : sc->E_add[q];
#pragma unroll
                for (int q = 0; q < 3; ++q) F[q] = F[q] - c.Zstar * E[q];
                ede += c.Zstar * (u[0] * E[0] + u[1] * E[1] + u[2] * E[2]);
            }
            // homogeneous coupling (dVddi-homo-coupling.F:28-44)
            {
                const double* eh = sc->e_homo;
                F[0] = F[0] + 2 * eh[0] * u[0] + eh[4] * u[2] + eh[5] * u[1];
                F[1] = F[1] + 2 * eh[1] * u[1] + eh[3] * u[2] + eh[5] * u[0];
                F[2] = F[2] + 2 * eh[2] * u[2] + eh[3] * u[1] + eh[4] * u[0];
            }
            // inhomogeneous coupling (dVddi_inho_coupling.F:54-82)
            {
                double ei[6];
#pragma unroll
                for (int m = 0; m < 6; ++m) {
                    double sum = 0.0;
#pragma unroll
                    for (int jj = 0; jj < 6; ++jj) sum = sum + acc[i][3 + jj] * c.Ninv * c.hc2[jj + 6 * m];
                    ei[m] = sum;
                }
                F[0] = F[0] + 2 * ei[0] * u[0] + ei[4] * u[2] + ei[5] * u[1];
                F[1] = F[1] + 2 * ei[1] * u[1] + ei[3] * u[2] + ei[5] * u[0];
                F[2] = F[2] + 2 * ei[2] * u[2] + ei[3] * u[1] + ei[4] * u[0];
            }
            // kicks
            if (a.mode == XI_NP) {
                const double dt2s = c.dt_2 * sc->s;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double p1 = f.pk[q][cc] - dt2s * F[q];            // nose-poincare.F:66-71
                    const double p2 = p1 - dt2s * F[q];                      // nose-poincare.F:24-29 (next step)
                    if (counted) { ke1 += p1 * p1 * minv; ke2 += p2 * p2 * minv; }
                    f.pk[q][cc] = p2;
                    if (a.materialize) f.p[q][cc] = p1;
                }
            } else if (a.mode == XI_LF) {
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double pq = f.p[q][cc] - c.dt * F[q];             // leapfrog.F:134-139
                    if (counted) ke1 += pq * pq * minv;
                    f.p[q][cc] = pq;
                }
            } else if (a.mode == XI_VS) {
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double pq = counted ? f.p[q][cc] - c.dt * F[q] : 0.0;   // velocity_scaling.F:62-76
                    if (counted) ke1 += pq * pq * minv;
                    f.p[q][cc] = pq;
                }
            }
            if (a.materialize || a.mode == XI_FORCE) {
#pragma unroll
                for (int q = 0; q < 3; ++q) f.F[q][cc] = F[q];
            }
            if constexpr (NF == 12) {
                const double mwinv = c.mwinv[s];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double g = acc[i][9 + q] * c.Ninv;                  // dVdac.F:56-62
                    if (a.mode == XI_LF || a.mode == XI_VS) {
                        const double pq = f.pw[q][cc] - c.dt * g;
                        kea += pq * pq * mwinv;
                        f.pw[q][cc] = pq;
                    }
                    if (a.materialize || a.mode == XI_FORCE) f.dVdw[q][cc] = g;
                }
            }
        }
    }
    if (a.mode == XI_GENERIC) return;
    const int nb = a.maxblocks;
    double r;
    r = block_sum(eunh, red); if (tid == 0) a.partials[P_EUNH * nb + blockIdx.x] = r;
    r = block_sum(ede, red);  if (tid == 0) a.partials[P_EDE * nb + blockIdx.x] = r;
    r = block_sum(ke1, red);  if (tid == 0) a.partials[P_KE1 * nb + blockIdx.x] = r;
    r = block_sum(ke2, red);  if (tid == 0) a.partials[P_KE2 * nb + blockIdx.x] = r;
    r = block_sum(kea, red);  if (tid == 0) a.partials[P_KEA * nb + blockIdx.x] = r;
}

// ---------------------------------------------------------------------------
// Kinetic energy of p (optionally after the first Nose-Poincare half kick).
// ---------------------------------------------------------------------------
template <int KICK>
__global__ __launch_bounds__(NT) void k_kick(Consts c, Fields f, const Scalars* __restrict__ sc,
                                             double* partials, int maxblocks)
{
    __shared__ double red[NT];
    const size_t cc = (size_t)blockIdx.x * NT + threadIdx.x;
    double ke = 0.0, kea = 0.0;
    if (cc < (size_t)c.N) {
        const unsigned char code = f.code[cc];
        const int s = code & 3;
        const double minv = (code & 4) ? 0.0 : c.minv[s];
        const double dt2s = c.dt_2 * sc->s;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            double pq = f.p[q][cc];
            if (KICK) {
                pq = pq - dt2s * f.F[q][cc];                            // nose-poincare.F:24-29
                f.pk[q][cc] = pq;
            }
            if (code & 8) ke += pq * pq * minv;
        }
        if (c.acoustic) {
#pragma unroll
            for (int q = 0; q < 3; ++q) kea += f.pw[q][cc] * f.pw[q][cc] * c.mwinv[s];
        }
    }
    double r = block_sum(ke, red);
    if (threadIdx.x == 0) partials[P_KEPK * maxblocks + blockIdx.x] = r;
    r = block_sum(kea, red);
    if (threadIdx.x == 0) partials[P_KEA * maxblocks + blockIdx.x] = r;
}

// Deterministic fold of the nb per-block partials of NS slots by one
// workgroup: all loads are issued first, then one tree per slot.
template <int NS>
__device__ __forceinline__ void fold_slots(const double* part, int maxblocks, int nb, const int (&slots)[NS],
                                           double (&out)[NS], double* red)
{
    double s[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) s[q] = 0.0;
    for (int i = threadIdx.x; i < nb; i += NT) {
        double v[NS];
#pragma unroll
        for (int q = 0; q < NS; ++q) v[q] = part[slots[q] * maxblocks + i];
#pragma unroll
        for (int q = 0; q < NS; ++q) s[q] += v[q];
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) out[q] = block_sum(s[q], red);
}

// Sum of the nine potential energies (energy_module.F:73-87).
__device__ __forceinline__ double potential(const double* en)
{
    return en[E_SR] + en[E_LR] + en[E_DE] + en[E_UNH] + en[E_HS] + en[E_HC] + en[E_IS] + en[E_IC] + en[E_IM];
}

// Constants of the scalar stage (thermostat, homogeneous strain).
struct ScalarConsts {
    int geometry;            // 0 bulk, 2 epit, 3 e110
    int method;              // 0 md, 1 vs, 2 lf
    int modulation_energy;   // modulation_constant != 0
    int n_sum_p;
    int N;
    double Ninv, dt, dt_2, Q, T_eV, n_freedom, three_N, target_ke, epi_strain;
    double pNa03[3];
    double H[36], G[36], Ginv[6], hc2[36];
};

// Nose-Poincare Hamiltonian (energy_module.F:128-148).
__device__ __forceinline__ double np_hamiltonian(const ScalarConsts& k, const Scalars* sc, const double* en)
{
#pragma clang fp contract(off)
    const double s = sc->s;
    const double H = en[E_DIPO_KIN] / (s * s) + potential(en) + sc->pi * sc->pi / 2 / k.Q
                   + k.three_N * k.T_eV * log(s) - sc->H0;
    return H * s;
}

// First half of a Nose-Poincare step up to the drift factor
// (nose-poincare.F:30-41); en[E_DIPO_KIN] <- KE after the first half kick.
__device__ __forceinline__ void np_begin_scalar(const ScalarConsts& k, Scalars* sc)
{
#pragma clang fp contract(off)
    double* en = sc->en;
    en[E_DIPO_KIN] = sc->ke_pk;
    double s = sc->s;
    double CB = k.n_freedom * k.T_eV * (1.0 + log(s)) - en[E_DIPO_KIN] / (s * s) + potential(en) - sc->H0;
    CB = k.dt_2 * CB - sc->pi;
    const double pi = -2 * CB / (1.0 + sqrt(1.0 - CB * k.dt / k.Q));
    const double s_old = s;
    s = s * (1.0 + k.dt_2 * pi / k.Q) / (1.0 - k.dt_2 * pi / k.Q);
    sc->pi = pi;
    sc->s = s;
    sc->s_old = s_old;
    sc->c_drift = k.dt_2 * (1 / s + 1 / s_old);
}

// Energies of the two z-pass groups, homogeneous strain and its coupling
// (optimize-homo-strain.F:27-85, dVddi-homo-coupling.F:21-26).
__global__ __launch_bounds__(NT) void k_homo(ScalarConsts k, Scalars* sc, const double* partials,
                                             int maxblocks, int zblocks)
{
#pragma clang fp contract(off)
    __shared__ double red[NT];
    const int slots[5] = {P_ESR, P_ELR, P_EIS, P_EIC, P_EIM};
    double r[5];
    fold_slots<5>(partials, maxblocks, zblocks, slots, r, red);
    if (threadIdx.x != 0) return;
    double* en = sc->en;
    en[E_SR] = r[0] * k.Ninv / 2;
    en[E_LR] = r[1] * k.Ninv / 2;
    en[E_IS] = r[2] * k.Ninv;
    en[E_IC] = r[3] * 2 * k.Ninv;
    en[E_IM] = k.modulation_energy ? r[4] * 2 * k.Ninv : 0.0;

    double c2uu[6];
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        for (int j = 0; j < 6; ++j) s = s + k.hc2[i + 6 * j] * sc->dc[j];
        c2uu[i] = s;
    }
    for (int i = 0; i < 3; ++i) c2uu[i] = c2uu[i] + k.pNa03[i];
    double* st = sc->strain;
    const double* H = k.H;
    if (k.geometry == 2) {
        st[0] = k.epi_strain;
        st[1] = k.epi_strain;
        st[2] = (-(H[2 + 6 * 0] + H[2 + 6 * 1]) * k.epi_strain - c2uu[2]) / H[2 + 6 * 2];
        st[3] = -c2uu[3] / H[3 + 6 * 3];
        st[4] = -c2uu[4] / H[4 + 6 * 4];
        st[5] = 0.0;
    } else if (k.geometry == 3) {
        st[2] = k.epi_strain;
        st[3] = 0.0;
        st[4] = 0.0;
        const double num = -(H[0] + H[7] + 2 * H[7]) * k.epi_strain - (H[2] + H[2 + 6]) * st[2]
                         - c2uu[0] - c2uu[1] + 2 * c2uu[5];
        const double den = H[0] + H[7] + 2 * H[1] + 4 * H[35];
        const double delta = num / den;
        st[0] = k.epi_strain + delta;
        st[1] = k.epi_strain + delta;
        st[5] = -2 * delta;
    } else {
        for (int i = 0; i < 6; ++i) st[i] = -c2uu[i];
        for (int i = 0; i < 6; ++i) {
            double t = 0.0;
            for (int j = 0; j < i; ++j) t = t + k.G[i + 6 * j] * st[j];
            st[i] = (st[i] - t) * k.Ginv[i];
        }
        for (int i = 5; i >= 0; --i) {
            double t = 0.0;
            for (int j = i + 1; j < 6; ++j) t = t + st[j] * k.G[j + 6 * i];
            st[i] = (st[i] - t) * k.Ginv[i];
        }
    }
    double hs = 0.0, hc = 0.0;
    for (int i = 0; i < 6; ++i) {
        double d = 0.0;
        for (int j = 0; j < 6; ++j) d = d + H[i + 6 * j] * st[j];
        hs = hs + st[i] * d;
    }
    for (int i = 0; i < 6; ++i) hc = hc + st[i] * c2uu[i];
    en[E_HS] = hs / 2;
    en[E_HC] = hc;
    for (int j = 0; j < 6; ++j) {
        double e = 0.0;
        for (int i = 0; i < 6; ++i) e = e + st[i] * k.hc2[i + 6 * j];
        sc->e_homo[j] = e;
    }
}

enum FinishMode { FIN_FORCE = 0, FIN_NP = 1, FIN_LF = 2, FIN_VS = 3 };

// Local energies and kinetic energies of the xinv pass, then the scalar part
// of the integrator that follows the force evaluation.  With FIN_NP and
// next_step set, the first half of the next Nose-Poincare step (eq. 38 for
// pi and s, the drift factor) follows in the same launch.
__global__ __launch_bounds__(NT) void k_finish(ScalarConsts k, Scalars* sc, const double* partials,
                                               int maxblocks, int xblocks, int mode, int acoustic,
                                               int next_step)
{
#pragma clang fp contract(off)
    __shared__ double red[NT];
    const int slots[5] = {P_EUNH, P_EDE, P_KE1, P_KE2, P_KEA};
    double r[5];
    fold_slots<5>(partials, maxblocks, xblocks, slots, r, red);
    if (threadIdx.x != 0) return;
    const double eunh = r[0], ede = r[1], ke1 = r[2], ke2 = r[3], kea = r[4];
    double* en = sc->en;
    en[E_UNH] = eunh;
    en[E_DE] = -ede;
    const double nsp = (double)k.n_sum_p, Nd = (double)k.N;
    if (mode == FIN_NP) {
        // nose-poincare.F:59-75; en[E_DIPO_KIN] holds KE after the first half kick
        double H = np_hamiltonian(k, sc, en);
        en[E_HNP] = H;
        const double s = sc->s;
        const double dH = H / s;
        sc->pi = sc->pi + k.dt_2 * (2 * en[E_DIPO_KIN] / (s * s) - k.n_freedom * k.T_eV) - k.dt_2 * dH;
        en[E_DIPO_KIN] = ke1 / nsp * Nd / 2;
        en[E_HNP] = np_hamiltonian(k, sc, en);
        sc->ke_pk = ke2 / nsp * Nd / 2;
    } else if (mode == FIN_LF) {
        en[E_DIPO_KIN] = ke1 / nsp * Nd / 2;                            // leapfrog.F:151-154
        if (acoustic) en[E_ACOU_KIN] = kea / 2;
    } else if (mode == FIN_VS) {
        en[E_DIPO_KIN] = ke1 / nsp * Nd / 2;                            // velocity_scaling.F:77-83
        sc->scale_p = sqrt(k.target_ke / en[E_DIPO_KIN]);
        if (acoustic) {
            en[E_ACOU_KIN] = kea / 2;                                   // velocity_scaling.F:86-100
            sc->scale_pw = sqrt(k.target_ke / en[E_ACOU_KIN]);
        }
    }
    en[E_TOTAL] = en[E_DIPO_KIN] / (sc->s * sc->s) + en[E_ACOU_KIN] + potential(en);
    if (mode == FIN_NP && next_step) np_begin_scalar(k, sc);
}

// First half of the first Nose-Poincare step of a call: KE from the kick
// kernel's partials, then eq. 38.
__global__ __launch_bounds__(NT) void k_np_begin(ScalarConsts k, Scalars* sc, const double* partials,
                                                 int maxblocks, int kblocks)
{
#pragma clang fp contract(off)
    __shared__ double red[NT];
    const int slots[1] = {P_KEPK};
    double r[1];
    fold_slots<1>(partials, maxblocks, kblocks, slots, r, red);
    if (threadIdx.x != 0) return;
    sc->ke_pk = r[0] / (double)k.n_sum_p * (double)k.N / 2;
    np_begin_scalar(k, sc);
}

// Kinetic energies of the current momenta for the energy report
// (kinetic_energy_dipo.F, kinetic_energy_acou.F).
__global__ __launch_bounds__(NT) void k_ke_report(ScalarConsts k, Scalars* sc, const double* partials,
                                                  int maxblocks, int kblocks)
{
#pragma clang fp contract(off)
    __shared__ double red[NT];
    const int slots[2] = {P_KEPK, P_KEA};
    double r[2];
    fold_slots<2>(partials, maxblocks, kblocks, slots, r, red);
    if (threadIdx.x != 0) return;
    sc->en[E_DIPO_KIN] = r[0] / (double)k.n_sum_p * (double)k.N / 2;
    sc->en[E_ACOU_KIN] = r[1] / 2;
}

// Apply the pending velocity-scaling factors so the host sees scaled momenta.
__global__ __launch_bounds__(NT) void k_apply_scale(Consts c, Fields f, Scalars* sc)
{
    const size_t cc = (size_t)blockIdx.x * NT + threadIdx.x;
    if (cc >= (size_t)c.N) return;
    const double sp = sc->scale_p;
#pragma unroll
    for (int q = 0; q < 3; ++q) f.p[q][cc] = f.p[q][cc] * sp;
    if (c.acoustic) {
        const double sw = sc->scale_pw;
#pragma unroll
        for (int q = 0; q < 3; ++q) f.pw[q][cc] = f.pw[q][cc] * sw;
    }
}

__global__ void k_reset_scale(Scalars* sc)
{
    sc->scale_p = 1.0;
    sc->scale_pw = 1.0;
}

// ---------------------------------------------------------------------------
// Set-up: Ewald real-space sum (dipole-dipole-long-range.F:47-91) and the
// reciprocal-space sum plus scaling (dipole-dipole-long-range.F:107-150).
// ---------------------------------------------------------------------------
struct EwaldArgs {
    int max_x, max_y, max_z;
    double kappa, a0, Lx_a0, Ly_a0, Lz_a0;
    double four_pi_over_v, two_pi_over_a, four_kappa2, scale;
};

__device__ __forceinline__ double B_decay(double r, double rr, double kappa)
{
    const double M2S = 1.12837916709551257390;
    return erfc(kappa * r) / (r * rr) + M2S * kappa * exp(-(kappa * kappa) * rr) / rr;
}
__device__ __forceinline__ double C_decay(double r, double rr, double kappa)
{
    const double M2S = 1.12837916709551257390;
    const double Crr = 3 * erfc(kappa * r) / (r * rr)
                     + M2S * kappa * exp(-(kappa * kappa) * rr) * (2 * (kappa * kappa) + 3 / rr);
    return Crr / rr;
}

// Real-space part for the 6 upper-triangle components (xx yy zz yz zx xy).
__global__ __launch_bounds__(NT) void k_ewald_real(Consts c, EwaldArgs e, double* out /*[6][N]*/)
{
#pragma clang fp contract(off)
    const size_t cc = (size_t)blockIdx.x * NT + threadIdx.x;
    if (cc >= (size_t)c.N) return;
    const int ix = (int)(cc % c.Lx), iy = (int)((cc / c.Lx) % c.Ly), iz = (int)(cc / ((size_t)c.Lx * c.Ly));
    const int ia[6] = {0, 1, 2, 1, 2, 0}, ib[6] = {0, 1, 2, 2, 0, 1};
    // self-interaction correction on the diagonal at R = 0
    const double self = -(2 * 1.12837916709551257390 * (e.kappa * e.kappa * e.kappa) / 3);
    double T[6];
    for (int m = 0; m < 6; ++m) T[m] = (cc == 0 && m < 3) ? self : 0.0;
    const double r1 = ix * e.a0, r2 = iy * e.a0, r3 = iz * e.a0;
    for (int jz = -1 - e.max_z; jz <= e.max_z; ++jz) {
        const double n3 = jz * e.Lz_a0;
        for (int jy = -1 - e.max_y; jy <= e.max_y; ++jy) {
            const double n2 = jy * e.Ly_a0;
            for (int jx = -1 - e.max_x; jx <= e.max_x; ++jx) {
                if (jx == 0 && jy == 0 && jz == 0 && cc == 0) continue;
                const double n1 = jx * e.Lx_a0;
                const double rn[3] = {r1 + n1, r2 + n2, r3 + n3};
                const double rnrn = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
                const double rn_norm = sqrt(rnrn);
                const double B = B_decay(rn_norm, rnrn, e.kappa);
                const double C = C_decay(rn_norm, rnrn, e.kappa);
                for (int m = 0; m < 6; ++m) {
                    if (m < 3) T[m] = T[m] + B;
                    T[m] = T[m] - rn[ia[m]] * rn[ib[m]] * C;
                }
            }
        }
    }
    for (int m = 0; m < 6; ++m) out[(size_t)m * c.N + cc] = T[m];
}

// L(k) = Re FFT(real part) + reciprocal sum, times 14.399644/eps_inf; stored
// as 6 words per point, in FP64 and in the spectral type.
template <typename T>
__global__ __launch_bounds__(NT) void k_ewald_recip(Consts c, EwaldArgs e, Spectra<double> sp, double* Ltab,
                                                    T* Lt)
{
#pragma clang fp contract(off)
    const size_t kk = (size_t)blockIdx.x * NT + threadIdx.x;
    if (kk >= (size_t)c.Kh) return;
    int ix, iy, iz;
    if (kk < c.Mh) {
        ix = (int)(kk % c.Mx);
        const size_t r = kk / c.Mx;
        iy = (int)(r % c.Ly);
        iz = (int)(r / c.Ly);
    } else {
        const size_t r = kk - c.Mh;
        ix = c.Mx;
        iy = (int)(r % c.Ly);
        iz = (int)(r / c.Ly);
    }
    double k3 = (double)iz / c.Lz; if (k3 > 0.5) k3 = k3 - 1.0;
    double k2 = (double)iy / c.Ly; if (k2 > 0.5) k2 = k2 - 1.0;
    const double k1 = (double)ix / c.Lx;
    const int ia[6] = {0, 1, 2, 1, 2, 0}, ib[6] = {0, 1, 2, 2, 0, 1};
    for (int m = 0; m < 6; ++m) {
        double acc = sp.s[m][kk].x;
        for (int nz = -e.max_z; nz <= e.max_z; ++nz) {
            const double kG3 = e.two_pi_over_a * (k3 + nz);
            for (int ny = -e.max_y; ny <= e.max_y; ++ny) {
                const double kG2 = e.two_pi_over_a * (k2 + ny);
                for (int nx = -e.max_x; nx <= e.max_x; ++nx) {
                    if (ix == 0 && iy == 0 && iz == 0 && nx == 0 && ny == 0 && nz == 0) continue;
                    const double kG1 = e.two_pi_over_a * (k1 + nx);
                    const double kG[3] = {kG1, kG2, kG3};
                    const double kGkG = kG1 * kG1 + kG2 * kG2 + kG3 * kG3;
                    acc = acc + e.four_pi_over_v * kG[ia[m]] * kG[ib[m]] * exp(-kGkG / e.four_kappa2) / kGkG;
                }
            }
        }
        Ltab[kk * 6 + m] = acc * e.scale;
        if (Lt) Lt[kk * 6 + m] = (T)(acc * e.scale);
    }
}

// w reconstructed from the optimised w(k): slots 0..2 <- -i acouK
// (coord_module.F:457-469).
template <typename T>
__global__ void k_prep_w(Consts c, Spectra<T> sp, const cpx<double>* acouK)
{
    const size_t kk = (size_t)blockIdx.x * NT + threadIdx.x;
    if (kk >= (size_t)c.Kh) return;
    for (int q = 0; q < 3; ++q) {
        const cpx<double> a = acouK[(size_t)q * c.Kh + kk];
        sp.s[q][kk] = cpx<T>{(T)a.y, (T)(-a.x)};
    }
}

}  // namespace fhip
