// fhip_kernels.hpp -- the per-step kernels of the force-and-integrate path.
//
// Device layout (structure of arrays, unpadded, x fastest):
//   real fields   f[c],  c = x + Lx*(y + Ly*z)                 (one array per component)
//   spectra       packed half spectra of Mx = Lx/2 columns, s[(kz*Ly + ky)*Mx + kx],
//                 Mh = Mx*Ly*Lz = N/2 points: the real kx = 0 and kx = Lx/2
//                 values of every x row share column 0 as A[0] + i A[Lx/2]
//                 (the y and z passes describe how that column is carried).
//                 Rows of Mx points keep every tile of the y and z passes
//                 aligned to whole cache lines.
//   k arrays      L(k), the modulation spectrum and w(k) hold all
//                 (Mx+1)*Ly*Lz = Kh points of the half spectrum: kx < Mx at
//                 (kz*Ly + ky)*Mx + kx, kx = Mx at Mh + ky*Lz + kz.
// Spectral slots of one step:
//   0..2  u_x,u_y,u_z        -> (S + Z*^2 L) u   (dipole-dipole force before 1/N)
//   3..8  uu_xx..uu_xy       -> eta(k) = k (x) w(k)  (inhomogeneous strain)
//   9..11 w (acoustic MD)    -> dV/dw(k)
// With species of different Z* the dipole-dipole group splits as the
// reference does: u -> S u in slots 0..2 and Z*(R) u -> L (Z* u) in the three
// slots from zu_slot (9, or 12 with acoustic MD), multiplied by Z*(R) after
// the inverse transform.
// One step is: xfwd (drift + r2c along x) -> y forward -> z pass (forward,
// k-space stage, inverse) -> scalar stage -> y inverse -> xinv (c2r along x,
// local forces, kick).  Reductions go through per-block partial sums that a
// single workgroup folds in a fixed order, so every run is bitwise repeatable.
#pragma once
#include "fhip_fft.hpp"

namespace fhip {

constexpr int NT = 256;          // threads per workgroup in every kernel
constexpr int MAX_SLOTS = 15;    // spectral slots
enum PartialSlot {
    P_ESR = 0, P_ELR, P_EIS, P_EIC, P_EIM, P_EUNH, P_EDE, P_KE1, P_KE2, P_KEA, P_KEPK,
    P_AVG,                       // 12 slots: sums of u, u^2, Z* u, (Z* u)^2 over the counted cells
    P_NSLOT = P_AVG + 12
};

// Constants shared by all kernels (by value in the kernel arguments).
struct Consts {
    int Lx, Ly, Lz, Mx, N, Kh;
    size_t Mh;           // Mx*Ly*Lz: start of the z-major Nyquist column (Kh: the y-major one)
    int acoustic;        // mass_acou > 0: w is dynamical
    int efield;          // !E_field_always_zero
    int has_eloc;        // per-cell local field present
    int modulation;      // modulation_constant != 0 with a .modulation field
    int mir_lo, mir_src; // film/epit: layers iz >= mir_lo mirror layer mir_src - iz (mir_lo = Lz: none)
    int zsp;             // species with different Z*: Z*(R) u has its own transforms (dipoFFT.F:35-44)
    int zu_slot;         // spectral slot of (Z* u)_x when zsp
    double Ninv, dt, dt_2, a0, Z2, Zstar;
    double Zs[3];        // Z* per species (read_relaxor.F:35)
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
    double en_rep[13];               // energies and s at the end of the last step, kept when
    double s_rep;                    //   the next step's eq. 38 has already started
    // E-field schedule (integers held exactly as doubles): E_add of step i_step
    double e_amp[3];                 // amplitude (external_E_field)
    double e_type;                   // ESchedule
    double e_period, e_ntherm, e_navg;
    double i_step;                   // index of the next step
    double e_mir;                    // 1: the field of mirrored layers is mirrored too (film/epit)
};

// Time dependence of the uniform field.
enum ESchedule { ES_CONSTANT = 0, ES_TRI_COS = 1, ES_TRI_SIN = 2, ES_RAMP_ON = 3, ES_RAMP_OFF = 4, ES_HYSTERESIS = 5 };

// E_add at step i: wave() of wave_module.F:8-67 (triangular_cos/_sin,
// ramping_on/_off) and the linear sweep of hysteresis_loop.F:47-49, with the
// reference's integer and real operation order.  ES_CONSTANT leaves E alone.
__host__ __device__ inline void e_field_at(int type, const double amp[3], long long period, long long ntherm,
                                           long long navg, long long i_step, double out[3])
{
    if (type == ES_CONSTANT) return;
    if (amp[0] == 0.0 && amp[1] == 0.0 && amp[2] == 0.0) {   // wave_module.F:19-24
        for (int q = 0; q < 3; ++q) out[q] = amp[q];
        return;
    }
    switch (type) {
    case ES_TRI_COS:
    case ES_TRI_SIN: {
        long long i = i_step - ntherm;
        if (i < 0) i = 0;
        const long long m = (type == ES_TRI_COS) ? (i % period) : ((i + 3 * period / 4) % period);
        long long d = m - period / 2;
        if (d < 0) d = -d;
        const long long tmp = d - period / 4;
        for (int q = 0; q < 3; ++q) out[q] = amp[q] * (double)tmp / ((double)period / 4.0);
        return;
    }
    case ES_RAMP_ON:
        for (int q = 0; q < 3; ++q) out[q] = i_step < period ? amp[q] * (double)i_step / (double)period : amp[q];
        return;
    case ES_RAMP_OFF:
        for (int q = 0; q < 3; ++q)
            out[q] = i_step < period ? amp[q] * (double)(period - i_step) / (double)period : 0.0;
        return;
    case ES_HYSTERESIS:
        for (int q = 0; q < 3; ++q)
            out[q] = i_step >= ntherm ? amp[q] - 2.0 * amp[q] * (double)(i_step - ntherm) / (double)navg : amp[q];
        return;
    default:
        for (int q = 0; q < 3; ++q) out[q] = amp[q];
        return;
    }
}

__host__ __device__ inline void e_field_next(Scalars* sc)
{
    e_field_at((int)sc->e_type, sc->e_amp, (long long)sc->e_period, (long long)sc->e_ntherm,
               (long long)sc->e_navg, (long long)sc->i_step, sc->E_add);
    // initialize_E_field mirrors E_field (initialize_E_field.F:68-71); the
    // per-step updates of molecular-dynamics.F:66-76 and hysteresis_loop.F:47-58
    // rewrite it unmirrored
    const int type = (int)sc->e_type;
    const bool rewritten = type == ES_HYSTERESIS ? sc->i_step >= sc->e_ntherm
                                                 : (type != ES_CONSTANT && sc->i_step >= 1.0);
    if (rewritten) sc->e_mir = 0.0;
}
enum EnergyIdx {
    E_ACOU_KIN = 0, E_DIPO_KIN, E_SR, E_LR, E_DE, E_UNH, E_HS, E_HC, E_IS, E_IC, E_IM, E_TOTAL, E_HNP
};

// Per-cell fields.  The state (u, p, the pre-kicked p, dV/du and the
// acoustic w, p_w, dV/dw) is stored in S: FP64 in the plain mode, FP32 in the
// FP32 mode, where every update is computed in FP64 registers and rounded
// once on the store.  The local field and the dipoRavg sums stay FP64.
template <typename S>
struct FieldsT {
    S* u[3];
    S* p[3];
    S* pk[3];            // Nose-Poincare: momentum after the first half kick of the next step
    S* F[3];             // materialised dV/du
    S* w[3];
    S* pw[3];
    S* dVdw[3];          // materialised dV/dw
    double* eloc[3];
    double* uavg[3];     // c%dipoRavg: sum of u over the averaging steps
    const unsigned char* code;   // bits 0-1 species, bit 2 frozen (1/m = 0), bit 3 counted (sum_p)
};
using Fields = FieldsT<double>;

template <typename T>
struct Spectra {
    cpx<T>* s[MAX_SLOTS];
};

template <typename T>
__device__ __forceinline__ cpx<double> todbl(cpx<T> a) { return cpx<double>{(double)a.x, (double)a.y}; }

// Index of half-spectrum point (kx, ky, kz) in the k-space order.
__host__ __device__ __forceinline__ size_t kindex(int Mx, int Ly, int Lz, size_t Mh, int kx, int ky, int kz)
{
    return kx < Mx ? ((size_t)kz * Ly + ky) * Mx + kx : Mh + (size_t)ky * Lz + kz;
}

__device__ __forceinline__ int kindex32(const Consts& c, int kx, int ky, int kz)
{
    return kx < c.Mx ? (kz * c.Ly + ky) * c.Mx + kx : (int)c.Mh + ky * c.Lz + kz;
}

// One component of the k vector of index i along an axis of length L
// (elastic.F:83-89): folded to (-1/2, 1/2] except along x.
__host__ __device__ __forceinline__ double kcomp(int i, int L, double a0, bool fold)
{
    const double M2PI = 6.28318530717958647693;
    double f = (double)i / L;
    if (fold && f > 0.5) f = f - 1.0;
    return f * M2PI / a0;
}

// Workgroup sums of NS values: a butterfly across the 64 lanes of each wave,
// then the NT/64 wave sums in wave order, so that every run adds in the same
// order.  red holds at least NS * NT/64 doubles.
template <int NS>
__device__ __forceinline__ void block_sums(double (&v)[NS], double* red)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NS; ++q) red[q * NW + wv] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        double r = red[q * NW];
#pragma unroll
        for (int u = 1; u < NW; ++u) r += red[q * NW + u];
        v[q] = r;
    }
    __syncthreads();
}

__device__ __forceinline__ double block_sum(double v, double* red)
{
    double a[1] = {v};
    block_sums<1>(a, red);
    return a[0];
}

// ---------------------------------------------------------------------------
// xfwd: drift, then r2c along x of every forward field, two real fields per
// complex transform.
// ---------------------------------------------------------------------------
enum XfwdMode { XF_FORCE = 0, XF_NP = 1, XF_LF = 2, XF_VS = 3, XF_GENERIC = 4 };

// Film/epit (mirror.F:9-36): phase 1 drifts and transforms the rows below
// Consts::mir_lo, phase 2 then fills the mirrored rows from the updated
// source rows, (u_x, u_y, u_z) -> (-u_x, -u_y, u_z), and transforms them.
enum XfwdPhase { XP_ALL = 0, XP_SOURCE = 1, XP_MIRROR = 2 };

struct XfwdArgs {
    int mode;
    int nf;                       // number of forward fields
    int phase;                    // XfwdPhase
    const double* gen[12];        // XF_GENERIC: real input fields
};

// Rows per xinv workgroup: the x-transform tile (tile_of<LX>), cut for
// edges up to 32 to the fewest rows (a power of two) that give every thread
// at least one cell, so that a small supercell's x-inverse pass spreads over
// more workgroups (16^3: 16 instead of 4).
constexpr int xinv_rows_of(int LX)
{
    int rb = 1;
    while (2 * rb * LX <= 1024) rb *= 2;          // tile_of<LX>
    if (LX <= 32) {
        int small = 1;
        while (small * LX < NT) small *= 2;
        if (small < rb) rb = small;
    }
    return rb;
}

// Rows per xfwd workgroup: half the x-transform tile (tile_of<LX>).  Shorter
// drift-and-transform chains per workgroup overlap better (128^3: FP32
// 50 -> 43 us, FP64 61 -> 55 us); xinv keeps the full tile.
template <int LX>
struct xfwd_rows {
    static constexpr int value = tile_of<LX>::value > 1 ? tile_of<LX>::value / 2 : 1;
};

// Row stride of the x-pass LDS tiles (element x of row r at buf[x*ST + r]):
// RB + XPAD.
#ifndef FHIP_XPAD
#define FHIP_XPAD 1
#endif

// GEN: the body can run the generic set-up transform (XF_GENERIC, input
// fields a.gen); replica batches never do, and leaving it out keeps their
// argument copy out of scratch memory.
template <typename T, int LX, int RB = xfwd_rows<LX>::value, bool GEN = true>
__device__ __forceinline__ void xfwd_body(const Consts& c, const FieldsT<T>& f, const Spectra<T>& sp, const XfwdArgs& a,
                                             const Scalars* __restrict__ sc,
                                             const cpx<T>* twx, int bx)
{
    constexpr int ST = RB + FHIP_XPAD;
    constexpr int CELLS = RB * LX;
    constexpr int CPT = (CELLS + NT - 1) / NT;
    constexpr int MX = LX / 2;
    __shared__ cpx<T> buf[LX * ST];
    __shared__ cpx<T> tws[LX];
    const int tid = threadIdx.x;
    const int nrows = c.Ly * c.Lz;
    const int row0 = bx * RB;
    const int mrow = c.mir_lo * c.Ly;             // first mirrored row
    if (a.phase == XP_SOURCE && row0 >= mrow) return;       // uniform over the block
    if (a.phase == XP_MIRROR && row0 + RB <= mrow) return;
    for (int i = tid; i < LX; i += NT) tws[i] = twx[i];

    // the drifted u and w as stored (the state type), so the spectra are
    // those of the stored state; FP32 mode: half the registers
    T u[CPT][3], w[CPT][3];
    double zs[CPT];
    bool ok[CPT];
    int cell[CPT];                                  // N < 2^31 for every supported supercell
    // every cell's code, displacement and momentum are loaded before any
    // drift is stored: a cell's loads cannot move above the previous cell's
    // stores to u, so loading in the drift loop waits once per cell
    const bool pre = a.phase != XP_MIRROR && (a.mode == XF_NP || a.mode == XF_LF || a.mode == XF_VS);
    unsigned char cpre[CPT];
    T upre[CPT][3], ppre[CPT][3];
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
        const int e = tid + i * NT;
        const int r = e / LX, x = e - r * LX;
        ok[i] = (CELLS % NT == 0 || e < CELLS) && (row0 + r < nrows)
                && (a.phase == XP_ALL || ((row0 + r >= mrow) == (a.phase == XP_MIRROR)));
        cell[i] = (row0 + r) * LX + x;
        cpre[i] = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) upre[i][q] = ppre[i][q] = T(0);
        if (pre && ok[i]) {
            const size_t cc = cell[i];
            cpre[i] = f.code[cc];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                upre[i][q] = f.u[q][cc];
                ppre[i][q] = (a.mode == XF_NP ? f.pk[q] : f.p[q])[cc];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
        const int e = tid + i * NT;
        const int r = e / LX, x = e - r * LX;
        if (a.mode == XF_GENERIC) continue;
        zs[i] = 0.0;
        if (!ok[i]) {
            u[i][0] = u[i][1] = u[i][2] = 0.0;
            w[i][0] = w[i][1] = w[i][2] = 0.0;
            continue;
        }
        const size_t cc = cell[i];
        const unsigned char code = pre ? cpre[i] : f.code[cc];
        const int spc = code & 3;
        zs[i] = c.Zs[spc];
        const double minv = (code & 4) ? 0.0 : c.minv[spc];
        if (a.phase == XP_MIRROR) {
            const int rr = row0 + r, iz = rr / c.Ly, iy = rr - iz * c.Ly;
            const size_t cs = ((size_t)(c.mir_src - iz) * c.Ly + iy) * LX + x;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double uq = q < 2 ? -f.u[q][cs] : f.u[q][cs];         // mirror.F:18-24
                u[i][q] = uq;
                f.u[q][cc] = uq;
            }
            if (a.mode == XF_VS) {
#pragma unroll
                for (int q = 0; q < 3; ++q) f.p[q][cc] = f.p[q][cc] * sc->scale_p;
            }
        } else if (a.mode == XF_NP) {
            const double cd = sc->c_drift;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                double uq = upre[i][q] + cd * ppre[i][q] * minv;     // nose-poincare.F:41-53
                u[i][q] = uq;
                f.u[q][cc] = uq;
            }
        } else if (a.mode == XF_LF || a.mode == XF_VS) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                double pq = ppre[i][q];
                if (a.mode == XF_VS) {                                 // velocity_scaling.F:77-83
                    pq = pq * sc->scale_p;
                    f.p[q][cc] = pq;
                }
                double uq = upre[i][q] + c.dt * pq * minv;            // leapfrog.F:104-114
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
                    if (GEN && a.mode == XF_GENERIC) {
                        if constexpr (GEN) v = a.gen[fi][cell[i]];
                    } else if (fi < 3) {
                        v = fi == 0 ? u[i][0] : (fi == 1 ? u[i][1] : u[i][2]);
                    } else if (fi < 9) {                                 // dipoFFT.F:20-33
                        const int m = fi - 3;
                        const double ua = (m == 0 || m == 5) ? u[i][0] : (m == 1 || m == 3) ? u[i][1] : u[i][2];
                        const double ub = (m == 0 || m == 4) ? u[i][0] : (m == 1 || m == 5) ? u[i][1] : u[i][2];
                        v = ua * ub;
                    } else if (c.acoustic && fi < 12) {
                        const int q = fi - 9;
                        v = q == 0 ? w[i][0] : (q == 1 ? w[i][1] : w[i][2]);
                    } else {                                             // Z_u_R (dipoFFT.F:35-43)
                        const int q = fi - c.zu_slot;
                        v = zs[i] * (q == 0 ? u[i][0] : (q == 1 ? u[i][1] : u[i][2]));
                    }
                    if (h == 0) va = v; else vb = v;
                }
            }
            buf[x * ST + r] = cpx<T>{T(va), T(vb)};
        }
        __syncthreads();
        fft_lds<T, LX, RB, NT, -1, ST>(buf, tws, tid);
        // unpack Z = A + iB: A[k] = (Z[k] + conj Z[-k])/2, B[k] = (Z[k] - conj Z[-k])/(2i);
        // the real A[0] and A[Lx/2] are packed as A[0] + i A[Lx/2] at kx = 0
        for (int e = tid; e < RB * MX; e += NT) {
            const int r = e / MX, kx = e - r * MX;
            if (row0 + r >= nrows) continue;
            if (a.phase != XP_ALL && ((row0 + r >= mrow) != (a.phase == XP_MIRROR))) continue;
            const cpx<T> zk = buf[kx * ST + r];
            const cpx<T> zn = buf[(kx == 0 ? MX : LX - kx) * ST + r];
            const size_t o = (size_t)(row0 + r) * MX + kx;
            if (kx == 0) {
                sp.s[f0][o] = cpx<T>{zk.x, zn.x};
                if (two) sp.s[f0 + 1][o] = cpx<T>{zk.y, zn.y};
            } else {
                sp.s[f0][o] = cpx<T>{(zk.x + zn.x) * T(0.5), (zk.y - zn.y) * T(0.5)};
                if (two) sp.s[f0 + 1][o] = cpx<T>{(zk.y + zn.y) * T(0.5), (zn.x - zk.x) * T(0.5)};
            }
        }
        __syncthreads();
    }
}

template <typename T, int LX>
__global__ __launch_bounds__(NT) void k_xfwd(Consts c, FieldsT<T> f, Spectra<T> sp, XfwdArgs a,
                                             const Scalars* __restrict__ sc,
                                             const cpx<T>* __restrict__ twx)
{
    xfwd_body<T, LX>(c, f, sp, a, sc, twx, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Line transform along y, in place, one field per blockIdx.y: TB columns
// (z, kx) of stride Mx per tile.  The kx = 0 column carries C = P + iQ, the
// y transform of the packed real kx = 0 and kx = Lx/2 values.  The forward
// transform splits it into P(ky) at ky < Ly/2 and Q(ky) at Ly - ky, leaving
// P(0) + iQ(0) and P(Ly/2) + iQ(Ly/2) (all four real) packed at 0 and Ly/2;
// the inverse transform reassembles C.
// ---------------------------------------------------------------------------

// Tiles of the y and z passes hold 16 KB: 8 FP64 or 16 FP32 complex lines at
// N = 128, so that every tile row is a whole 128-byte cache line.
template <typename T>
constexpr int line_tile_cap() { return 16384 / (int)sizeof(cpx<T>); }

// Pairs (v at t, m at L - t) of a packed line: split C into P (t < L/2) or
// Q(L - t) (t > L/2), and the reverse.
template <typename T>
__device__ __forceinline__ cpx<T> split_pq(cpx<T> v, cpx<T> m, bool lower)
{
    return lower ? cpx<T>{(v.x + m.x) * T(0.5), (v.y - m.y) * T(0.5)}
                 : cpx<T>{(m.y + v.y) * T(0.5), (v.x - m.x) * T(0.5)};
}
template <typename T>
__device__ __forceinline__ cpx<T> join_pq(cpx<T> v, cpx<T> m, bool lower)
{
    return lower ? cpx<T>{v.x - m.y, v.y + m.x} : cpx<T>{m.x + v.y, v.x - m.y};
}

template <typename T, int LY, int SIGN, int TB = tile_of<LY, line_tile_cap<T>()>::value>
__device__ __forceinline__ void ypass_body(const Consts& c, const Spectra<T>& sp, int slot0,
                                              const cpx<T>* twy, int bx, int by)
{
    __shared__ cpx<T> buf[LY * TB];
    __shared__ cpx<T> tws[LY];
    const int tid = threadIdx.x;
    for (int i = tid; i < LY; i += NT) tws[i] = twy[i];
    const int l0 = bx * TB;
    const int nlines = c.Mx * c.Lz;
    cpx<T>* s = sp.s[slot0 + by];
    if constexpr (NT % TB == 0) {
        // a thread keeps one column b = tid % TB and every (NT / TB)-th row:
        // the column's (z, kx) is found once, and all its loads are issued
        // before the first LDS store
        constexpr int RPI = NT / TB;
        constexpr int IT = (LY + RPI - 1) / RPI;
        const int b = tid % TB, t0 = tid / TB, l = l0 + b;
        const bool okl = l < nlines;
        const int z = okl ? l / c.Mx : 0, kx = okl ? l - z * c.Mx : 1;
        const size_t col = (size_t)z * LY * c.Mx + kx;
        cpx<T> v[IT];
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int t = t0 + i * RPI;
            v[i] = cpx<T>{T(0), T(0)};
            if (okl && (LY % RPI == 0 || t < LY)) v[i] = s[col + (size_t)t * c.Mx];
        }
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int t = t0 + i * RPI;
            if (LY % RPI == 0 || t < LY) buf[t * TB + b] = v[i];
        }
        __syncthreads();
        if constexpr (SIGN > 0) {
            // a kx = 0 column carries P (t < Ly/2) and Q (Ly - t): rejoin
            // C = P + iQ pairwise in LDS (the pair's owner reads both rows
            // before writing either), only in tiles that hold such a column
            const int lend = min(l0 + TB, nlines) - 1;
            if (l0 % c.Mx == 0 || l0 / c.Mx != lend / c.Mx) {
                if (okl && kx == 0) {
                    for (int t = t0; 2 * t < LY; t += RPI) {
                        if (t == 0) continue;
                        const cpx<T> a = buf[t * TB + b], m = buf[(LY - t) * TB + b];
                        buf[t * TB + b] = join_pq(a, m, true);
                        buf[(LY - t) * TB + b] = join_pq(m, a, false);
                    }
                }
                __syncthreads();
            }
        }
        fft_lds<T, LY, TB, NT, SIGN>(buf, tws, tid);
        if (okl) {
#pragma unroll
            for (int i = 0; i < IT; ++i) {
                const int t = t0 + i * RPI;
                if (LY % RPI == 0 || t < LY) {
                    cpx<T> w = buf[t * TB + b];
                    if (SIGN < 0 && kx == 0 && t != 0 && 2 * t != LY)
                        w = split_pq(w, buf[(LY - t) * TB + b], 2 * t < LY);
                    s[col + (size_t)t * c.Mx] = w;
                }
            }
        }
    } else {
        for (int e = tid; e < LY * TB; e += NT) {
            const int t = e / TB, b = e - t * TB, l = l0 + b;
            if (l < nlines) {
                const int z = l / c.Mx, kx = l - z * c.Mx;
                const size_t col = (size_t)z * LY * c.Mx + kx;
                cpx<T> v = s[col + (size_t)t * c.Mx];
                if (SIGN > 0 && kx == 0 && t != 0 && 2 * t != LY)
                    v = join_pq(v, s[col + (size_t)(LY - t) * c.Mx], 2 * t < LY);
                buf[e] = v;
            }
        }
        __syncthreads();
        fft_lds<T, LY, TB, NT, SIGN>(buf, tws, tid);
        for (int e = tid; e < LY * TB; e += NT) {
            const int t = e / TB, b = e - t * TB, l = l0 + b;
            if (l < nlines) {
                const int z = l / c.Mx, kx = l - z * c.Mx;
                cpx<T> v = buf[e];
                if (SIGN < 0 && kx == 0 && t != 0 && 2 * t != LY) v = split_pq(v, buf[(LY - t) * TB + b], 2 * t < LY);
                s[((size_t)z * LY + t) * c.Mx + kx] = v;
            }
        }
    }
}

template <typename T, int LY, int SIGN>
__global__ __launch_bounds__(NT) void k_ypass(Consts c, Spectra<T> sp, int slot0,
                                              const cpx<T>* __restrict__ twy)
{
    ypass_body<T, LY, SIGN>(c, sp, slot0, twy, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// z pass: forward transform along z, k-space stage, inverse transform, on
// tiles of TB pencils (ky, kx) of stride Mx*Ly.  Pencils with kx = 0 carry,
// after the y pass, P(ky) (ky < Ly/2: the points (0, ky, kz)), Q(Ly - ky)
// (ky > Ly/2: the points (Lx/2, Ly - ky, kz)), or, at ky = 0 and Ly/2, R =
// P + iQ with both real before the z transform; the k-space stage splits R
// into P(kz) (kz < Lz/2) and Q(Lz - kz) (kz > Lz/2) and joins it back.
//
// Every stored point stands for itself and its mirror partner -k.  The
// reference evaluates both on its half spectrum and its c2r keeps the
// Hermitian part; where the partner's k vector is not exactly -k (the
// kx = Lx/2 plane, ky = Ly/2, kz = Lz/2) the strain stage is evaluated at
// both and averaged, elsewhere once with the partner's energy share.
// ---------------------------------------------------------------------------
// Loads of data that the step reads once and does not need again soon: the
// L(k) table (read once per step) and the spectra the x-inverse pass consumes
// (rewritten by the next step before anyone reads them).  In FP64 the rest of
// the step's working set (u, the pre-kicked p and the nine spectra, 250 MB at
// 128^3) is about the size of the device's 256 MB last-level cache, so these
// loads are nontemporal and leave that set cached (FP64 128^3: x passes
// 4-5 us faster each).  FP32 state fits the cache whole and keeps ordinary
// loads, which are faster there.
template <typename T>
__device__ __forceinline__ T once_load(const T* p)
{
    if constexpr (sizeof(T) == 8) return __builtin_nontemporal_load(p);
    else return *p;
}
template <typename T>
__device__ __forceinline__ cpx<T> once_load(const cpx<T>* p)
{
    return cpx<T>{once_load(&p->x), once_load(&p->y)};
}

enum ZGroup { ZG_PLAIN = 0, ZG_DD = 1, ZG_STRAIN = 2, ZG_ACOU = 3, ZG_DDZ = 4 };

template <typename T>
struct ZArgs {
    const T* Ltab;                   // [Kh][6] long-range kernel L(k): xx yy zz yz zx xy
    const cpx<double>* csx;          // cos/sin tables (2 pi i / L) per axis
    const cpx<double>* csy;
    const cpx<double>* csz;
    const double* kvx;               // k components per axis (kcomp), Mx + 1 along x
    const double* kvy;
    const double* kvz;
    const cpx<double>* modk;         // FFT of the modulation field on the k array (or null)
    cpx<double>* acouK;              // materialised optimised w(k) on the k array (or null)
    double* partials;                // [P_NSLOT][maxblocks]
    int maxblocks;
    int nplain;                      // ZG_PLAIN: number of slots
    int plain_sign;                  // ZG_PLAIN: -1 forward, +1 inverse
    int split;                       // pair kernel: workgroup 2w + g runs group g of tile w alone
};

// coup2(a, i) (elastic.F:148-175)
template <typename KT>
__device__ __forceinline__ KT coup2(const Consts& c, const KT k[3], int a, int i)
{
    if (i < 3) return (i == a) ? k[a] * KT(c.B1xx) / 2 : k[a] * KT(c.B1yy) / 2;
    // ki = ((0,3,2),(3,0,1),(2,1,0)) by column: component of k multiplying w_a in eta_{i}
    const int m = i - 3;
    int q;
    if (m == 0) q = (a == 0) ? -1 : (a == 1 ? 2 : 1);
    else if (m == 1) q = (a == 0) ? 2 : (a == 1 ? -1 : 0);
    else q = (a == 0) ? 1 : (a == 1 ? 0 : -1);
    return (q < 0 ? KT(0) : k[q]) * KT(c.B4yz);
}

// Modes of one evaluation (see kspace_elem): N, the point itself; C, its
// mirror partner on the conjugate input, conjugated back; RP / RQ, the real
// P or Q of an R pencil at kz = 0 or Lz/2, each its own partner.
enum { EV_N = 0, EV_C = 1, EV_RP = 2, EV_RQ = 3 };

// One evaluation of the k-space stage at the point (kx, ky, kz), k-array
// index kk, arithmetic in KT, on the NF inputs X; outputs to O.  The point's
// energies are added with weight wm times that of a point off the kx = 0
// and kx = Lx/2 planes, and w(k) goes to acouK when it is materialised
// (with -conj w at kkp, the implied partner, when kkp >= 0).
// Table values the k-space stage reads at one point: the per-axis cos/sin
// and the L(k) entries of the dipole-dipole groups, the k vector of the
// strain and acoustic groups.  A z pass that prefetches the next group's
// tile loads them for all its points before that prefetch is issued, since
// vector-memory loads complete in issue order and a table load issued after
// the prefetch would wait for the whole prefetch.
template <typename KT>
struct KTab {
    KT cx, cy, cz, sx, sy, sz;
    KT L[6];
    KT k[3];
    cpx<KT> mk;                   // modulation spectrum at the point (0 without modulation)
};

// MOD: 0 no modulation spectrum, 1 one is applied, -1 c.modulation decides.
template <typename KT, int GROUP, typename T, int MOD = -1>
__device__ __forceinline__ KTab<KT> ktab_load(const Consts& c, const ZArgs<T>& a, int kx, int ky, int kz, int kk)
{
    KTab<KT> tb;
    if constexpr (GROUP == ZG_DD || GROUP == ZG_DDZ) {
        const cpx<double> ex = a.csx[kx], ey = a.csy[ky], ez = a.csz[kz];
        tb.cx = KT(ex.x); tb.cy = KT(ey.x); tb.cz = KT(ez.x);
        tb.sx = KT(ex.y); tb.sy = KT(ey.y); tb.sz = KT(ez.y);
        const T* lt = a.Ltab + kk * 6;
        // one Z*: the table is applied as Z*^2 L(k)
        const KT Z2 = KT(c.Z2);
#pragma unroll
        for (int m = 0; m < 6; ++m) tb.L[m] = GROUP == ZG_DD ? KT(once_load(lt + m)) * Z2 : KT(once_load(lt + m));
    } else {
        tb.k[0] = KT(a.kvx[kx]);
        tb.k[1] = KT(a.kvy[ky]);
        tb.k[2] = KT(a.kvz[kz]);
        tb.mk = cpx<KT>{KT(0), KT(0)};
        if (MOD == 1 || (MOD < 0 && c.modulation)) {
            const cpx<double> m0 = a.modk[kk];
            tb.mk = cpx<KT>{KT(m0.x), KT(m0.y)};
        }
    }
    return tb;
}

template <typename KT, int GROUP, typename T, int NF>
__device__ __forceinline__ void keval(const Consts& c, const ZArgs<T>& a, Scalars* sc, int kx, int ky, int kz,
                                      int kk, int kkp, const KTab<KT>& tb, const cpx<KT> (&X)[NF], cpx<KT> (&O)[NF],
                                      double wm, double& acc0, double& acc1, double& acc2)
{
    auto put = [&](int q, cpx<KT> o) { O[q] = o; };
    if constexpr (GROUP == ZG_DD) {
        // S(k): Fourier transform of the short-range couplings
        // (dipole-dipole-short-range.F:35-83) in closed form.
        const KT cx = tb.cx, cy = tb.cy, cz = tb.cz, sx = tb.sx, sy = tb.sy, szz = tb.sz;
        const KT j1 = KT(c.j[0]), j2 = KT(c.j[1]), j3 = KT(c.j[2]), j4 = KT(c.j[3]), j5 = KT(c.j[4]),
                 j6 = KT(c.j[5]), j7 = KT(c.j[6]);
        const KT k2t = KT(2.0 * c.kappa2[0]);
        const KT cxyz = cx * cy * cz;
        KT S[6];
        S[0] = k2t + 2 * j2 * cx + 2 * j1 * (cy + cz) + 4 * j3 * cx * (cy + cz) + 4 * j4 * cy * cz + 8 * j6 * cxyz;
        S[1] = k2t + 2 * j2 * cy + 2 * j1 * (cz + cx) + 4 * j3 * cy * (cz + cx) + 4 * j4 * cz * cx + 8 * j6 * cxyz;
        S[2] = k2t + 2 * j2 * cz + 2 * j1 * (cx + cy) + 4 * j3 * cz * (cx + cy) + 4 * j4 * cx * cy + 8 * j6 * cxyz;
        S[3] = -4 * j5 * sy * szz - 8 * j7 * sy * szz * cx;   // yz
        S[4] = -4 * j5 * szz * sx - 8 * j7 * szz * sx * cy;   // zx
        S[5] = -4 * j5 * sx * sy - 8 * j7 * sx * sy * cz;     // xy
        const KT (&L)[6] = tb.L;
        const cpx<KT> ux = X[0], uy = X[1], uz = X[2];
        // F = (S + Z^2 L) u ; energies via Parseval
        cpx<KT> fs[3], fl[3];
        fs[0] = {S[0] * ux.x + S[5] * uy.x + S[4] * uz.x, S[0] * ux.y + S[5] * uy.y + S[4] * uz.y};
        fs[1] = {S[5] * ux.x + S[1] * uy.x + S[3] * uz.x, S[5] * ux.y + S[1] * uy.y + S[3] * uz.y};
        fs[2] = {S[4] * ux.x + S[3] * uy.x + S[2] * uz.x, S[4] * ux.y + S[3] * uy.y + S[2] * uz.y};
        fl[0] = {L[0] * ux.x + L[5] * uy.x + L[4] * uz.x, L[0] * ux.y + L[5] * uy.y + L[4] * uz.y};
        fl[1] = {L[5] * ux.x + L[1] * uy.x + L[3] * uz.x, L[5] * ux.y + L[1] * uy.y + L[3] * uz.y};
        fl[2] = {L[4] * ux.x + L[3] * uy.x + L[2] * uz.x, L[4] * ux.y + L[3] * uy.y + L[2] * uz.y};
        KT es = KT(0), el = KT(0);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            es += X[q].x * fs[q].x + X[q].y * fs[q].y;
            el += X[q].x * fl[q].x + X[q].y * fl[q].y;
            put(q, cpx<KT>{fs[q].x + fl[q].x, fs[q].y + fl[q].y});
        }
        acc0 += 2.0 * wm * (double)es;
        acc1 += 2.0 * wm * (double)el;
    } else if constexpr (GROUP == ZG_DDZ) {
        // species with different Z*: S(k) u^ and L(k) (Z* u)^ separately
        // (dVddi-dipole-dipole.F:20-30, 40-50), E_sr and E_lr by Parseval
        const KT cx = tb.cx, cy = tb.cy, cz = tb.cz, sx = tb.sx, sy = tb.sy, szz = tb.sz;
        const KT j1 = KT(c.j[0]), j2 = KT(c.j[1]), j3 = KT(c.j[2]), j4 = KT(c.j[3]), j5 = KT(c.j[4]),
                 j6 = KT(c.j[5]), j7 = KT(c.j[6]);
        const KT k2t = KT(2.0 * c.kappa2[0]);
        const KT cxyz = cx * cy * cz;
        KT S[6];
        S[0] = k2t + 2 * j2 * cx + 2 * j1 * (cy + cz) + 4 * j3 * cx * (cy + cz) + 4 * j4 * cy * cz + 8 * j6 * cxyz;
        S[1] = k2t + 2 * j2 * cy + 2 * j1 * (cz + cx) + 4 * j3 * cy * (cz + cx) + 4 * j4 * cz * cx + 8 * j6 * cxyz;
        S[2] = k2t + 2 * j2 * cz + 2 * j1 * (cx + cy) + 4 * j3 * cz * (cx + cy) + 4 * j4 * cx * cy + 8 * j6 * cxyz;
        S[3] = -4 * j5 * sy * szz - 8 * j7 * sy * szz * cx;
        S[4] = -4 * j5 * szz * sx - 8 * j7 * szz * sx * cy;
        S[5] = -4 * j5 * sx * sy - 8 * j7 * sx * sy * cz;
        const KT (&L)[6] = tb.L;
        auto sym = [&](const KT (&M)[6], const cpx<KT>* v, cpx<KT>* o) {
            o[0] = {M[0] * v[0].x + M[5] * v[1].x + M[4] * v[2].x, M[0] * v[0].y + M[5] * v[1].y + M[4] * v[2].y};
            o[1] = {M[5] * v[0].x + M[1] * v[1].x + M[3] * v[2].x, M[5] * v[0].y + M[1] * v[1].y + M[3] * v[2].y};
            o[2] = {M[4] * v[0].x + M[3] * v[1].x + M[2] * v[2].x, M[4] * v[0].y + M[3] * v[1].y + M[2] * v[2].y};
        };
        cpx<KT> fs[3], fl[3];
        sym(S, X, fs);
        sym(L, X + 3, fl);
        KT es = KT(0), el = KT(0);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            es += X[q].x * fs[q].x + X[q].y * fs[q].y;
            el += X[3 + q].x * fl[q].x + X[3 + q].y * fl[q].y;
            put(q, fs[q]);
            put(3 + q, fl[q]);
        }
        acc0 += 2.0 * wm * (double)es;
        acc1 += 2.0 * wm * (double)el;
    } else {
        const KT (&k)[3] = tb.k;
        const bool k0 = (kx == 0 && ky == 0 && kz == 0);
        if (k0) {
#pragma unroll
            for (int m = 0; m < 6; ++m) sc->dc[m] = (double)X[m].x;
        }
        // coupling rhs: c_mod + C uu  (optimize-inho-strain.F:36-42, dVdac.F:46-52)
        cpx<KT> cm[3], Cu[3];
        const cpx<KT> mk = tb.mk;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const KT fq = KT(c.mod_c) * k[q];
            cm[q] = c.modulation ? cpx<KT>{fq * mk.x, fq * mk.y} : cpx<KT>{KT(0), KT(0)};
            KT rr = KT(0), ri = KT(0);
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                const KT cq = coup2<KT>(c, k, q, m);
                rr = rr + cq * X[m].x;
                ri = ri + cq * X[m].y;
            }
            Cu[q] = {rr, ri};
        }
        // K(k) (elastic.F:90-104)
        const KT B11 = KT(c.B11), B12 = KT(c.B12), B44 = KT(c.B44);
        KT K[3][3];
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) {
            const int o1 = (bb + 2) % 3, o2 = (bb + 1) % 3;
#pragma unroll
            for (int aa = 0; aa < 3; ++aa) {
                if (aa == bb)
                    K[aa][bb] = B11 * k[aa] * k[aa] + B44 * k[o1] * k[o1] + B44 * k[o2] * k[o2];
                else
                    K[aa][bb] = B12 * k[aa] * k[bb] + B44 * k[aa] * k[bb];
            }
        }
        cpx<KT> w[3];
        if constexpr (GROUP == ZG_STRAIN) {
            if (k0) {
                w[0] = w[1] = w[2] = cpx<KT>{KT(0), KT(0)};
            } else {
                // w = -K^-1 (c_mod + C uu) (optimize-inho-strain.F:36-57,
                // which solves it with the Cholesky factor of elastic.F:131-141):
                // here through the adjugate of the symmetric 3 x 3 K(k), one
                // reciprocal of its determinant instead of three square roots
                // and three divisions on the critical path.  K(k) is positive
                // definite for k != 0 and well-conditioned (eigenvalues between
                // about B44 |k|^2 and B11 |k|^2 for a cubic elastic tensor), so
                // both solves agree to rounding.
                const KT a00 = K[1][1] * K[2][2] - K[1][2] * K[1][2];
                const KT a01 = K[0][2] * K[1][2] - K[0][1] * K[2][2];
                const KT a02 = K[0][1] * K[1][2] - K[0][2] * K[1][1];
                const KT a11 = K[0][0] * K[2][2] - K[0][2] * K[0][2];
                const KT a12 = K[0][1] * K[0][2] - K[0][0] * K[1][2];
                const KT a22 = K[0][0] * K[1][1] - K[0][1] * K[0][1];
                const KT det = K[0][0] * a00 + K[0][1] * a01 + K[0][2] * a02;
                const KT id = KT(-1) / det;
                cpx<KT> r[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) r[q] = {cm[q].x + Cu[q].x, cm[q].y + Cu[q].y};
                w[0] = {(a00 * r[0].x + a01 * r[1].x + a02 * r[2].x) * id, (a00 * r[0].y + a01 * r[1].y + a02 * r[2].y) * id};
                w[1] = {(a01 * r[0].x + a11 * r[1].x + a12 * r[2].x) * id, (a01 * r[0].y + a11 * r[1].y + a12 * r[2].y) * id};
                w[2] = {(a02 * r[0].x + a12 * r[1].x + a22 * r[2].x) * id, (a02 * r[0].y + a12 * r[1].y + a22 * r[2].y) * id};
            }
        } else {
            // acoustic MD: w(k) = i FFT(w) (acouFFT.F:19-24);
            // dV/dw = -i [c_mod + C uu + K w] (dVdac.F:44-54)
#pragma unroll
            for (int q = 0; q < 3; ++q) w[q] = {-X[6 + q].y, X[6 + q].x};
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                KT kr = KT(0), ki = KT(0);
#pragma unroll
                for (int bb = 0; bb < 3; ++bb) { kr += K[q][bb] * w[bb].x; ki += K[q][bb] * w[bb].y; }
                const KT vr = ((KT(0) + cm[q].x) + Cu[q].x) + kr;
                const KT vi = ((KT(0) + cm[q].y) + Cu[q].y) + ki;
                put(6 + q, cpx<KT>{vi, -vr});
            }
        }
        if constexpr (GROUP == ZG_STRAIN) {
            if (a.acouK) {
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    a.acouK[(size_t)q * c.Kh + kk] = cpx<double>{(double)w[q].x, (double)w[q].y};
                    if (kkp >= 0)                     // w(-k) = -conj w(k)
                        a.acouK[(size_t)q * c.Kh + kkp] = cpx<double>{-(double)w[q].x, (double)w[q].y};
                }
            }
        }
        // energies with weight 1/2 on the kx = 0 and kx = Lx/2 planes
        // (energy_module.F:162-239)
        // FP32 mode with optimised strain: w solves K w = -(c_mod + C uu), so
        // Re(w^H K w) = -(E_ic + E_im) per point and the quadratic form is
        // not evaluated; FP64 keeps the reference's sum (energy_module.F:162-177)
        constexpr bool EIS_BY_SOLVE = sizeof(KT) == 4 && GROUP == ZG_STRAIN;
        KT eis = KT(0), eic = KT(0), eim = KT(0);
#pragma unroll
        for (int aa = 0; aa < 3; ++aa) {
            if constexpr (!EIS_BY_SOLVE) {
#pragma unroll
                for (int bb = 0; bb < 3; ++bb)
                    eis += K[aa][bb] * (w[aa].x * w[bb].x + w[aa].y * w[bb].y);
            }
            eic += w[aa].x * Cu[aa].x + w[aa].y * Cu[aa].y;
            eim += w[aa].x * cm[aa].x + w[aa].y * cm[aa].y;
        }
        if constexpr (EIS_BY_SOLVE) eis = -(eic + eim);
        acc0 += wm * (double)eis;
        acc1 += wm * (double)eic;
        acc2 += wm * (double)eim;
        // eta(k) = k (x) w (dVddi_inho_coupling.F:41-47)
        put(0, cpx<KT>{k[0] * w[0].x, k[0] * w[0].y});
        put(1, cpx<KT>{k[1] * w[1].x, k[1] * w[1].y});
        put(2, cpx<KT>{k[2] * w[2].x, k[2] * w[2].y});
        put(3, cpx<KT>{k[2] * w[1].x + k[1] * w[2].x, k[2] * w[1].y + k[1] * w[2].y});
        put(4, cpx<KT>{k[0] * w[2].x + k[2] * w[0].x, k[0] * w[2].y + k[2] * w[0].y});
        put(5, cpx<KT>{k[1] * w[0].x + k[0] * w[1].x, k[1] * w[0].y + k[0] * w[1].y});
    }
}

// Evaluation plan of a point of a kx = 0 pencil (slot s, kz index t): one
// evaluation, or two where the partner's k vector is not exactly -k.
struct KPlan {
    int nev, x0, y0, z0, m0, x1, y1, z1, m1;
    double wm;                    // energy weight of each evaluation
    int yp, zp;                   // implied partner (0, yp, zp) of a single evaluation (yp < 0: none)
};

// Input of an evaluation in mode md, and its output folded into the stored value.
template <typename KT>
__device__ __forceinline__ cpx<KT> ev_in(cpx<KT> x, int md)
{
    return md == EV_N ? x : md == EV_C ? cpx<KT>{x.x, -x.y} : md == EV_RP ? cpx<KT>{x.x, KT(0)} : cpx<KT>{x.y, KT(0)};
}
template <typename KT>
__device__ __forceinline__ cpx<KT> ev_out(cpx<KT> acc, cpx<KT> o, int md, KT half)
{
    if (md == EV_N) return {acc.x + half * o.x, acc.y + half * o.y};
    if (md == EV_C) return {acc.x + half * o.x, acc.y - half * o.y};
    if (md == EV_RP) return {o.x, acc.y};
    return {acc.x, o.x};
}

// The dipole-dipole group evaluates both points of every pair: the reference
// applies L(k) at each point of the kx = 0 and kx = Lx/2 planes of its half
// spectrum from that point's own table entry (dVddi-dipole-dipole.F:40-50),
// and the table is symmetric only up to the asymmetric image ranges of the
// real-space Ewald sum (dipole-dipole-long-range.F:41-91), so L(-k) is not
// L(k) to rounding on anisotropic supercells; the c2r keeps the average.
// The strain group needs the second evaluation only where the partner's
// folded k vector is not exactly -k.
template <bool DD>
__device__ __forceinline__ KPlan kplan(const Consts& c, int s, int t)
{
    const int Ly = c.Ly, Lz = c.Lz, Mx = c.Mx;
    const int tp = (t == 0) ? 0 : Lz - t;          // kz of the mirror partner
    KPlan p{1, 0, s, t, EV_N, 0, 0, 0, EV_C, 1.0, -1, 0};
    if (s != 0 && 2 * s < Ly) {                    // P pencil: (0, s, t)
        if (DD || 2 * t == Lz) { p.nev = 2; p.y1 = Ly - s; p.z1 = tp; p.wm = 0.5; }
        else { p.yp = Ly - s; p.zp = tp; }
    } else if (2 * s > Ly) {                       // Q pencil: (Lx/2, Ly - s, t)
        p.x0 = Mx; p.y0 = Ly - s;
        p.nev = 2; p.x1 = Mx; p.y1 = s; p.z1 = tp; p.wm = 0.5;
    } else if (t == 0 || 2 * t == Lz) {            // R pencil, real P and Q
        p.nev = 2; p.m0 = EV_RP; p.m1 = EV_RQ; p.x1 = Mx; p.y1 = s; p.z1 = t; p.wm = 0.5;
    } else if (2 * t < Lz) {                       // R pencil, P(t): (0, s, t)
        if (DD || s != 0) { p.nev = 2; p.y1 = s; p.z1 = tp; p.wm = 0.5; }
        else { p.yp = s; p.zp = tp; }
    } else {                                       // R pencil, Q(Lz - t): (Lx/2, s, Lz - t)
        p.x0 = Mx; p.z0 = Lz - t;
        p.nev = 2; p.x1 = Mx; p.y1 = s; p.z1 = t; p.wm = 0.5;
    }
    return p;
}

// Arithmetic type of the k-space stage: FP64 with FP64 spectra, FP32 with FP32.
template <typename T> struct kspace_type { using type = double; };
template <> struct kspace_type<float> { using type = float; };

// z-pass tile geometry: TB pencils of LZ points per tile (at most 64, which
// bounds the kx = 0 pencils a tile holds on small supercells), LDS rows of
// TB+1 (so that the line-order reads of the k-space stage hit every bank),
// and the side buffer of the kx = 0 pencils, two NF-field values per point,
// that the host sizes as dynamic LDS.
constexpr int ZTILE_MAX = 64;
constexpr int ZSTRAIN_FB = 3;   // strain fields transformed together (DD and acoustic groups: 3)

// Pencils per z tile for an edge LZ and complex points of CB bytes (host and
// device): the y pass's 16 KB tile, at most ZTILE_MAX pencils.  Edges of 32
// and less take instead the fewest pencils that give every thread the same
// number of points (LZ TB a multiple of NT) in full 128-byte rows, so that a
// small supercell's z pass spreads over more workgroups (16^3 FP64: 8
// instead of 2); edges with a factor 3 or 5 keep the larger tile, which no
// smaller one divides evenly among the threads.
constexpr int ztile_pencils(int LZ, int CB)
{
    int tb = 1;
    while (2 * tb * LZ <= 16384 / CB) tb *= 2;          // tile_of<LZ, 16 KB>
    if (tb > ZTILE_MAX) tb = ZTILE_MAX;
    if (LZ <= 32) {
        int small = 128 / CB;                            // one 128-byte row
        while (small < tb && (small * LZ) % NT != 0) small *= 2;
        if (small < tb) tb = small;
    }
    return tb;
}

template <typename T, int LZ>
struct ZTile {
    static constexpr int TB = ztile_pencils(LZ, (int)sizeof(cpx<T>));
    static constexpr int ST = TB + 1;
    static constexpr int EL = LZ * TB;
    static constexpr int PPT = (EL + NT - 1) / NT;
    // the first and last pass of each z transform run in registers (see zpass_body)
    static constexpr bool REG = NT % TB == 0 && EL % NT == 0 && dft_radix(PPT) && LZ % (PPT * PPT) == 0;
};

// Thread-to-element maps of a z tile (TB pencils of LZ points, NTH threads),
// host and device: the kernels index through them and the host-side bounds
// check enumerates every (thread, iteration) pair through the same code.
template <int NTH, int TB, int LZ>
struct ZMap {
    static constexpr int EL = LZ * TB;
    static constexpr int PPT = (EL + NTH - 1) / NTH;
    // iteration i of thread tid holds element tid + i*NTH of the tile
    __host__ __device__ static constexpr bool live(int tid, int i) { return EL % NTH == 0 || tid + i * NTH < EL; }
    // element e in HBM order (pencil b fastest, coalesced): kz index t
    __host__ __device__ static void hbm_e(int e, int& t, int& b)
    {
        t = e / TB;
        b = e - t * TB;
    }
    // element tid + i*NTH in HBM order: a per-thread base plus a uniform step in i
    __host__ __device__ static void hbm(int tid, int i, int& t, int& b)
    {
        if constexpr (NTH % TB == 0) {
            b = tid % TB;
            t = tid / TB + i * (NTH / TB);
        } else {
            hbm_e(tid + i * NTH, t, b);
        }
    }
    // element tid + i*NTH in line order (t fastest), for the k-space stage
    __host__ __device__ static void line(int tid, int i, int& t, int& b)
    {
        if constexpr (NTH % LZ == 0) {
            t = tid % LZ;
            b = tid / LZ + i * (NTH / LZ);
        } else if constexpr (LZ % NTH == 0) {
            b = i / (LZ / NTH);
            t = (i % (LZ / NTH)) * NTH + tid;
        } else {
            const int e = tid + i * NTH;
            b = e / LZ;
            t = e - b * LZ;
        }
    }
};

// Fields of a z group, fields transformed together (FB tiles side by side in
// LDS, so that each Stockham pass has FB times the butterflies between two
// barriers) and the LDS row stride that follows.
template <int GROUP>
struct ZGroupShape {
    static constexpr int NF = (GROUP == ZG_DD) ? 3 : (GROUP == ZG_STRAIN || GROUP == ZG_DDZ ? 6 : (GROUP == ZG_ACOU ? 9 : 1));
    static constexpr int F0 = (GROUP == ZG_DD || GROUP == ZG_DDZ || GROUP == ZG_PLAIN) ? 0 : 3;
    static constexpr int FB = (GROUP == ZG_PLAIN) ? 1 : (GROUP == ZG_STRAIN ? ZSTRAIN_FB : 3);
};

// LDS of a z-pass workgroup, declared by the kernel (one pool when two groups
// run in the same workgroup).
template <typename T, int LZ, int NBUF>
struct ZLds {
    cpx<T> buf[NBUF];
    cpx<T> tws[LZ];
    double red[NT];
    int lkx[ZTile<T, LZ>::TB], lky[ZTile<T, LZ>::TB], lord[ZTile<T, LZ>::TB];
};

// Spectral slot of field fi of a z group (the split dipole-dipole group reads
// u^ from slots 0..2 and (Z* u)^ from zu_slot on).
template <int GROUP>
__device__ __forceinline__ int zslot(const Consts& c, int fi)
{
    return (GROUP == ZG_DDZ && fi >= 3) ? c.zu_slot + fi - 3 : ZGroupShape<GROUP>::F0 + fi;
}

// Prefetch of the second group's first ZPF field tiles in a paired z pass.
// While the first group evaluates its k-space stage, the LDS pool of the
// transforms is idle; the tiles of the next group's first field batch are
// copied into it from HBM by LDS-DMA loads (global_load_lds_dwordx4: each
// lane moves 16 bytes, a wave's 64 lanes fill 1 KB of LDS from a uniform
// base), which hold no registers while in flight.  Before the first group's
// inverse transforms take the pool back the tiles move into registers, in
// the layout the ordinary loads give, and the next group starts from them.
// Staging layout: field f, element (t, b) of the tile at f*EL + t*TB + b
// (HBM order, unpadded).  The 16-byte chunks stay inside one tile row (TB a
// multiple of CH); a chunk that would reach past the last pencil of the
// spectrum (the last tile when Mx*Ly is not a multiple of TB) or that does
// not start on 16 bytes is not copied, and its pencils below that end are
// read from HBM when the tiles move into registers.
constexpr int ZPF = 3;
template <typename T, int LZ>
struct ZPrefetch {
    static constexpr int TB = ZTile<T, LZ>::TB, EL = ZTile<T, LZ>::EL, PPT = ZTile<T, LZ>::PPT;
    static constexpr int CH = 16 / (int)sizeof(cpx<T>);     // tile elements per 16-byte chunk
    static constexpr int NCH = EL / CH;                      // chunks per field tile
    static constexpr bool OK = (16 % sizeof(cpx<T>)) == 0 && TB % CH == 0;
};

// Whether the chunk of pencil b (its first pencil) of tile w is copied: it
// lies inside the spectrum, and its rows start on 16 bytes (Mx*Ly a multiple
// of CH, which FP32 spectra with odd Mx and Ly are not).
template <typename T, int LZ>
__host__ __device__ __forceinline__ bool zchunk_ok(int w, int b0, int nlines)
{
    constexpr int TB = ZPrefetch<T, LZ>::TB, CH = ZPrefetch<T, LZ>::CH;
    return w * TB + b0 + CH <= nlines && nlines % CH == 0;
}

template <typename T, int LZ, int GROUP>
__device__ __forceinline__ void zprefetch_issue(const Consts& c, const Spectra<T>& sp, int w, cpx<T>* stage)
{
    using P = ZPrefetch<T, LZ>;
    constexpr int TB = P::TB, EL = P::EL, CH = P::CH, NCH = P::NCH;
    static_assert(P::OK, "prefetch needs 16-byte chunks inside a tile row");
    const int nlines = c.Mx * c.Ly;
    const size_t zstride = (size_t)nlines;
    const int tid = threadIdx.x;
    const int wbase = tid & ~63;                             // first thread of this wave
#pragma unroll
    for (int f = 0; f < ZPF; ++f) {
        const cpx<T>* src = sp.s[zslot<GROUP>(c, f)];
#pragma unroll
        for (int it = 0; it < (NCH + NT - 1) / NT; ++it) {
            const int j = it * NT + tid;
            if (NCH % NT == 0 || j < NCH) {
                const int e = j * CH, t = e / TB, b = e - t * TB;
                // every lane issues its copy (a chunk past the end copies the
                // field's first chunk instead, and is not used), so the number
                // of loads in flight after the prefetch is known to the compiler
                // and later waits need not drain it
                const size_t in = zchunk_ok<T, LZ>(w, b, nlines) ? 1 : 0;
                const cpx<T>* g = src + ((size_t)(w * TB + b) + (size_t)t * zstride) * in;
                cpx<T>* l = stage + f * EL + (it * NT + wbase) * CH;   // wave-uniform LDS base
                __builtin_amdgcn_global_load_lds((const void*)g, (__attribute__((address_space(3))) void*)l, 16, 0, 0);
            }
        }
    }
}

// Wait for this wave's LDS-DMA loads (vmcnt 0; expcnt and lgkmcnt left free).
__device__ __forceinline__ void wait_vmem_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

template <typename T, int LZ, int GROUP>
__device__ __forceinline__ void zprefetch_collect(const Consts& c, const Spectra<T>& sp, int w, const cpx<T>* stage,
                                                  cpx<T> (&pre)[ZPF][ZTile<T, LZ>::PPT])
{
    using P = ZPrefetch<T, LZ>;
    constexpr int TB = P::TB, EL = P::EL, PPT = P::PPT, CH = P::CH;
    using Map = ZMap<NT, TB, LZ>;
    const int nlines = c.Mx * c.Ly;
    const int tid = threadIdx.x;
#pragma unroll
    for (int f = 0; f < ZPF; ++f) {
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            pre[f][i] = cpx<T>{T(0), T(0)};
            if (Map::live(tid, i)) {
                int t, b;
                Map::hbm(tid, i, t, b);
                const int l = w * TB + b;
                if (zchunk_ok<T, LZ>(w, b - b % CH, nlines)) pre[f][i] = stage[f * EL + t * TB + b];
                else if (l < nlines) pre[f][i] = sp.s[zslot<GROUP>(c, f)][(size_t)l + (size_t)t * nlines];
            }
        }
    }
}

// PIN: the first PIN fields of the group come in pre (prefetched);
// NEXT >= 0: the group prefetches the first ZPF tiles of group NEXT into pre
// (only without a modulation spectrum, whose loads would sit in the k-space
// stage behind the prefetch).
template <typename T, int LZ, int GROUP, int PIN, int NEXT, typename L>
__device__ __forceinline__ void zpass_body(const Consts& c, const Spectra<T>& sp, const ZArgs<T>& a, Scalars* sc,
                                              const cpx<T>* twz, int bx, L& lds,
                                              cpx<T> (&pre)[ZPF][ZTile<T, LZ>::PPT])
{
    constexpr int TB = ZTile<T, LZ>::TB, ST = ZTile<T, LZ>::ST, EL = ZTile<T, LZ>::EL, PPT = ZTile<T, LZ>::PPT;
    constexpr int NF = ZGroupShape<GROUP>::NF;
    constexpr int F0 = ZGroupShape<GROUP>::F0;
    constexpr int FB = ZGroupShape<GROUP>::FB;
    static_assert(NF % FB == 0, "field batches must divide the fields");
    constexpr int STB = FB * TB + 1;
    static_assert(sizeof(lds.buf) >= sizeof(cpx<T>) * LZ * STB, "LDS pool too small for the group");
    using KT = typename kspace_type<T>::type;
    cpx<T>* buf = lds.buf;
    cpx<T>* tws = lds.tws;
    double* red = lds.red;
    int* lkx = lds.lkx;   // per pencil: kx, ky, ordinal among the kx = 0 pencils (or -1)
    int* lky = lds.lky;
    int* lord = lds.lord;
    extern __shared__ char zdyn[];
    cpx<T>* side = reinterpret_cast<cpx<T>*>(zdyn);   // [ordinal*LZ + t][NF]
    const int tid = threadIdx.x;
    const int w = bx;
    const int nlines = c.Mx * c.Ly;
    const size_t zstride = (size_t)nlines;
    for (int i = tid; i < LZ; i += NT) tws[i] = twz[i];
    auto slot_of = [&](int fi) { return zslot<GROUP>(c, fi); };

    using Map = ZMap<NT, TB, LZ>;
    // element e of the tile in HBM order: LDS slot and HBM offset
    auto hbm_at = [&](int e, int& lds, size_t& off) -> bool {
        int t, b;
        Map::hbm_e(e, t, b);
        const int l = w * TB + b;
        lds = t * STB + b;
        off = (size_t)l + (size_t)t * zstride;
        return l < nlines;
    };
    auto hbm_i = [&](int i, int& lds, size_t& off) -> bool {
        int t, b;
        Map::hbm(tid, i, t, b);
        const int l = w * TB + b;
        lds = t * STB + b;
        off = (size_t)l + (size_t)t * zstride;
        return l < nlines;
    };

    if constexpr (GROUP == ZG_PLAIN) {
        for (int sl = 0; sl < a.nplain; ++sl) {
            cpx<T>* s = sp.s[sl];
            for (int e = tid; e < EL; e += NT) {
                int lds; size_t off;
                const bool ok = hbm_at(e, lds, off);
                buf[lds] = ok ? s[off] : cpx<T>{T(0), T(0)};
            }
            __syncthreads();
            if (a.plain_sign < 0) fft_lds<T, LZ, TB, NT, -1, STB>(buf, tws, tid);
            else fft_lds<T, LZ, TB, NT, +1, STB>(buf, tws, tid);
            for (int e = tid; e < EL; e += NT) {
                int lds; size_t off;
                if (hbm_at(e, lds, off)) s[off] = buf[lds];
            }
            __syncthreads();
        }
        return;
    } else {
        // pencil metadata; the kx = 0 pencils of the tile are numbered in order
        const int l0 = w * TB;
        const int b0 = (c.Mx - l0 % c.Mx) % c.Mx;     // first kx = 0 pencil of the tile
        for (int b = tid; b < TB; b += NT) {
            const int l = l0 + b;
            const int ky = l / c.Mx, kx = l - ky * c.Mx;
            const bool ok = l < nlines;
            lkx[b] = ok ? kx : -1;
            lky[b] = ky;
            lord[b] = (ok && kx == 0) ? (b - b0) / c.Mx : -1;
        }
        const int nsp = (b0 < TB && b0 < nlines - l0) ? ((min(TB, nlines - l0) - 1 - b0) / c.Mx + 1) * LZ : 0;

        cpx<T> v[NF][PPT];
        // every field's tile is requested from HBM up front, so that one load
        // latency covers the whole tile; the forward transforms then run
        // field by field through LDS and the results stay in registers
#pragma unroll
        for (int fi = 0; fi < NF; ++fi) {
            const cpx<T>* __restrict__ src = sp.s[slot_of(fi)];
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                const int e = tid + i * NT;
                v[fi][i] = cpx<T>{T(0), T(0)};
                if (fi < PIN) {
                    v[fi][i] = pre[fi < PIN ? fi : 0][i];
                } else if (EL % NT == 0 || e < EL) {
                    int lds; size_t off;
                    if (hbm_i(i, lds, off)) v[fi][i] = src[off];
                }
            }
        }
        // The first and the last pass of every z transform run in registers
        // where the tile allows: a thread's PPT points of its pencil b,
        // t0 + i S (S = NT / TB), are the inputs of one radix-PPT butterfly of
        // the first Stockham pass (stride 1) and the outputs of one of the
        // last (stride LZ / PPT), so the tile goes from HBM to the k-space
        // stage and back through LDS only for the middle passes; v[.][i] then
        // holds the point (t0 + i S, b).  Otherwise the whole transform runs
        // in LDS and the k-space stage reads the tile in line order.
        constexpr int S = NT / TB;
        constexpr bool REG = ZTile<T, LZ>::REG;
        const int zt0 = tid / TB, zb = tid % TB;
        // the point (t, b) that iteration i of this thread holds after the forward transforms
        auto pt_i = [&](int i, int& t, int& b) {
            if constexpr (REG) Map::hbm(tid, i, t, b);
            else Map::line(tid, i, t, b);
        };
        // A group that prefetches the next one loads the table values of its
        // points here, with its tile: the prefetch is issued after the
        // forward transforms, once every ordinary load has landed, because
        // the compiler does not order LDS-DMA loads with ordinary ones and
        // would otherwise drain the prefetch at the first table value used.
        // Every point gets its load (kx = 0 and empty pencils at a valid
        // entry, unused) so that no load depends on a branch.
        constexpr bool PRETAB = NEXT >= 0;
        KTab<KT> tab[PRETAB ? PPT : 1];
        if constexpr (PRETAB) {
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                int t, b;
                pt_i(i, t, b);
                t = min(t, LZ - 1);
                const int l = l0 + b;
                const int ky = l / c.Mx, kx = l - ky * c.Mx;
                const bool ok = l < nlines;
                tab[i] = ktab_load<KT, GROUP, T, 0>(c, a, ok ? kx : 0, ok ? ky : 0, t, ok ? l + t * nlines : 0);
            }
        }
        if constexpr (REG) {
#pragma unroll
            for (int fb = 0; fb < NF; fb += FB) {
#pragma unroll
                for (int f = 0; f < FB; ++f) {
                    cpx<T> wv[PPT];
#pragma unroll
                    for (int q = 0; q < PPT; ++q) wv[q] = v[fb + f][q];
                    dft<PPT, -1>(wv);
#pragma unroll
                    for (int q = 0; q < PPT; ++q) buf[(PPT * zt0 + q) * STB + f * TB + zb] = wv[q];
                }
                __syncthreads();
                fft_lds_span<T, LZ, FB * TB, NT, -1, STB, PPT, LZ / PPT>(buf, tws, tid);
#pragma unroll
                for (int f = 0; f < FB; ++f) {
                    cpx<T> wv[PPT];
#pragma unroll
                    for (int q = 0; q < PPT; ++q) wv[q] = buf[(zt0 + q * S) * STB + f * TB + zb];
#pragma unroll
                    for (int q = 1; q < PPT; ++q) wv[q] = cmul(wv[q], tws[zt0 * q]);
                    dft<PPT, -1>(wv);
                    const int ord = lord[zb];
#pragma unroll
                    for (int q = 0; q < PPT; ++q) {
                        v[fb + f][q] = wv[q];
                        if (ord >= 0) side[(ord * LZ + zt0 + q * S) * NF + fb + f] = wv[q];
                    }
                }
                __syncthreads();
            }
            // kx = 0 pencils: R pencils (ky = 0, Ly/2) carry C = P + iQ; split
            // into P(t), t < Lz/2, and Q(Lz - t), pairwise in the side buffer
            constexpr int npair = (LZ - 1) / 2;
            const int npen = nsp / LZ;
            for (int e = tid; e < npen * npair * NF; e += NT) {
                const int fi = e % NF, rest = e / NF, t = 1 + rest % npair, ord = rest / npair;
                const int ky = lky[b0 + ord * c.Mx];
                if (!(ky == 0 || 2 * ky == c.Ly)) continue;
                cpx<T>* lo = side + (ord * LZ + t) * NF + fi;
                cpx<T>* hi = side + (ord * LZ + LZ - t) * NF + fi;
                const cpx<T> x = *lo, m = *hi;
                *lo = split_pq(x, m, true);
                *hi = split_pq(m, x, false);
            }
            // the side buffer is complete before the k-space stage reads it
            // (this barrier comes before a prefetch is issued: a barrier after
            // it would wait for the prefetch's LDS writes)
            __syncthreads();
        } else {
#pragma unroll
            for (int fb = 0; fb < NF; fb += FB) {
#pragma unroll
                for (int f = 0; f < FB; ++f) {
#pragma unroll
                    for (int i = 0; i < PPT; ++i) {
                        const int e = tid + i * NT;
                        if (EL % NT == 0 || e < EL) {
                            int lds; size_t off;
                            hbm_i(i, lds, off);
                            buf[lds + f * TB] = v[fb + f][i];
                        }
                    }
                }
                __syncthreads();
                fft_lds<T, LZ, FB * TB, NT, -1, STB>(buf, tws, tid);
#pragma unroll
                for (int f = 0; f < FB; ++f) {
                    const int fi = fb + f;
#pragma unroll
                    for (int i = 0; i < PPT; ++i) {
                        const int e = tid + i * NT;
                        if (EL % NT == 0 || e < EL) {
                            int t, b;
                            pt_i(i, t, b);
                            const cpx<T> x = buf[t * STB + f * TB + b];
                            const int ord = lord[b];
                            if (ord < 0) {
                                v[fi][i] = x;
                            } else {
                                // kx = 0 pencil: R pencils (ky = 0, Ly/2) are split into P and Q here
                                const int ky = lky[b];
                                const bool rp = (ky == 0 || 2 * ky == c.Ly) && t != 0 && 2 * t != LZ;
                                side[(ord * LZ + t) * NF + fi] =
                                    rp ? split_pq(x, buf[(LZ - t) * STB + f * TB + b], 2 * t < LZ) : x;
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }

        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
        if constexpr (PRETAB) {
            wait_vmem_all();
            zprefetch_issue<T, LZ, NEXT>(c, sp, w, buf);
        }
        // points off the kx = 0 pencils: one evaluation each, in place
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            const int e = tid + i * NT;
            if (!(EL % NT == 0 || e < EL)) continue;
            int t, b;
            pt_i(i, t, b);
            const int kx = lkx[b];
            if (kx <= 0) continue;
            const int ky = lky[b];
            const int kk = l0 + b + t * nlines;
            cpx<KT> X[NF], O[NF];
#pragma unroll
            for (int q = 0; q < NF; ++q) X[q] = {KT(v[q][i].x), KT(v[q][i].y)};
            const KTab<KT> tb = PRETAB ? tab[PRETAB ? i : 0] : ktab_load<KT, GROUP>(c, a, kx, ky, t, kk);
            keval<KT, GROUP>(c, a, sc, kx, ky, t, kk, -1, tb, X, O, 1.0, acc0, acc1, acc2);
#pragma unroll
            for (int q = 0; q < NF; ++q) v[q][i] = cpx<T>{T(O[q].x), T(O[q].y)};
        }
        // points of the kx = 0 pencils, from the side buffer: the first
        // evaluation of every point, then the second where the plan has one,
        // its first share kept in the second half of the side buffer
        cpx<T>* first = side + nsp * NF;
#pragma unroll 1
        for (int j = tid; j < nsp; j += NT) {
            const int ord = j / LZ, t = j - ord * LZ;
            const KPlan pl = kplan<GROUP == ZG_DD || GROUP == ZG_DDZ>(c, lky[b0 + ord * c.Mx], t);
            cpx<KT> X[NF], O[NF];
#pragma unroll
            for (int q = 0; q < NF; ++q) X[q] = ev_in(cpx<KT>{KT(side[j * NF + q].x), KT(side[j * NF + q].y)}, pl.m0);
            const int kkp = pl.yp >= 0 ? kindex32(c, 0, pl.yp, pl.zp) : -1;
            const int kk0 = kindex32(c, pl.x0, pl.y0, pl.z0);
            keval<KT, GROUP>(c, a, sc, pl.x0, pl.y0, pl.z0, kk0, kkp, ktab_load<KT, GROUP>(c, a, pl.x0, pl.y0, pl.z0, kk0),
                             X, O, pl.wm, acc0, acc1, acc2);
            const KT half = pl.nev == 2 ? KT(0.5) : KT(1);
            cpx<T>* dst = pl.nev == 2 ? first + j * NF : side + j * NF;
#pragma unroll
            for (int q = 0; q < NF; ++q) {
                const cpx<KT> r = ev_out(cpx<KT>{KT(0), KT(0)}, O[q], pl.m0, half);
                dst[q] = cpx<T>{T(r.x), T(r.y)};
            }
        }
#pragma unroll 1
        for (int j = tid; j < nsp; j += NT) {
            const int ord = j / LZ, t = j - ord * LZ;
            const KPlan pl = kplan<GROUP == ZG_DD || GROUP == ZG_DDZ>(c, lky[b0 + ord * c.Mx], t);
            if (pl.nev < 2) continue;
            cpx<KT> X[NF], O[NF];
#pragma unroll
            for (int q = 0; q < NF; ++q) X[q] = ev_in(cpx<KT>{KT(side[j * NF + q].x), KT(side[j * NF + q].y)}, pl.m1);
            const int kk1 = kindex32(c, pl.x1, pl.y1, pl.z1);
            keval<KT, GROUP>(c, a, sc, pl.x1, pl.y1, pl.z1, kk1, -1, ktab_load<KT, GROUP>(c, a, pl.x1, pl.y1, pl.z1, kk1),
                             X, O, pl.wm, acc0, acc1, acc2);
#pragma unroll
            for (int q = 0; q < NF; ++q) {
                const cpx<KT> r = ev_out(cpx<KT>{KT(first[j * NF + q].x), KT(first[j * NF + q].y)}, O[q], pl.m1,
                                         KT(0.5));
                side[j * NF + q] = cpx<T>{T(r.x), T(r.y)};
            }
        }
        __syncthreads();
        if constexpr (NEXT >= 0) {
            // every wave's prefetch has landed before any wave reads it, and
            // every read is done before the inverse transforms reuse the pool
            wait_vmem_all();
            __syncthreads();
            zprefetch_collect<T, LZ, NEXT>(c, sp, w, buf, pre);
            __syncthreads();
        }
        // inverse transforms; kx = 0 pencils take their values from the side
        // buffer, R pencils joining P(t) (t < Lz/2) and Q(t) (held at Lz - t) back
        auto inv_in = [&](int fi, int i, int t, int b) -> cpx<T> {
            const int ord = lord[b];
            if (ord < 0) return v[fi][i];
            const int ky = lky[b];
            cpx<T> x = side[(ord * LZ + t) * NF + fi];
            if ((ky == 0 || 2 * ky == c.Ly) && t != 0 && 2 * t != LZ)
                x = join_pq(x, side[(ord * LZ + LZ - t) * NF + fi], 2 * t < LZ);
            return x;
        };
        if constexpr (REG) {
#pragma unroll
            for (int fb = 0; fb < NF; fb += FB) {
#pragma unroll
                for (int f = 0; f < FB; ++f) {
                    cpx<T> wv[PPT];
#pragma unroll
                    for (int q = 0; q < PPT; ++q) wv[q] = inv_in(fb + f, q, zt0 + q * S, zb);
                    dft<PPT, +1>(wv);
#pragma unroll
                    for (int q = 0; q < PPT; ++q) buf[(PPT * zt0 + q) * STB + f * TB + zb] = wv[q];
                }
                __syncthreads();
                fft_lds_span<T, LZ, FB * TB, NT, +1, STB, PPT, LZ / PPT>(buf, tws, tid);
#pragma unroll
                for (int f = 0; f < FB; ++f) {
                    cpx<T>* __restrict__ dst = sp.s[slot_of(fb + f)];
                    cpx<T> wv[PPT];
#pragma unroll
                    for (int q = 0; q < PPT; ++q) wv[q] = buf[(zt0 + q * S) * STB + f * TB + zb];
#pragma unroll
                    for (int q = 1; q < PPT; ++q) {
                        cpx<T> tw = tws[zt0 * q];
                        tw.y = -tw.y;
                        wv[q] = cmul(wv[q], tw);
                    }
                    dft<PPT, +1>(wv);
#pragma unroll
                    for (int q = 0; q < PPT; ++q) {
                        int lds; size_t off;
                        if (hbm_i(q, lds, off)) dst[off] = wv[q];
                    }
                }
                __syncthreads();
            }
        } else {
#pragma unroll
            for (int fb = 0; fb < NF; fb += FB) {
#pragma unroll
                for (int f = 0; f < FB; ++f) {
                    const int fi = fb + f;
#pragma unroll
                    for (int i = 0; i < PPT; ++i) {
                        const int e = tid + i * NT;
                        if (EL % NT == 0 || e < EL) {
                            int t, b;
                            pt_i(i, t, b);
                            buf[t * STB + f * TB + b] = inv_in(fi, i, t, b);
                        }
                    }
                }
                __syncthreads();
                fft_lds<T, LZ, FB * TB, NT, +1, STB>(buf, tws, tid);
#pragma unroll
                for (int f = 0; f < FB; ++f) {
                    cpx<T>* __restrict__ dst = sp.s[slot_of(fb + f)];
#pragma unroll
                    for (int i = 0; i < PPT; ++i) {
                        const int e = tid + i * NT;
                        if (EL % NT == 0 || e < EL) {
                            int lds; size_t off;
                            if (hbm_i(i, lds, off)) dst[off] = buf[lds + f * TB];
                        }
                    }
                }
                __syncthreads();
            }
        }
        const int nb = a.maxblocks;
        if constexpr (GROUP == ZG_DD || GROUP == ZG_DDZ) {
            double r[2] = {acc0, acc1};
            block_sums<2>(r, red);
            if (tid == 0) {
                a.partials[P_ESR * nb + bx] = r[0];
                a.partials[P_ELR * nb + bx] = r[1];
            }
        } else {
            double r[3] = {acc0, acc1, acc2};
            block_sums<3>(r, red);
            if (tid == 0) {
                a.partials[P_EIS * nb + bx] = r[0];
                a.partials[P_EIC * nb + bx] = r[1];
                a.partials[P_EIM * nb + bx] = r[2];
            }
        }
    }
}

template <typename T, int LZ, int GROUP>
constexpr int zbuf_size() { return LZ * (ZGroupShape<GROUP>::FB * ZTile<T, LZ>::TB + 1); }

template <typename T, int LZ, int GROUP>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(GROUP == ZG_ACOU ? 1 : 2))) void k_zpass(Consts c, Spectra<T> sp, ZArgs<T> a, Scalars* sc,
                                              const cpx<T>* __restrict__ twz)
{
    __shared__ ZLds<T, LZ, zbuf_size<T, LZ, GROUP>()> lds;
    cpx<T> pre[ZPF][ZTile<T, LZ>::PPT];
    zpass_body<T, LZ, GROUP, 0, -1>(c, sp, a, sc, twz, blockIdx.x, lds, pre);
}

// Both z groups of a step on one tile in one workgroup (the dipole-dipole and
// the strain or acoustic group read and write disjoint spectra); even
// workgroups run the dipole-dipole group first and odd ones second, so that
// half the grid is loading or storing while the other half is in its
// transforms.
template <typename T, int LZ, int G2>
constexpr int zpair_buf_size()
{
    return zbuf_size<T, LZ, ZG_DD>() > zbuf_size<T, LZ, G2>() ? zbuf_size<T, LZ, ZG_DD>() : zbuf_size<T, LZ, G2>();
}

template <typename T, int LZ, int G2, int DDG = ZG_DD>
__device__ __forceinline__ void zpair_body(const Consts& c, const Spectra<T>& sp, const ZArgs<T>& a, Scalars* sc,
                                           const cpx<T>* twz, int bx)
{
    __shared__ ZLds<T, LZ, zpair_buf_size<T, LZ, G2>()> lds;
    static_assert(sizeof(lds.buf) >= ZPF * ZTile<T, LZ>::EL * sizeof(cpx<T>), "prefetch staging exceeds the LDS pool");
    // the first group prefetches the second group's first field batch (one
    // Z*, no modulation spectrum; only edges of 64 and more with register
    // first/last passes, where it was measured: below, a step is launch-bound)
    constexpr bool PF = ZPrefetch<T, LZ>::OK && DDG == ZG_DD && ZTile<T, LZ>::REG && LZ >= 64;
    cpx<T> pre[ZPF][ZTile<T, LZ>::PPT];
    if (a.split) {
        // a grid of fewer tiles than the device has CUs: the two groups of a
        // tile run side by side in two workgroups instead of one after the
        // other (the partial sums are indexed by tile either way)
        if (bx & 1) zpass_body<T, LZ, G2, 0, -1>(c, sp, a, sc, twz, bx >> 1, lds, pre);
        else zpass_body<T, LZ, DDG, 0, -1>(c, sp, a, sc, twz, bx >> 1, lds, pre);
        return;
    }
    if (PF && !c.modulation) {
        if (bx & 1) {
            zpass_body<T, LZ, G2, 0, PF ? DDG : -1>(c, sp, a, sc, twz, bx, lds, pre);
            __syncthreads();
            zpass_body<T, LZ, DDG, PF ? ZPF : 0, -1>(c, sp, a, sc, twz, bx, lds, pre);
        } else {
            zpass_body<T, LZ, DDG, 0, PF ? G2 : -1>(c, sp, a, sc, twz, bx, lds, pre);
            __syncthreads();
            zpass_body<T, LZ, G2, PF ? ZPF : 0, -1>(c, sp, a, sc, twz, bx, lds, pre);
        }
    } else {
        if (bx & 1) {
            zpass_body<T, LZ, G2, 0, -1>(c, sp, a, sc, twz, bx, lds, pre);
            __syncthreads();
            zpass_body<T, LZ, DDG, 0, -1>(c, sp, a, sc, twz, bx, lds, pre);
        } else {
            zpass_body<T, LZ, DDG, 0, -1>(c, sp, a, sc, twz, bx, lds, pre);
            __syncthreads();
            zpass_body<T, LZ, G2, 0, -1>(c, sp, a, sc, twz, bx, lds, pre);
        }
    }
}

// DDG: ZG_DD (one Z*, folded into L(k)) or ZG_DDZ (species with different Z*)
template <typename T, int LZ, int G2, int DDG = ZG_DD>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(G2 == ZG_ACOU ? 1 : 2)))
void k_zpair(Consts c, Spectra<T> sp, ZArgs<T> a, Scalars* sc, const cpx<T>* __restrict__ twz)
{
    zpair_body<T, LZ, G2, DDG>(c, sp, a, sc, twz, blockIdx.x);
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

// c2r along x of NF spectral slots into real device fields (XI_GENERIC,
// set-up and output work), unnormalised: two fields per complex transform.
template <typename T, int LX, int NF>
__device__ __forceinline__ void xinv_body(const Consts& c, const FieldsT<T>& f, const Spectra<T>& sp, const XinvArgs& a,
                                             const Scalars* __restrict__ sc,
                                             const cpx<T>* twx, int bx)
{
    constexpr int RB = xinv_rows_of(LX);
    constexpr int ST = RB + FHIP_XPAD;
    constexpr int CELLS = RB * LX;
    constexpr int MX = LX / 2;
    constexpr int HP = RB * MX;
    __shared__ cpx<T> buf[LX * ST];
    __shared__ cpx<T> tws[LX];
    const int tid = threadIdx.x;
    const int nrows = c.Ly * c.Lz;
    const int row0 = bx * RB;
    for (int i = tid; i < LX; i += NT) tws[i] = twx[i];
    for (int f0 = 0; f0 < NF; f0 += 2) {
        const bool two = (f0 + 1 < NF);
        // Z = A + i B with both half spectra completed Hermitian; kx = 0
        // holds the real A[0] + i A[Lx/2]
        for (int e = tid; e < HP; e += NT) {
            const int r = e / MX, kx = e - r * MX;
            cpx<T> A{T(0), T(0)}, B{T(0), T(0)};
            if (row0 + r < nrows) {
                A = sp.s[f0][(size_t)row0 * MX + e];
                if (two) B = sp.s[f0 + 1][(size_t)row0 * MX + e];
            }
            if (kx == 0) {
                buf[r] = cpx<T>{A.x, B.x};
                buf[MX * ST + r] = cpx<T>{A.y, B.y};
            } else {
                buf[kx * ST + r] = cpx<T>{A.x - B.y, A.y + B.x};
                buf[(LX - kx) * ST + r] = cpx<T>{A.x + B.y, B.x - A.y};
            }
        }
        __syncthreads();
        fft_lds<T, LX, RB, NT, +1, ST>(buf, tws, tid);
        for (int e = tid; e < CELLS; e += NT) {
            const int r = e / LX, x = e - r * LX;
            if (row0 + r >= nrows) continue;
            const size_t cc = (size_t)(row0 + r) * LX + x;
            const cpx<T> z = buf[x * ST + r];
            a.gen[f0][cc] = (double)z.x;
            if (two) a.gen[f0 + 1][cc] = (double)z.y;
        }
        __syncthreads();
    }
}

// xinv of a step (NF = 9 inverse fields, plus 3 with acoustic MD and 3 with
// species of different Z*): the field
// pairs are transformed PB at a time side by side in LDS (PB * RB lines, so a
// 128-point transform takes radix-8 passes with every thread busy) and each
// transformed value is folded into the cell's force as it comes out of LDS:
// only the force, the cell's state and one batch of spectra are live in
// registers, and the next batch's spectra load while this one transforms
// (FP64 128^3: 90 -> 82 us against holding every field's values).  The force
// is the sum of the same terms as dVddi-dipole-dipole.F,
// dVddi_inho_coupling.F, dVddi-unharmonic.F, mirror_force_mirror.F:39-65 and
// dVddi-homo-coupling.F, the inhomogeneous coupling added field by field.
template <typename T, int LX, int NF>
__device__ __forceinline__ void xinv_step_body(const Consts& c, const FieldsT<T>& f, const Spectra<T>& sp,
                                               const XinvArgs& a, const Scalars* __restrict__ sc,
                                               const cpx<T>* twx, int bx)
{
    static_assert(NF == 9 || NF == 12 || NF == 15, "step fields");
    constexpr int RB = xinv_rows_of(LX);            // rows per workgroup
    constexpr int PB = 2;                           // field pairs per LDS transform
    constexpr int NL = PB * RB;                     // lines of a full batch
    constexpr int ST = NL + FHIP_XPAD;
    constexpr int CELLS = RB * LX;
    constexpr int CPT = (CELLS + NT - 1) / NT;
    constexpr int MX = LX / 2;
    constexpr int HP = RB * MX;
    constexpr int HPT = (HP + NT - 1) / NT;
    constexpr int NPAIR = (NF + 1) / 2;
    constexpr int NBATCH = (NPAIR + PB - 1) / PB;
    __shared__ cpx<T> buf[LX * ST];
    __shared__ cpx<T> tws[LX];
    __shared__ double red[NT];
    const int tid = threadIdx.x;
    const int nrows = c.Ly * c.Lz;
    const int row0 = bx * RB;
    for (int i = tid; i < LX; i += NT) tws[i] = twx[i];
    const bool kick = a.mode == XI_NP || a.mode == XI_LF || a.mode == XI_VS;

    // the cells' state: u and the codes now (the folds need u), the momenta
    // the kick reads before the last transform
    // u, the momenta and the force accumulators are held in the state type:
    // in the FP32 mode the local forces accumulate in FP32 (every term is
    // formed in FP64 and rounded once as it is added), which keeps the pass
    // at 126 VGPRs, 4 waves per SIMD (FP32 128^3: 58.4 -> 56.4 us)
    bool live[CPT];
    int cell[CPT];                                  // N < 2^31 for every supported supercell
    T u[CPT][3], F[CPT][3], pp[CPT][3];
    unsigned char code[CPT];
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
        const int e = tid + i * NT;
        const int r = e / LX, x = e - r * LX;
        live[i] = (CELLS % NT == 0 || e < CELLS) && row0 + r < nrows;
        cell[i] = (row0 + r) * LX + x;
        code[i] = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) u[i][q] = F[i][q] = pp[i][q] = T(0);
        if (live[i]) {
            code[i] = f.code[cell[i]];
#pragma unroll
            for (int q = 0; q < 3; ++q) u[i][q] = f.u[q][cell[i]];
        }
    }
    double g[NF >= 12 ? CPT : 1][3];                // acoustic dV/dw (dVdac.F:56-62)

    cpx<T> h[PB][2][HPT];
    auto load_batch = [&](int b) {
#pragma unroll
        for (int p = 0; p < PB; ++p) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int fi = 2 * (b * PB + p) + k;
#pragma unroll
                for (int i = 0; i < HPT; ++i) {
                    const int e = tid + i * NT;
                    h[p][k][i] = cpx<T>{T(0), T(0)};
                    if (fi >= NF || (HP % NT != 0 && e >= HP)) continue;
                    if (row0 + e / MX < nrows) h[p][k][i] = once_load(sp.s[fi] + (size_t)row0 * MX + e);
                }
            }
        }
    };
    // value v of inverse field fi at cell i, normalised by 1/N
    auto fold = [&](int i, int fi, double v) {
        v = v * c.Ninv;
        if (fi < 3) {
            F[i][fi] = F[i][fi] + v;                                        // dVddi-dipole-dipole.F:34-38
        } else if (fi >= 9 && c.zsp && fi >= c.zu_slot) {
            // long-range force of species with different Z*: Z*(R) (L (Z* u)^)(R) / N
            // (dVddi-dipole-dipole.F:52-62)
            F[i][fi - c.zu_slot] = F[i][fi - c.zu_slot] + c.Zs[code[i] & 3] * v;
        } else if (fi < 9) {
            // eta_jj(R) / N through homogeneous_coupling_2 (dVddi_inho_coupling.F:54-82)
            const int jj = fi - 3;
            double e[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) e[m] = v * c.hc2[jj + 6 * m];
            F[i][0] = F[i][0] + (2 * e[0] * u[i][0] + e[4] * u[i][2] + e[5] * u[i][1]);
            F[i][1] = F[i][1] + (2 * e[1] * u[i][1] + e[3] * u[i][2] + e[5] * u[i][0]);
            F[i][2] = F[i][2] + (2 * e[2] * u[i][2] + e[3] * u[i][1] + e[4] * u[i][0]);
        } else if constexpr (NF >= 12) {
            g[i][fi - 9] = v;
        }
    };

    load_batch(0);
#pragma unroll
    for (int b = 0; b < NBATCH; ++b) {
        const int npb = (b * PB + PB <= NPAIR) ? PB : NPAIR - b * PB;   // pairs in this batch
        // Z = A + i B of each pair with both half spectra completed Hermitian;
        // kx = 0 holds the real A[0] + i A[Lx/2]
#pragma unroll
        for (int p = 0; p < PB; ++p) {
            if (p >= npb) break;
            const bool two = 2 * (b * PB + p) + 1 < NF;
#pragma unroll
            for (int i = 0; i < HPT; ++i) {
                const int e = tid + i * NT;
                if (HP % NT != 0 && e >= HP) continue;
                const int r = e / MX, kx = e - r * MX, line = p * RB + r;
                const cpx<T> A = h[p][0][i], B = two ? h[p][1][i] : cpx<T>{T(0), T(0)};
                if (kx == 0) {
                    buf[line] = cpx<T>{A.x, B.x};
                    buf[MX * ST + line] = cpx<T>{A.y, B.y};
                } else {
                    buf[kx * ST + line] = cpx<T>{A.x - B.y, A.y + B.x};
                    buf[(LX - kx) * ST + line] = cpx<T>{A.x + B.y, B.x - A.y};
                }
            }
        }
        // the next batch's spectra (or, before the last transform, the
        // momenta) are requested while this batch transforms
        if (b + 1 < NBATCH) {
            load_batch(b + 1);
        } else if (kick) {
#pragma unroll
            for (int i = 0; i < CPT; ++i) {
                if (!live[i]) continue;
#pragma unroll
                for (int q = 0; q < 3; ++q) pp[i][q] = (a.mode == XI_NP ? f.pk[q] : f.p[q])[cell[i]];
            }
        }
        __syncthreads();
        if (npb == PB) fft_lds<T, LX, NL, NT, +1, ST>(buf, tws, tid);
        else fft_lds<T, LX, RB, NT, +1, ST>(buf, tws, tid);
#pragma unroll
        for (int i = 0; i < CPT; ++i) {
            const int e = tid + i * NT;
            if (CELLS % NT != 0 && e >= CELLS) continue;
            const int r = e / LX, x = e - r * LX;
#pragma unroll
            for (int p = 0; p < PB; ++p) {
                if (p >= npb) break;
                const int f0 = 2 * (b * PB + p);
                const cpx<T> z = buf[x * ST + p * RB + r];
                fold(i, f0, (double)z.x);
                if (f0 + 1 < NF) fold(i, f0 + 1, (double)z.y);
            }
        }
        __syncthreads();
    }

    double eunh = 0.0, ede = 0.0, ke1 = 0.0, ke2 = 0.0, kea = 0.0;
    const bool film = c.mir_lo < c.Lz;
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
        if (!live[i]) continue;
        const int e = tid + i * NT;
        const int r = e / LX, x = e - r * LX;
        const size_t cc = cell[i];
        const int s = code[i] & 3;
        const bool counted = (code[i] & 8) != 0;
        // film/epit: the dipole force and kick of a mirrored cell come from
        // its source cell (mirror_force_mirror.F:80-84, k_mirror_kick)
        const bool mirrored = film && row0 + r >= c.mir_lo * c.Ly;
        const double minv = (code[i] & 4) ? 0.0 : c.minv[s];
        T* Fi = F[i];
        const double ui[3] = {(double)u[i][0], (double)u[i][1], (double)u[i][2]};
        // on-site anharmonic terms (dVddi-unharmonic.F:34-65)
        const double d0 = ui[0] * ui[0], d1 = ui[1] * ui[1], d2 = ui[2] * ui[2];
        const double d[3] = {d0, d1, d2};
        const double ss = d0 + d1 + d2;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int q1 = (q + 1) % 3, q2 = (q + 2) % 3;
            Fi[q] = Fi[q] + ui[q] * 2 * (c.kappa2[s] - c.kappa2[0]);
            Fi[q] = Fi[q] + ui[q] * (c.fa2g[s] * ss - c.tg[s] * d[q]);
            Fi[q] = Fi[q] + ui[q] * (c.sk[s] * (ss * ss) + c.k4x8[s] * (ss * ss * ss)
                                     - c.tk2[s] * (d[q] * d[q]) - c.fk[s] * d[q1] * d[q2]);
        }
        eunh += (c.kappa2[s] - c.kappa2[0]) * ss + c.alpha[s] * (ss * ss)
              + c.gamma[s] * (d0 * d1 + d1 * d2 + d2 * d0) + c.k1[s] * (ss * ss * ss)
              + c.k2[s] * ((d0 * d0) * (d1 + d2) + (d1 * d1) * (d2 + d0) + (d2 * d2) * (d0 + d1))
              + c.k3[s] * d0 * d1 * d2 + c.k4[s] * ((ss * ss) * (ss * ss));
        // external and local field (mirror_force_mirror.F:39-65)
        if (c.efield) {
            double E[3];
            if (mirrored && sc->e_mir != 0.0) {                     // initialize_E_field.F:68-71
                const int rr = row0 + r, iz = rr / c.Ly, iy = rr - iz * c.Ly;
                const size_t cs = ((size_t)(c.mir_src - iz) * c.Ly + iy) * LX + x;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double Es = c.has_eloc ? f.eloc[q][cs] + sc->E_add[q] : sc->E_add[q];
                    E[q] = q < 2 ? -Es : Es;
                }
            } else {
#pragma unroll
                for (int q = 0; q < 3; ++q) E[q] = c.has_eloc ? f.eloc[q][cc] + sc->E_add[q] : sc->E_add[q];
            }
            const double zs = c.Zs[s];
#pragma unroll
            for (int q = 0; q < 3; ++q) Fi[q] = Fi[q] - zs * E[q];
            ede += zs * (ui[0] * E[0] + ui[1] * E[1] + ui[2] * E[2]);
        }
        // homogeneous coupling (dVddi-homo-coupling.F:28-44)
        {
            const double* eh = sc->e_homo;
            Fi[0] = Fi[0] + 2 * eh[0] * ui[0] + eh[4] * ui[2] + eh[5] * ui[1];
            Fi[1] = Fi[1] + 2 * eh[1] * ui[1] + eh[3] * ui[2] + eh[5] * ui[0];
            Fi[2] = Fi[2] + 2 * eh[2] * ui[2] + eh[3] * ui[1] + eh[4] * ui[0];
        }
        // kicks
        if (mirrored) {
        } else if (a.mode == XI_NP) {
            const double dt2s = c.dt_2 * sc->s;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double p1 = pp[i][q] - dt2s * Fi[q];                 // nose-poincare.F:66-71
                const double p2 = p1 - dt2s * Fi[q];                       // nose-poincare.F:24-29 (next step)
                if (counted) { ke1 += p1 * p1 * minv; ke2 += p2 * p2 * minv; }
                f.pk[q][cc] = p2;
                if (a.materialize) f.p[q][cc] = p1;
            }
        } else if (a.mode == XI_LF) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double pq = pp[i][q] - c.dt * Fi[q];                 // leapfrog.F:134-139
                if (counted) ke1 += pq * pq * minv;
                f.p[q][cc] = pq;
            }
        } else if (a.mode == XI_VS) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double pq = counted ? pp[i][q] - c.dt * Fi[q] : 0.0;      // velocity_scaling.F:62-76
                if (counted) ke1 += pq * pq * minv;
                f.p[q][cc] = pq;
            }
        }
        if ((a.materialize || a.mode == XI_FORCE || film) && !mirrored) {
#pragma unroll
            for (int q = 0; q < 3; ++q) f.F[q][cc] = Fi[q];
        }
        if constexpr (NF >= 12) {
            if (!c.acoustic) continue;
            const double mwinv = c.mwinv[s];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if (a.mode == XI_LF || a.mode == XI_VS) {
                    const double pq = f.pw[q][cc] - c.dt * g[i][q];
                    kea += pq * pq * mwinv;
                    f.pw[q][cc] = pq;
                }
                if (a.materialize || a.mode == XI_FORCE) f.dVdw[q][cc] = g[i][q];
            }
        }
    }
    const int nb = a.maxblocks;
    double r[5] = {eunh, ede, ke1, ke2, kea};
    block_sums<5>(r, red);
    if (tid == 0) {
        a.partials[P_EUNH * nb + bx] = r[0];
        a.partials[P_EDE * nb + bx] = r[1];
        a.partials[P_KE1 * nb + bx] = r[2];
        a.partials[P_KE2 * nb + bx] = r[3];
        a.partials[P_KEA * nb + bx] = r[4];
    }
}

template <typename T, int LX, int NF>
__global__ __launch_bounds__(NT) void k_xinv(Consts c, FieldsT<T> f, Spectra<T> sp, XinvArgs a,
                                             const Scalars* __restrict__ sc,
                                             const cpx<T>* __restrict__ twx)
{
    if constexpr (NF >= 9) xinv_step_body<T, LX, NF>(c, f, sp, a, sc, twx, blockIdx.x);
    else xinv_body<T, LX, NF>(c, f, sp, a, sc, twx, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Film/epit: dV/du of the mirrored layers is the mirror of their source
// layers' (mirror_force_mirror.F:80-84, mirror.F:9-36), and their momenta are
// kicked with it as xinv kicks the other cells.  Mirrored layers are never
// counted (read_defects_u.F:52-54), so they add nothing to the kinetic sums.
// ---------------------------------------------------------------------------
template <typename S>
__device__ __forceinline__ void mirror_kick_body(const Consts& c, const FieldsT<S>& f, const Scalars* sc, int mode,
                                                    int materialize, int bx)
{
    const size_t plane = (size_t)c.Lx * c.Ly;
    const size_t i = (size_t)bx * NT + threadIdx.x;
    if (i >= (size_t)(c.Lz - c.mir_lo) * plane) return;
    const int iz = c.mir_lo + (int)(i / plane);
    const size_t xy = i - (size_t)(iz - c.mir_lo) * plane;
    const size_t cc = (size_t)iz * plane + xy;
    const size_t cs = (size_t)(c.mir_src - iz) * plane + xy;
    double F[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        F[q] = q < 2 ? -f.F[q][cs] : f.F[q][cs];
        f.F[q][cc] = F[q];
    }
    const bool counted = (f.code[cc] & 8) != 0;
    if (mode == XI_NP) {
        const double dt2s = c.dt_2 * sc->s;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double p1 = f.pk[q][cc] - dt2s * F[q];                    // nose-poincare.F:66-71
            const double p2 = p1 - dt2s * F[q];
            f.pk[q][cc] = p2;
            if (materialize) f.p[q][cc] = p1;
        }
    } else if (mode == XI_LF) {
#pragma unroll
        for (int q = 0; q < 3; ++q) f.p[q][cc] = f.p[q][cc] - c.dt * F[q];     // leapfrog.F:134-139
    } else if (mode == XI_VS) {
#pragma unroll
        for (int q = 0; q < 3; ++q) f.p[q][cc] = counted ? f.p[q][cc] - c.dt * F[q] : 0.0;   // velocity_scaling.F:62-76
    }
}

template <typename S>
__global__ __launch_bounds__(NT) void k_mirror_kick(Consts c, FieldsT<S> f, const Scalars* __restrict__ sc,
                                                           int mode, int materialize)
{
    mirror_kick_body(c, f, sc, mode, materialize, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Kinetic energy of p (optionally after the first Nose-Poincare half kick).
// ---------------------------------------------------------------------------
template <int KICK, typename S>
__device__ __forceinline__ void kick_body(const Consts& c, const FieldsT<S>& f, const Scalars* sc,
                                             double* partials, int maxblocks, int bx)
{
    __shared__ double red[NT];
    const size_t cc = (size_t)bx * NT + threadIdx.x;
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
            for (int q = 0; q < 3; ++q) kea += (double)f.pw[q][cc] * f.pw[q][cc] * c.mwinv[s];
        }
    }
    double r[2] = {ke, kea};
    block_sums<2>(r, red);
    if (threadIdx.x == 0) {
        partials[P_KEPK * maxblocks + bx] = r[0];
        partials[P_KEA * maxblocks + bx] = r[1];
    }
}

template <int KICK, typename S>
__global__ __launch_bounds__(NT) void k_kick(Consts c, FieldsT<S> f, const Scalars* __restrict__ sc,
                                             double* partials, int maxblocks)
{
    kick_body<KICK, S>(c, f, sc, partials, maxblocks, blockIdx.x);
}

// Deterministic fold of the nb per-block partials of NS slots by one
// workgroup: all loads are issued first, then one tree per slot.
template <int NS>
__device__ __forceinline__ void fold_slots(const double* part, int maxblocks, int nb, const int (&slots)[NS],
                                           double (&out)[NS], double* red)
{
    constexpr int U = 8;                 // loads in flight per slot and thread
    double s[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) s[q] = 0.0;
    for (int i0 = threadIdx.x; i0 < nb; i0 += U * NT) {
        double v[U][NS];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * NT;
#pragma unroll
            for (int q = 0; q < NS; ++q) v[u][q] = i < nb ? part[slots[q] * maxblocks + i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int q = 0; q < NS; ++q) s[q] += v[u][q];
        }
    }
    block_sums<NS>(s, red);
#pragma unroll
    for (int q = 0; q < NS; ++q) out[q] = s[q];
}

// The scalar stage works on an LDS copy of the device scalars: one parallel
// load, the serial arithmetic of one thread at LDS latency, one parallel
// store.
__device__ __forceinline__ void scalars_in(Scalars* ls, const Scalars* sc)
{
    constexpr int W = (int)(sizeof(Scalars) / sizeof(double));
    static_assert(sizeof(Scalars) % sizeof(double) == 0, "Scalars holds doubles only");
    for (int i = threadIdx.x; i < W; i += NT) reinterpret_cast<double*>(ls)[i] = reinterpret_cast<const double*>(sc)[i];
}
__device__ __forceinline__ void scalars_out(Scalars* sc, const Scalars* ls)
{
    constexpr int W = (int)(sizeof(Scalars) / sizeof(double));
    __syncthreads();
    for (int i = threadIdx.x; i < W; i += NT) reinterpret_cast<double*>(sc)[i] = reinterpret_cast<const double*>(ls)[i];
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
__device__ __forceinline__ void homo_scalar(const ScalarConsts& k, Scalars* sc, const double (&r)[3]);

__device__ __forceinline__ void homo_body(const ScalarConsts& k, Scalars* gsc, const double* partials,
                                             int maxblocks, int zblocks, int bx)
{
#pragma clang fp contract(off)
    __shared__ double red[NT];
    __shared__ Scalars lsc;
    scalars_in(&lsc, gsc);
    const int slots[3] = {P_EIS, P_EIC, P_EIM};
    double r[3];
    fold_slots<3>(partials, maxblocks, zblocks, slots, r, red);
    if (threadIdx.x == 0) homo_scalar(k, &lsc, r);
    scalars_out(gsc, &lsc);
}

static __global__ __launch_bounds__(NT) void k_homo(ScalarConsts k, Scalars* gsc, const double* partials,
                                             int maxblocks, int zblocks)
{
    homo_body(k, gsc, partials, maxblocks, zblocks, blockIdx.x);
}

__device__ __forceinline__ void homo_scalar(const ScalarConsts& k, Scalars* sc, const double (&r)[3])
{
#pragma clang fp contract(off)
    double* en = sc->en;
    en[E_IS] = r[0] * k.Ninv;
    en[E_IC] = r[1] * 2 * k.Ninv;
    en[E_IM] = k.modulation_energy ? r[2] * 2 * k.Ninv : 0.0;

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
__device__ __forceinline__ void finish_scalar(const ScalarConsts& k, Scalars* sc, const double (&r)[5],
                                              const double (&rz)[2], int mode, int acoustic, int next_step);

__device__ __forceinline__ void finish_body(const ScalarConsts& k, Scalars* gsc, const double* partials,
                                               int maxblocks, int xblocks, int zblocks, int mode, int acoustic,
                                               int next_step, int bx)
{
#pragma clang fp contract(off)
    __shared__ double red[NT];
    __shared__ Scalars lsc;
    scalars_in(&lsc, gsc);
    const int slots[5] = {P_EUNH, P_EDE, P_KE1, P_KE2, P_KEA};
    double r[5];
    fold_slots<5>(partials, maxblocks, xblocks, slots, r, red);
    const int zslots[2] = {P_ESR, P_ELR};
    double rz[2];
    fold_slots<2>(partials, maxblocks, zblocks, zslots, rz, red);
    if (threadIdx.x == 0) finish_scalar(k, &lsc, r, rz, mode, acoustic, next_step);
    scalars_out(gsc, &lsc);
}

static __global__ __launch_bounds__(NT) void k_finish(ScalarConsts k, Scalars* gsc, const double* partials,
                                               int maxblocks, int xblocks, int zblocks, int mode, int acoustic,
                                               int next_step)
{
    finish_body(k, gsc, partials, maxblocks, xblocks, zblocks, mode, acoustic, next_step, blockIdx.x);
}

__device__ __forceinline__ void finish_scalar(const ScalarConsts& k, Scalars* sc, const double (&r)[5],
                                              const double (&rz)[2], int mode, int acoustic, int next_step)
{
#pragma clang fp contract(off)
    const double eunh = r[0], ede = r[1], ke1 = r[2], ke2 = r[3], kea = r[4];
    double* en = sc->en;
    en[E_SR] = rz[0] * k.Ninv / 2;
    en[E_LR] = rz[1] * k.Ninv / 2;
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
    for (int i = 0; i < 13; ++i) sc->en_rep[i] = en[i];
    sc->s_rep = sc->s;
    if (mode != FIN_FORCE) {            // the next step's field (molecular-dynamics.F:66-76)
        sc->i_step = sc->i_step + 1.0;
        e_field_next(sc);
    }
    if (mode == FIN_NP && next_step) np_begin_scalar(k, sc);
}

// First half of the first Nose-Poincare step of a call: KE from the kick
// kernel's partials, then eq. 38.
__device__ __forceinline__ void np_begin_body(const ScalarConsts& k, Scalars* gsc, const double* partials,
                                                 int maxblocks, int kblocks, int bx)
{
#pragma clang fp contract(off)
    __shared__ double red[NT];
    __shared__ Scalars lsc;
    scalars_in(&lsc, gsc);
    const int slots[1] = {P_KEPK};
    double r[1];
    fold_slots<1>(partials, maxblocks, kblocks, slots, r, red);
    if (threadIdx.x == 0) {
        lsc.ke_pk = r[0] / (double)k.n_sum_p * (double)k.N / 2;
        np_begin_scalar(k, &lsc);
    }
    scalars_out(gsc, &lsc);
}

static __global__ __launch_bounds__(NT) void k_np_begin(ScalarConsts k, Scalars* gsc, const double* partials,
                                                 int maxblocks, int kblocks)
{
    np_begin_body(k, gsc, partials, maxblocks, kblocks, blockIdx.x);
}

// Kinetic energies of the current momenta for the energy report
// (kinetic_energy_dipo.F, kinetic_energy_acou.F).
static __global__ __launch_bounds__(NT) void k_ke_report(ScalarConsts k, Scalars* sc, const double* partials,
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
template <typename S>
__device__ __forceinline__ void apply_scale_body(const Consts& c, const FieldsT<S>& f, const Scalars* sc, int bx)
{
    const size_t cc = (size_t)bx * NT + threadIdx.x;
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

template <typename S>
__global__ __launch_bounds__(NT) void k_apply_scale(Consts c, FieldsT<S> f, Scalars* sc)
{
    apply_scale_body(c, f, sc, blockIdx.x);
}

static __global__ void k_reset_scale(Scalars* sc)
{
    sc->scale_p = 1.0;
    sc->scale_pw = 1.0;
}

// ---------------------------------------------------------------------------
// Heat-source and sink layers of leapfrog (leapfrog.F:38-83): the kinetic
// energy of every z layer (one workgroup per layer), then the momenta of
// layer Lz/4 scaled up and of layer 3Lz/4 down so that their energies change
// by +-Lx Ly delta_e (and delta_e_acou for the acoustic momenta).
// ---------------------------------------------------------------------------
template <typename S>
__device__ __forceinline__ void layer_ke_body(const Consts& c, const FieldsT<S>& f, const int* __restrict__ nlayer,
                                              double* el, int iz)
{
    __shared__ double red[NT];
    const int LxLy = c.Lx * c.Ly;
    double ke = 0.0, kea = 0.0;
    for (int i = threadIdx.x; i < LxLy; i += NT) {
        const size_t cc = (size_t)iz * LxLy + i;
        const unsigned char code = f.code[cc];
        const int s = code & 3;
        const double minv = (code & 4) ? 0.0 : c.minv[s];
#pragma unroll
        for (int q = 0; q < 3; ++q) ke += (double)f.p[q][cc] * f.p[q][cc] * minv;   // leapfrog.F:47-54
        if (c.acoustic) {
#pragma unroll
            for (int q = 0; q < 3; ++q) kea += (double)f.pw[q][cc] * f.pw[q][cc] * c.mwinv[s];   // :61-68
        }
    }
    double r[2] = {ke, kea};
    block_sums<2>(r, red);
    if (threadIdx.x == 0) {
        el[iz] = r[0] / 2 * c.Lx * c.Ly / nlayer[iz];                                // :57
        el[c.Lz + iz] = c.acoustic ? r[1] / 2 : 0.0;                                 // :69, :72
    }
}

template <typename S>
__global__ __launch_bounds__(NT) void k_layer_ke(Consts c, FieldsT<S> f, const int* __restrict__ nlayer, double* el)
{
    layer_ke_body(c, f, nlayer, el, blockIdx.x);
}

template <typename S>
__device__ __forceinline__ void layer_scale_body(const Consts& c, const FieldsT<S>& f, const double* __restrict__ el,
                                                 double delta_e, double delta_e_acou, int bx, int by)
{
    const int LxLy = c.Lx * c.Ly;
    const int i = bx * NT + threadIdx.x;
    if (i >= LxLy) return;
    const int q1 = c.Lz / 4, q3 = q1 * 3;
    const int iz = by ? q3 : q1;
    const double sgn = by ? -1.0 : 1.0;
    const size_t cc = (size_t)iz * LxLy + i;
    if (delta_e != 0.0) {                                                            // :75-78
        const double e = el[iz];
        const double fac = sqrt((e + sgn * ((double)(c.Lx * c.Ly) * delta_e)) / e);
#pragma unroll
        for (int q = 0; q < 3; ++q) f.p[q][cc] = f.p[q][cc] * fac;
    }
    if (delta_e_acou != 0.0 && c.acoustic) {                                         // :80-83
        const double e = el[c.Lz + iz];
        const double fac = sqrt((e + sgn * ((double)(c.Lx * c.Ly) * delta_e_acou)) / e);
#pragma unroll
        for (int q = 0; q < 3; ++q) f.pw[q][cc] = f.pw[q][cc] * fac;
    }
}

template <typename S>
__global__ __launch_bounds__(NT) void k_layer_scale(Consts c, FieldsT<S> f, const double* __restrict__ el,
                                                    double delta_e, double delta_e_acou)
{
    layer_scale_body(c, f, el, delta_e, delta_e_acou, blockIdx.x, blockIdx.y);
}

// Sums of the layer energies over the averaging steps (average_module.F:82-85).
__device__ __forceinline__ void layer_acc_body(const double* __restrict__ el, double* sum, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) sum[i] = sum[i] + el[i];
}
static __global__ void k_layer_acc(const double* __restrict__ el, double* sum, int n)
{
    layer_acc_body(el, sum, n);
}

// ---------------------------------------------------------------------------
// Averages of the MD steps (average_module.F:57-163): per step, the sums over
// the counted cells of u, u^2, Z* u and (Z* u)^2, the histograms of u, and
// the accumulation of energies, strain and the products of the sums.
// ---------------------------------------------------------------------------
constexpr int DIST_K = 99;                 // max_k (average_module.F:15)
constexpr int DIST_W = 2 * DIST_K + 1;
constexpr double DIST_STEP_INV = 200.0;    // 1/step, step = 0.005 (average_module.F:13-14)

// Layout shared with fhip_average (all doubles, n_add last).
struct AvgAcc {
    double energy[13];
    double e2, acou_kinetic_true, dipo_kinetic_true;
    double strain[6];
    double u[3], uu[6], p[3], pp[6];
    double u_tmp[3], uu_tmp[3], p_tmp[3], pp_tmp[3];
    double n_add;
};

template <int HIST, typename S>
__device__ __forceinline__ void avg_cells_body(const Consts& c, const FieldsT<S>& f, double* partials, int maxblocks,
                                                  unsigned long long* dist1, int* dist3, int ravg, int bx)
{
    __shared__ double red[NT];
    const size_t cc = (size_t)bx * NT + threadIdx.x;
    double v[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) v[q] = 0.0;
    if (ravg && cc < (size_t)c.N) {                             // molecular-dynamics.F:144-148
#pragma unroll
        for (int q = 0; q < 3; ++q) f.uavg[q][cc] = f.uavg[q][cc] + f.u[q][cc];
    }
    if (cc < (size_t)c.N && (f.code[cc] & 8)) {
        double u[3];
        const double zs = c.Zs[f.code[cc] & 3];
#pragma unroll
        for (int q = 0; q < 3; ++q) u[q] = f.u[q][cc];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double zu = zs * u[q];                        // Z_u_R (dipoFFT.F:35-43)
            v[q] = u[q];                                        // average_module.F:94-95
            v[3 + q] = u[q] * u[q];
            v[6 + q] = zu;                                      // average_module.F:115-116
            v[9 + q] = zu * zu;
        }
        if (HIST) {                                             // average_module.F:130-160
            long k[3];
            bool in3 = true;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                k[q] = lround(u[q] * DIST_STEP_INV);            // Fortran nint: half away from zero
                const bool in = k[q] >= -DIST_K && k[q] <= DIST_K;
                if (in) atomicAdd(&dist1[q * DIST_W + (int)k[q] + DIST_K], 1ull);
                in3 = in3 && in;
            }
            if (in3)
                atomicAdd(&dist3[((int)k[0] + DIST_K) + DIST_W * (((int)k[1] + DIST_K) + DIST_W * ((int)k[2] + DIST_K))], 1);
        }
    }
    block_sums<12>(v, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 12; ++q) partials[(P_AVG + q) * maxblocks + bx] = v[q];
    }
}

template <int HIST, typename S>
__global__ __launch_bounds__(NT) void k_avg_cells(Consts c, FieldsT<S> f, double* partials, int maxblocks,
                                                  unsigned long long* dist1, int* dist3, int ravg)
{
    avg_cells_body<HIST, S>(c, f, partials, maxblocks, dist1, dist3, ravg, blockIdx.x);
}

__device__ __forceinline__ void avg_fold_body(const Scalars* sc, AvgAcc* acc,
                                                 const double* partials, int maxblocks, int kblocks, int bx)
{
#pragma clang fp contract(off)
    __shared__ double red[NT];
    int slots[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) slots[q] = P_AVG + q;
    double r[12];
    fold_slots<12>(partials, maxblocks, kblocks, slots, r, red);
    if (threadIdx.x != 0) return;
    AvgAcc a = *acc;
    const double* en = sc->en_rep;
    for (int i = 0; i < 13; ++i) a.energy[i] = a.energy[i] + en[i];                 // average_module.F:76
    a.e2 = a.e2 + en[E_TOTAL] * en[E_TOTAL];                                      // :77
    a.dipo_kinetic_true = a.dipo_kinetic_true + en[E_DIPO_KIN] / (sc->s_rep * sc->s_rep);   // :79
    for (int i = 0; i < 6; ++i) a.strain[i] = a.strain[i] + sc->strain[i];        // :80
    const double* ut = r;
    const double* pt = r + 6;
    for (int q = 0; q < 3; ++q) {
        a.u[q] = a.u[q] + ut[q];                                                  // :101
        a.uu[q] = a.uu[q] + ut[q] * ut[q];                                        // :102-104
        a.p[q] = a.p[q] + pt[q];                                                  // :122
        a.pp[q] = a.pp[q] + pt[q] * pt[q];                                        // :123-125
        a.u_tmp[q] = ut[q];
        a.uu_tmp[q] = r[3 + q];
        a.p_tmp[q] = pt[q];
        a.pp_tmp[q] = r[9 + q];
    }
    a.uu[3] = a.uu[3] + ut[1] * ut[2];                                            // :105-107
    a.uu[4] = a.uu[4] + ut[2] * ut[0];
    a.uu[5] = a.uu[5] + ut[0] * ut[1];
    a.pp[3] = a.pp[3] + pt[1] * pt[2];                                            // :126-128
    a.pp[4] = a.pp[4] + pt[2] * pt[0];
    a.pp[5] = a.pp[5] + pt[0] * pt[1];
    a.n_add = a.n_add + 1.0;
    *acc = a;
}

static __global__ __launch_bounds__(NT) void k_avg_fold(const Scalars* __restrict__ sc, AvgAcc* acc,
                                                 const double* partials, int maxblocks, int kblocks)
{
    avg_fold_body(sc, acc, partials, maxblocks, kblocks, blockIdx.x);
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
static __global__ __launch_bounds__(NT) void k_ewald_real(Consts c, EwaldArgs e, double* out /*[6][N]*/)
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

// Value at the point (kx, ky, kz), 0 <= kx <= Mx, of a transformed packed
// spectrum: pencils with kx = 0 carry P (the kx = 0 plane), Q (kx = Lx/2)
// and, at ky = 0 and Ly/2, R = P + iQ with P and Q Hermitian along kz.
template <typename T>
__device__ __forceinline__ cpx<double> spec_at(const Consts& c, const cpx<T>* s, int kx, int ky, int kz)
{
    auto at = [&](int x, int y, int z) {
        const cpx<T> v = s[((size_t)z * c.Ly + y) * c.Mx + x];
        return cpx<double>{(double)v.x, (double)v.y};
    };
    if (kx != 0 && kx != c.Mx) return at(kx, ky, kz);
    const int kzp = kz == 0 ? 0 : c.Lz - kz;
    cpx<double> P, Q;
    if (ky != 0 && 2 * ky != c.Ly) {
        if (2 * ky < c.Ly) {
            P = at(0, ky, kz);
            Q = at(0, c.Ly - ky, kz);
        } else {
            const cpx<double> p2 = at(0, c.Ly - ky, kzp), q2 = at(0, ky, kzp);
            P = {p2.x, -p2.y};
            Q = {q2.x, -q2.y};
        }
    } else {
        const cpx<double> r = at(0, ky, kz), rp = at(0, ky, kzp);
        P = {(r.x + rp.x) * 0.5, (r.y - rp.y) * 0.5};
        Q = {(r.y + rp.y) * 0.5, (rp.x - r.x) * 0.5};
    }
    return kx == 0 ? P : Q;
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
        iy = (int)(r / c.Lz);
        iz = (int)(r % c.Lz);
    }
    double k3 = (double)iz / c.Lz; if (k3 > 0.5) k3 = k3 - 1.0;
    double k2 = (double)iy / c.Ly; if (k2 > 0.5) k2 = k2 - 1.0;
    const double k1 = (double)ix / c.Lx;
    const int ia[6] = {0, 1, 2, 1, 2, 0}, ib[6] = {0, 1, 2, 2, 0, 1};
    for (int m = 0; m < 6; ++m) {
        double acc = spec_at(c, sp.s[m], ix, iy, iz).x;
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

// Point (kx, ky, kz), 0 <= kx <= Mx, of the transformed packed spectrum s.
__device__ __forceinline__ void kdecode(const Consts& c, size_t kk, int& kx, int& ky, int& kz)
{
    if (kk < c.Mh) {
        kx = (int)(kk % c.Mx);
        const size_t r = kk / c.Mx;
        ky = (int)(r % c.Ly);
        kz = (int)(r / c.Ly);
    } else {
        const size_t r = kk - c.Mh;
        kx = c.Mx;
        ky = (int)(r / c.Lz);
        kz = (int)(r % c.Lz);
    }
}

// The transformed packed spectrum s on the k array (kindex order).
template <typename T>
__global__ void k_unpack(Consts c, const cpx<T>* __restrict__ s, cpx<double>* __restrict__ out)
{
    const size_t kk = (size_t)blockIdx.x * NT + threadIdx.x;
    if (kk >= (size_t)c.Kh) return;
    int kx, ky, kz;
    kdecode(c, kk, kx, ky, kz);
    out[kk] = spec_at(c, s, kx, ky, kz);
}

// w reconstructed from the optimised w(k) (coord_module.F:457-469): slots
// 0..2 <- the packed Hermitian part of G = -i acouK, which is what a c2r of
// the half spectrum -i acouK returns.
template <typename T>
__global__ void k_prep_w(Consts c, Spectra<T> sp, const cpx<double>* acouK)
{
    const size_t p = (size_t)blockIdx.x * NT + threadIdx.x;
    if (p >= c.Mh) return;
    const int x = (int)(p % c.Mx);
    const size_t r = p / c.Mx;
    const int y = (int)(r % c.Ly), z = (int)(r / c.Ly);
    const int zp = z == 0 ? 0 : c.Lz - z;
    for (int q = 0; q < 3; ++q) {
        const cpx<double>* aK = acouK + (size_t)q * c.Kh;
        auto G = [&](int kx, int ky, int kz) {
            const cpx<double> a = aK[kindex(c.Mx, c.Ly, c.Lz, c.Mh, kx, ky, kz)];
            return cpx<double>{a.y, -a.x};
        };
        // (G(k) + conj G(partner)) / 2
        auto H = [&](int kx, int ky, int kz, int py, int pz) {
            const cpx<double> g = G(kx, ky, kz), h = G(kx, py, pz);
            return cpx<double>{(g.x + h.x) * 0.5, (g.y - h.y) * 0.5};
        };
        cpx<double> v;
        if (x != 0) v = G(x, y, z);
        else if (y != 0 && 2 * y < c.Ly) v = H(0, y, z, c.Ly - y, zp);
        else if (2 * y > c.Ly) v = H(c.Mx, c.Ly - y, z, y, zp);
        else {
            const cpx<double> hp = H(0, y, z, y, zp), hq = H(c.Mx, y, z, y, zp);
            v = {hp.x - hq.y, hp.y + hq.x};
        }
        sp.s[q][p] = cpx<T>{(T)v.x, (T)v.y};
    }
}

// ---------------------------------------------------------------------------
// Replica batches.  Replicas of one size class (edges, precision, acoustic
// degrees of freedom) share every kernel instantiation and grid, so each stage
// of a step is one launch for all of them: blockIdx.z picks the replica, whose
// arguments sit in a device array of Rep records; the stage bodies are the
// single-replica ones.
// ---------------------------------------------------------------------------
template <typename T>
struct Rep {
    Consts c;
    ScalarConsts k;
    FieldsT<T> f;
    Spectra<T> sp;
    ZArgs<T> za;                 // acouK set; a z pass given want_acouK = 0 drops it
    XfwdArgs xa;
    XinvArgs ia;                 // materialize set per launch
    Scalars* sc;
    double* partials;
    const cpx<T>* twx;
    const cpx<T>* twy;
    const cpx<T>* twz;
    AvgAcc* acc;
    unsigned long long* dist1;
    int* dist3;
    int maxblocks, xblocks, zblocks, kblocks, ravg;
    // heat-source layers of a leapfrog replica (delta_e = delta_e_acou = 0: none)
    double* el;
    double* el_sum;
    const int* nlayer;
    double delta_e, delta_e_acou;
};

template <typename T, int LX, int RB>
__global__ __launch_bounds__(NT) void kb_xfwd(const Rep<T>* __restrict__ reps, int phase, int mode)
{
    const Rep<T>& r = reps[blockIdx.z];
    XfwdArgs a{};                          // no generic input fields in a batch
    a.mode = mode >= 0 ? mode : r.xa.mode; // XF_FORCE: no drift (the tuner's timing runs)
    a.nf = r.xa.nf;
    a.phase = phase;
    xfwd_body<T, LX, RB, false>(r.c, r.f, r.sp, a, r.sc, r.twx, blockIdx.x);
}

template <typename T, int LY, int SIGN, int TB>
__global__ __launch_bounds__(NT) void kb_ypass(const Rep<T>* __restrict__ reps)
{
    const Rep<T>& r = reps[blockIdx.z];
    ypass_body<T, LY, SIGN, TB>(r.c, r.sp, 0, r.twy, blockIdx.x, blockIdx.y);
}

template <typename T, int LZ, int GROUP>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(GROUP == ZG_ACOU ? 1 : 2)))
void kb_zpass(const Rep<T>* __restrict__ reps, int want_acouK)
{
    const Rep<T>& r = reps[blockIdx.z];
    ZArgs<T> a = r.za;
    if (!want_acouK) a.acouK = nullptr;
    __shared__ ZLds<T, LZ, zbuf_size<T, LZ, GROUP>()> lds;
    cpx<T> pre[ZPF][ZTile<T, LZ>::PPT];
    zpass_body<T, LZ, GROUP, 0, -1>(r.c, r.sp, a, r.sc, r.twz, blockIdx.x, lds, pre);
}

template <typename T, int LZ, int G2, int DDG = ZG_DD>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(G2 == ZG_ACOU ? 1 : 2)))
void kb_zpair(const Rep<T>* __restrict__ reps, int want_acouK)
{
    const Rep<T>& r = reps[blockIdx.z];
    ZArgs<T> a = r.za;
    if (!want_acouK) a.acouK = nullptr;
    zpair_body<T, LZ, G2, DDG>(r.c, r.sp, a, r.sc, r.twz, blockIdx.x);
}

template <typename T, int LX, int NF>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NF == 12 ? 2 : 3))) void kb_xinv(const Rep<T>* __restrict__ reps, int materialize)
{
    const Rep<T>& r = reps[blockIdx.z];
    XinvArgs a = r.ia;
    a.materialize = materialize;
    if constexpr (NF >= 9) xinv_step_body<T, LX, NF>(r.c, r.f, r.sp, a, r.sc, r.twx, blockIdx.x);
    else xinv_body<T, LX, NF>(r.c, r.f, r.sp, a, r.sc, r.twx, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(NT) void kb_mirror_kick(const Rep<T>* __restrict__ reps, int mode, int materialize)
{
    const Rep<T>& r = reps[blockIdx.z];
    mirror_kick_body(r.c, r.f, r.sc, mode, materialize, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(NT) void kb_homo(const Rep<T>* __restrict__ reps)
{
    const Rep<T>& r = reps[blockIdx.z];
    homo_body(r.k, r.sc, r.partials, r.maxblocks, r.zblocks, 0);
}

template <typename T>
__global__ __launch_bounds__(NT) void kb_finish(const Rep<T>* __restrict__ reps, int mode, int next_step)
{
    const Rep<T>& r = reps[blockIdx.z];
    finish_body(r.k, r.sc, r.partials, r.maxblocks, r.xblocks, r.zblocks, mode, r.c.acoustic, next_step, 0);
}

template <typename T, int KICK>
__global__ __launch_bounds__(NT) void kb_kick(const Rep<T>* __restrict__ reps)
{
    const Rep<T>& r = reps[blockIdx.z];
    kick_body<KICK, T>(r.c, r.f, r.sc, r.partials, r.maxblocks, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(NT) void kb_np_begin(const Rep<T>* __restrict__ reps)
{
    const Rep<T>& r = reps[blockIdx.z];
    np_begin_body(r.k, r.sc, r.partials, r.maxblocks, r.kblocks, 0);
}

template <typename T>
__global__ __launch_bounds__(NT) void kb_apply_scale(const Rep<T>* __restrict__ reps)
{
    const Rep<T>& r = reps[blockIdx.z];
    apply_scale_body(r.c, r.f, r.sc, blockIdx.x);
}

template <typename T>
__global__ void kb_reset_scale(const Rep<T>* __restrict__ reps)
{
    Scalars* sc = reps[blockIdx.z].sc;
    sc->scale_p = 1.0;
    sc->scale_pw = 1.0;
}

template <typename T, int HIST>
__global__ __launch_bounds__(NT) void kb_avg_cells(const Rep<T>* __restrict__ reps)
{
    const Rep<T>& r = reps[blockIdx.z];
    avg_cells_body<HIST, T>(r.c, r.f, r.partials, r.maxblocks, r.dist1, r.dist3, r.ravg, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(NT) void kb_avg_fold(const Rep<T>* __restrict__ reps)
{
    const Rep<T>& r = reps[blockIdx.z];
    avg_fold_body(r.sc, r.acc, r.partials, r.maxblocks, r.kblocks, 0);
}

// Heat-source layers of the leapfrog replicas that have them (grid x: the
// largest Lz of the class for the layer energies, the x tiles of a layer for
// the rescale; other replicas return at once).
template <typename T>
__global__ __launch_bounds__(NT) void kb_layer_ke(const Rep<T>* __restrict__ reps)
{
    const Rep<T>& r = reps[blockIdx.z];
    if ((r.delta_e == 0.0 && r.delta_e_acou == 0.0) || (int)blockIdx.x >= r.c.Lz) return;
    layer_ke_body(r.c, r.f, r.nlayer, r.el, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(NT) void kb_layer_scale(const Rep<T>* __restrict__ reps)
{
    const Rep<T>& r = reps[blockIdx.z];
    if (r.delta_e == 0.0 && r.delta_e_acou == 0.0) return;
    layer_scale_body(r.c, r.f, r.el, r.delta_e, r.delta_e_acou, blockIdx.x, blockIdx.y);
}

template <typename T>
__global__ __launch_bounds__(NT) void kb_layer_acc(const Rep<T>* __restrict__ reps)
{
    const Rep<T>& r = reps[blockIdx.z];
    if (r.delta_e == 0.0) return;                   // as the per-context averages
    layer_acc_body(r.el, r.el_sum, 2 * r.c.Lz);
}

}  // namespace fhip
