/*
 * feram_oracle.c -- CPU FP64 restatement of feram's per-step force-and-integrate
 * path.  TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.
 *
 * Each function follows the reference routine cited above it in operation
 * order (including the left-to-right evaluation order of Fortran expressions),
 * so that sums and products round the way the reference rounds them.  The FFT
 * is this file's own mixed-radix DFT with FFTW's sign and normalisation
 * conventions; FFTW itself is not available here.
 */
#include "feram_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define M_PI_D       3.14159265358979323846
#define M_2PI_D      6.28318530717958647693
#define M_2_SQRTPI_D 1.12837916709551257390
#define HARTREE_BOHR 14.399644
#define KELVIN_EV    11604.45
#define ELEMENTARY_CHARGE 1.60217653

struct fo_ctx {
    fo_params p;
    int N, Kx, Kh, n_freedom;
    double Ninv, dt_2, kappa, unit_cell_volume, pNa03[3];
    double target_kinetic_energy, temperature_eV;
    /* real-space fields (3N unless noted) */
    double *dipoR, *dipoP, *dVddi, *dVddi_long, *acouR, *acouP, *dVdacR;
    double *E_field, *Z_u_R, *didiR /*6N*/, *inhoR /*6N*/;
    double *Z_starR, *mass_dipo_inv, *mass_acou_inv; /* N */
    int *relaxor;
    unsigned char *sum_p;
    int n_sum_p;
    /* spectra (interleaved complex) */
    double *dipoK, *Z_u_K, *dipoT, *didiK /*6*/, *acouK, *inhoK /*6*/, *dVdacK;
    double *modulation_k /*1*/, *c_k_modulation_k /*3*/;
    /* k tables */
    double *long_range, *short_range;   /* 9 per k */
    double *inhomo_K;                   /* 9 per k */
    double *inhomo_K_G;                 /* 12 per k: (a,b) a fastest, b=1..4 */
    double *inhomo_coup2_K;             /* 18 per k: (a,i) a fastest */
    /* homogeneous */
    double strain[6];
    double homogeneous[36], homogeneous_G[36], homogeneous_G_inv[6];
    double homogeneous_coupling_2[36];  /* col-major (i,j) -> [i + 6*j] */
    double nose[3];                     /* s_Nose, pi_Nose, H0 */
    /* heat-source layers (leapfrog.F:38-73) */
    double delta_e, delta_e_acou;
    double *e_layer_dipo, *e_layer_acou;   /* Lz */
    int *n_sum_layer;                      /* Lz: counted cells per z layer */
    /* per z-plane partial sums of the reductions (3 x Lz) */
    double *planes;
};

/* Reductions over cells or k points run in parallel as one partial sum per
 * z plane (each in the reference's loop order), and the planes are added in
 * order: the result does not depend on the number of threads. */
static double *plane_buf(fo_ctx *o, int which) { return o->planes + (size_t)which * o->p.Lz; }
static double plane_sum(const double *part, int Lz)
{
    double t = 0.0;
    for (int iz = 0; iz < Lz; ++iz) t = t + part[iz];
    return t;
}

#define H6(a, i, j) ((a)[(i) + 6 * (j)])   /* 0-based (i,j) */

/* ------------------------------------------------------------------------- */
/* 1-D mixed-radix complex DFT: Stockham autosort passes (radix 4, 2, then   */
/* 3, 5 and any other prime as a plain DFT of that length), applied to B     */
/* interleaved lines at once (element t of line b at index t*B + b), with    */
/* per-pass twiddle tables.  FFTW's conventions: sign -1 forward, +1 inverse, */
/* neither normalised.                                                       */
/* ------------------------------------------------------------------------- */
#define FFT_MAXPASS 40
#define FFT_MAXRADIX 256     /* largest prime factor of a transform length */
typedef struct {
    int n, sign, npass;
    int r[FFT_MAXPASS], ns[FFT_MAXPASS];
    double *tw[FFT_MAXPASS];     /* pass s: [k < ns][q = 1 .. r-1] of e^{sign 2 pi i q k / (ns r)} */
    double *root[FFT_MAXPASS];   /* pass s: [j < r] of e^{sign 2 pi i j / r} (plain DFT passes) */
} fft_plan;

static void plan_make(fft_plan *pl, int n, int sign)
{
    memset(pl, 0, sizeof(*pl));
    pl->n = n;
    pl->sign = sign;
    int rem = n, ns = 1;
    while (rem > 1) {
        int r;
        if (rem % 4 == 0) r = 4;
        else if (rem % 2 == 0) r = 2;
        else {
            r = rem;
            for (int f = 3; f * f <= rem; f += 2)
                if (rem % f == 0) { r = f; break; }
        }
        if (r > FFT_MAXRADIX) {
            fprintf(stderr, "feram_oracle: FFT length %d has a prime factor above %d\n", n, FFT_MAXRADIX);
            abort();
        }
        const int s = pl->npass++;
        pl->r[s] = r;
        pl->ns[s] = ns;
        pl->tw[s] = (double *)malloc(sizeof(double) * 2 * (size_t)ns * (r > 1 ? r - 1 : 1));
        for (int k = 0; k < ns; ++k)
            for (int q = 1; q < r; ++q) {
                const double a = M_2PI_D * (double)(q * k) / (double)(ns * r);
                pl->tw[s][2 * (k * (r - 1) + q - 1)] = cos(a);
                pl->tw[s][2 * (k * (r - 1) + q - 1) + 1] = sign * sin(a);
            }
        pl->root[s] = (double *)malloc(sizeof(double) * 2 * r);
        for (int j = 0; j < r; ++j) {
            const double a = M_2PI_D * (double)j / (double)r;
            pl->root[s][2 * j] = cos(a);
            pl->root[s][2 * j + 1] = sign * sin(a);
        }
        ns *= r;
        rem /= r;
    }
}

static void plan_free(fft_plan *pl)
{
    for (int s = 0; s < pl->npass; ++s) { free(pl->tw[s]); free(pl->root[s]); }
    pl->npass = 0;
}

/* one Stockham pass of radix r at stride ns on B lines: in -> out */
static void fft_pass(const fft_plan *pl, int s, const double *in, double *out, int B)
{
    const int n = pl->n, r = pl->r[s], ns = pl->ns[s], m = n / r, sg = pl->sign;
    const double *tw = pl->tw[s], *root = pl->root[s];
    double vr[FFT_MAXRADIX], vi[FFT_MAXRADIX];
    for (int j = 0; j < m; ++j) {
        const int k = j % ns;
        const int od = (j / ns) * ns * r + k;
        const double *w = tw + 2 * k * (r - 1);
        for (int b = 0; b < B; ++b) {
            for (int q = 0; q < r; ++q) {
                const size_t ii = 2 * ((size_t)(j + q * m) * B + b);
                double xr = in[ii], xi = in[ii + 1];
                if (q > 0 && ns > 1) {
                    const double wr = w[2 * (q - 1)], wi = w[2 * (q - 1) + 1];
                    const double tr = xr * wr - xi * wi;
                    xi = xr * wi + xi * wr;
                    xr = tr;
                }
                vr[q] = xr;
                vi[q] = xi;
            }
            if (r == 2) {
                const double ar = vr[0], ai = vi[0];
                vr[0] = ar + vr[1]; vi[0] = ai + vi[1];
                vr[1] = ar - vr[1]; vi[1] = ai - vi[1];
            } else if (r == 4) {
                /* X1 = (v0 - v2) + w (v1 - v3), w = sign i */
                const double s02r = vr[0] + vr[2], s02i = vi[0] + vi[2];
                const double d02r = vr[0] - vr[2], d02i = vi[0] - vi[2];
                const double s13r = vr[1] + vr[3], s13i = vi[1] + vi[3];
                const double d13r = vr[1] - vr[3], d13i = vi[1] - vi[3];
                const double er = -sg * d13i, ei = sg * d13r;     /* (sign i) d13 */
                vr[0] = s02r + s13r; vi[0] = s02i + s13i;
                vr[2] = s02r - s13r; vi[2] = s02i - s13i;
                vr[1] = d02r + er;   vi[1] = d02i + ei;
                vr[3] = d02r - er;   vi[3] = d02i - ei;
            } else {
                double xr[FFT_MAXRADIX], xi[FFT_MAXRADIX];
                for (int q = 0; q < r; ++q) { xr[q] = vr[q]; xi[q] = vi[q]; }
                for (int t = 0; t < r; ++t) {
                    double ar = 0.0, ai = 0.0;
                    for (int q = 0; q < r; ++q) {
                        const int e = (q * t) % r;
                        const double wr = root[2 * e], wi = root[2 * e + 1];
                        ar += xr[q] * wr - xi[q] * wi;
                        ai += xr[q] * wi + xi[q] * wr;
                    }
                    vr[t] = ar;
                    vi[t] = ai;
                }
            }
            for (int q = 0; q < r; ++q) {
                const size_t oo = 2 * ((size_t)(od + q * ns) * B + b);
                out[oo] = vr[q];
                out[oo + 1] = vi[q];
            }
        }
    }
}

/* transform B interleaved lines held in buf (work: same size); result in buf */
static void fft_lines(const fft_plan *pl, double *buf, double *work, int B)
{
    double *a = buf, *b = work;
    for (int s = 0; s < pl->npass; ++s) {
        fft_pass(pl, s, a, b, B);
        double *t = a; a = b; b = t;
    }
    if (a != buf) memcpy(buf, a, sizeof(double) * 2 * (size_t)pl->n * B);
}

void fo_fft1d(double *x, int n, int sign)
{
    fft_plan pl;
    plan_make(&pl, n, sign);
    double *work = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    fft_lines(&pl, x, work, 1);
    free(work);
    plan_free(&pl);
}

/* lines transformed together in the 3-D passes (cache lines of the strided
 * y and z gathers are used whole) */
#define FFT_B 8

/* Rank-3 r2c with FFTW's n=(Lz,Ly,Lx) row-major semantics: x is the real
 * (halved) dimension (coord_module.F:151-157, dipoFFT.F:19). */
void fo_r2c(fo_ctx *o, const double *in, int ncomp, int comp, double *out)
{
    const int Lx = o->p.Lx, Ly = o->p.Ly, Lz = o->p.Lz, Kx = o->Kx;
    fft_plan px, py, pz;
    plan_make(&px, Lx, -1);
    plan_make(&py, Ly, -1);
    plan_make(&pz, Lz, -1);
    int nmax = Lx > Ly ? Lx : Ly; if (Lz > nmax) nmax = Lz;
    const int nyb = (Ly + FFT_B - 1) / FFT_B, nkb = (Kx + FFT_B - 1) / FFT_B;
#define SPEC(kx, y, z) (out + 2 * ((size_t)ncomp * ((kx) + (size_t)Kx * ((y) + (size_t)Ly * (z))) + comp))
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)nmax * FFT_B);
        double *work = (double *)malloc(sizeof(double) * 2 * (size_t)nmax * FFT_B);
#pragma omp for collapse(2) schedule(static)
        for (int z = 0; z < Lz; ++z)
            for (int yb = 0; yb < nyb; ++yb) {
                const int y0 = yb * FFT_B, B = (Ly - y0 < FFT_B) ? Ly - y0 : FFT_B;
                for (int x = 0; x < Lx; ++x)
                    for (int b = 0; b < B; ++b) {
                        buf[2 * (x * B + b)] = in[(size_t)ncomp * (x + Lx * (y0 + b + (size_t)Ly * z)) + comp];
                        buf[2 * (x * B + b) + 1] = 0.0;
                    }
                fft_lines(&px, buf, work, B);
                for (int b = 0; b < B; ++b)
                    for (int kx = 0; kx < Kx; ++kx) {
                        double *d = SPEC(kx, y0 + b, z);
                        d[0] = buf[2 * (kx * B + b)];
                        d[1] = buf[2 * (kx * B + b) + 1];
                    }
            }
#pragma omp for collapse(2) schedule(static)
        for (int z = 0; z < Lz; ++z)
            for (int kb = 0; kb < nkb; ++kb) {
                const int k0 = kb * FFT_B, B = (Kx - k0 < FFT_B) ? Kx - k0 : FFT_B;
                for (int y = 0; y < Ly; ++y)
                    for (int b = 0; b < B; ++b) {
                        const double *d = SPEC(k0 + b, y, z);
                        buf[2 * (y * B + b)] = d[0];
                        buf[2 * (y * B + b) + 1] = d[1];
                    }
                fft_lines(&py, buf, work, B);
                for (int y = 0; y < Ly; ++y)
                    for (int b = 0; b < B; ++b) {
                        double *d = SPEC(k0 + b, y, z);
                        d[0] = buf[2 * (y * B + b)];
                        d[1] = buf[2 * (y * B + b) + 1];
                    }
            }
#pragma omp for collapse(2) schedule(static)
        for (int y = 0; y < Ly; ++y)
            for (int kb = 0; kb < nkb; ++kb) {
                const int k0 = kb * FFT_B, B = (Kx - k0 < FFT_B) ? Kx - k0 : FFT_B;
                for (int z = 0; z < Lz; ++z)
                    for (int b = 0; b < B; ++b) {
                        const double *d = SPEC(k0 + b, y, z);
                        buf[2 * (z * B + b)] = d[0];
                        buf[2 * (z * B + b) + 1] = d[1];
                    }
                fft_lines(&pz, buf, work, B);
                for (int z = 0; z < Lz; ++z)
                    for (int b = 0; b < B; ++b) {
                        double *d = SPEC(k0 + b, y, z);
                        d[0] = buf[2 * (z * B + b)];
                        d[1] = buf[2 * (z * B + b) + 1];
                    }
            }
        free(buf); free(work);
    }
#undef SPEC
    plan_free(&px); plan_free(&py); plan_free(&pz);
}

/* Rank-3 c2r (unnormalised): complex inverse along z and y, then the
 * half-complex inverse along x, which uses only Re of kx=0 and kx=Lx/2. */
void fo_c2r(fo_ctx *o, const double *in, int ncomp, int comp, double *out)
{
    const int Lx = o->p.Lx, Ly = o->p.Ly, Lz = o->p.Lz, Kx = o->Kx;
    const size_t Kh = (size_t)o->Kh;
    double *tmp = (double *)malloc(sizeof(double) * 2 * Kh);
    fft_plan px, py, pz;
    plan_make(&px, Lx, +1);
    plan_make(&py, Ly, +1);
    plan_make(&pz, Lz, +1);
    int nmax = Lx > Ly ? Lx : Ly; if (Lz > nmax) nmax = Lz;
    const int nyb = (Ly + FFT_B - 1) / FFT_B, nkb = (Kx + FFT_B - 1) / FFT_B;
    for (size_t k = 0; k < Kh; ++k) {
        tmp[2 * k] = in[2 * (ncomp * k + comp)];
        tmp[2 * k + 1] = in[2 * (ncomp * k + comp) + 1];
    }
#define TMP(kx, y, z) (tmp + 2 * ((kx) + (size_t)Kx * ((y) + (size_t)Ly * (z))))
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)nmax * FFT_B);
        double *work = (double *)malloc(sizeof(double) * 2 * (size_t)nmax * FFT_B);
#pragma omp for collapse(2) schedule(static)
        for (int y = 0; y < Ly; ++y)
            for (int kb = 0; kb < nkb; ++kb) {
                const int k0 = kb * FFT_B, B = (Kx - k0 < FFT_B) ? Kx - k0 : FFT_B;
                for (int z = 0; z < Lz; ++z)
                    for (int b = 0; b < B; ++b) {
                        const double *d = TMP(k0 + b, y, z);
                        buf[2 * (z * B + b)] = d[0];
                        buf[2 * (z * B + b) + 1] = d[1];
                    }
                fft_lines(&pz, buf, work, B);
                for (int z = 0; z < Lz; ++z)
                    for (int b = 0; b < B; ++b) {
                        double *d = TMP(k0 + b, y, z);
                        d[0] = buf[2 * (z * B + b)];
                        d[1] = buf[2 * (z * B + b) + 1];
                    }
            }
#pragma omp for collapse(2) schedule(static)
        for (int z = 0; z < Lz; ++z)
            for (int kb = 0; kb < nkb; ++kb) {
                const int k0 = kb * FFT_B, B = (Kx - k0 < FFT_B) ? Kx - k0 : FFT_B;
                for (int y = 0; y < Ly; ++y)
                    for (int b = 0; b < B; ++b) {
                        const double *d = TMP(k0 + b, y, z);
                        buf[2 * (y * B + b)] = d[0];
                        buf[2 * (y * B + b) + 1] = d[1];
                    }
                fft_lines(&py, buf, work, B);
                for (int y = 0; y < Ly; ++y)
                    for (int b = 0; b < B; ++b) {
                        double *d = TMP(k0 + b, y, z);
                        d[0] = buf[2 * (y * B + b)];
                        d[1] = buf[2 * (y * B + b) + 1];
                    }
            }
#pragma omp for collapse(2) schedule(static)
        for (int z = 0; z < Lz; ++z)
            for (int yb = 0; yb < nyb; ++yb) {
                const int y0 = yb * FFT_B, B = (Ly - y0 < FFT_B) ? Ly - y0 : FFT_B;
                for (int b = 0; b < B; ++b) {
                    const double *row = TMP(0, y0 + b, z);
                    for (int kx = 0; kx < Kx; ++kx) {
                        buf[2 * (kx * B + b)] = row[2 * kx];
                        buf[2 * (kx * B + b) + 1] = row[2 * kx + 1];
                    }
                    buf[2 * b + 1] = 0.0;
                    if (Lx % 2 == 0) buf[2 * ((Lx / 2) * B + b) + 1] = 0.0;
                    for (int kx = Kx; kx < Lx; ++kx) {
                        buf[2 * (kx * B + b)] = buf[2 * ((Lx - kx) * B + b)];
                        buf[2 * (kx * B + b) + 1] = -buf[2 * ((Lx - kx) * B + b) + 1];
                    }
                }
                fft_lines(&px, buf, work, B);
                for (int b = 0; b < B; ++b)
                    for (int x = 0; x < Lx; ++x)
                        out[(size_t)ncomp * (x + Lx * (y0 + b + (size_t)Ly * z)) + comp] = buf[2 * (x * B + b)];
            }
        free(buf); free(work);
    }
#undef TMP
    free(tmp);
    plan_free(&px); plan_free(&py); plan_free(&pz);
}

/* ------------------------------------------------------------------------- */
/* Parameters and allocation                                                  */
/* ------------------------------------------------------------------------- */
static double *zalloc(size_t n)
{
    double *a = (double *)calloc(n ? n : 1, sizeof(double));
    return a;
}

static void mirror(fo_ctx *o, double *f);

fo_ctx *fo_create(const fo_params *pp)
{
    fo_ctx *o = (fo_ctx *)calloc(1, sizeof(fo_ctx));
    o->p = *pp;
    const fo_params *p = &o->p;
    /* make_Param (param_module.F:452-480) */
    o->N = p->Lx * p->Ly * p->Lz;
    o->Ninv = 1.0 / o->N;
    o->dt_2 = p->dt / 2;
    o->n_freedom = 3 * o->N;
    o->kappa = p->kappa;
    if (o->kappa <= 0.0)
        o->kappa = sqrt(M_PI_D) / p->a0 / pow((double)o->N, 1.0 / 6);
    o->unit_cell_volume = p->a0 * p->a0 * p->a0;
    for (int i = 0; i < 3; ++i)
        o->pNa03[i] = p->GPa[i] * 0.01 / ELEMENTARY_CHARGE * o->N * o->unit_cell_volume;
    o->target_kinetic_energy = p->kelvin * 1.5 * o->N / KELVIN_EV;
    o->temperature_eV = p->kelvin / KELVIN_EV;

    o->Kx = p->Lx / 2 + 1;
    o->Kh = o->Kx * p->Ly * p->Lz;
    size_t N = o->N, Kh = o->Kh;
    o->dipoR = zalloc(3 * N); o->dipoP = zalloc(3 * N); o->dVddi = zalloc(3 * N);
    o->dVddi_long = zalloc(3 * N); o->acouR = zalloc(3 * N); o->acouP = zalloc(3 * N);
    o->dVdacR = zalloc(3 * N); o->E_field = zalloc(3 * N); o->Z_u_R = zalloc(3 * N);
    o->didiR = zalloc(6 * N); o->inhoR = zalloc(6 * N);
    o->Z_starR = zalloc(N); o->mass_dipo_inv = zalloc(N); o->mass_acou_inv = zalloc(N);
    o->relaxor = (int *)calloc(N, sizeof(int));
    o->sum_p = (unsigned char *)calloc(N, 1);
    o->dipoK = zalloc(6 * Kh); o->Z_u_K = zalloc(6 * Kh); o->dipoT = zalloc(6 * Kh);
    o->didiK = zalloc(12 * Kh); o->acouK = zalloc(6 * Kh); o->inhoK = zalloc(12 * Kh);
    o->dVdacK = zalloc(6 * Kh); o->modulation_k = zalloc(2 * Kh);
    o->c_k_modulation_k = zalloc(6 * Kh);
    o->long_range = zalloc(9 * Kh); o->short_range = zalloc(9 * Kh);
    o->inhomo_K = zalloc(9 * Kh); o->inhomo_K_G = zalloc(12 * Kh);
    o->inhomo_coup2_K = zalloc(18 * Kh);
    o->e_layer_dipo = zalloc(p->Lz); o->e_layer_acou = zalloc(p->Lz);
    o->n_sum_layer = (int *)calloc(p->Lz, sizeof(int));
    o->planes = zalloc(3 * (size_t)p->Lz);
    /* alloc_Coord (coord_module.F:99-143) and read_defects_u defaults */
    for (size_t c = 0; c < N; ++c) {
        o->mass_dipo_inv[c] = 1.0 / p->mass_dipo[0];
        o->mass_acou_inv[c] = 1.0 / p->mass_acou[0];
        o->relaxor[c] = 1;
        o->Z_starR[c] = p->Z_star[0];
        o->sum_p[c] = 1;
    }
    /* film/epit: upper half of z does not count (read_defects_u.F:52-54) */
    if (p->bulk_or_film == 1 || p->bulk_or_film == 2) {
        for (int z = (p->Lz + 1) / 2; z < p->Lz; ++z)
            for (int c = 0; c < p->Lx * p->Ly; ++c) o->sum_p[c + (size_t)p->Lx * p->Ly * z] = 0;
    }
    /* dead layers iz = 0 (gap_id 1, 2) and Lz/2 (gap_id 2): fixed u, 1/m = 0,
     * not counted (read_defects_u.F:56-69) */
    for (int g = 0; g < p->gap_id; ++g) {
        const int z = g == 0 ? 0 : p->Lz / 2;
        for (int c = 0; c < p->Lx * p->Ly; ++c) {
            const size_t cc = c + (size_t)p->Lx * p->Ly * z;
            for (int a = 0; a < 3; ++a) o->dipoR[3 * cc + a] = p->gap_dipole_u[a];
            o->mass_dipo_inv[cc] = 0.0;
            o->sum_p[cc] = 0;
        }
    }
    fo_recount_sum_p(o);
    /* E_field from a static external field (initialize_E_field.F:58-66),
     * mirrored for film/epit (:68-71) */
    for (size_t c = 0; c < N; ++c)
        for (int a = 0; a < 3; ++a) o->E_field[3 * c + a] = p->external_E_field[a];
    if (p->bulk_or_film == 1 || p->bulk_or_film == 2) mirror(o, o->E_field);
    o->nose[0] = 1.0; o->nose[1] = 0.0; o->nose[2] = 0.0;
    return o;
}

void fo_recount_sum_p(fo_ctx *o)
{
    int n = 0;
    for (int c = 0; c < o->N; ++c) n += o->sum_p[c] ? 1 : 0;
    o->n_sum_p = n;
    const int LxLy = o->p.Lx * o->p.Ly;                     /* read_defects_u.F:73-75 */
    for (int iz = 0; iz < o->p.Lz; ++iz) {
        int m = 0;
        for (int c = 0; c < LxLy; ++c) m += o->sum_p[iz * LxLy + c] ? 1 : 0;
        o->n_sum_layer[iz] = m;
    }
}

void fo_destroy(fo_ctx *o)
{
    if (!o) return;
    double *ptrs[] = {o->dipoR, o->dipoP, o->dVddi, o->dVddi_long, o->acouR, o->acouP,
                      o->dVdacR, o->E_field, o->Z_u_R, o->didiR, o->inhoR, o->Z_starR,
                      o->mass_dipo_inv, o->mass_acou_inv, o->dipoK, o->Z_u_K, o->dipoT,
                      o->didiK, o->acouK, o->inhoK, o->dVdacK, o->modulation_k,
                      o->c_k_modulation_k, o->long_range, o->short_range, o->inhomo_K,
                      o->inhomo_K_G, o->inhomo_coup2_K};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    free(o->e_layer_dipo); free(o->e_layer_acou); free(o->n_sum_layer); free(o->planes);
    free(o->relaxor); free(o->sum_p); free(o);
}

double fo_kappa(const fo_ctx *o) { return o->kappa; }
double fo_pNa03(const fo_ctx *o, int i) { return o->pNa03[i]; }
double fo_target_kinetic_energy(const fo_ctx *o) { return o->target_kinetic_energy; }
double fo_temperature_eV(const fo_ctx *o) { return o->temperature_eV; }

double *fo_field(fo_ctx *o, const char *n)
{
#define F(name) if (!strcmp(n, #name)) return o->name;
    F(dipoR) F(dipoP) F(dVddi) F(acouR) F(acouP) F(dVdacR) F(E_field) F(Z_starR)
    F(mass_dipo_inv) F(mass_acou_inv) F(long_range) F(short_range) F(inhomo_K)
    F(inhomo_K_G) F(inhomo_coup2_K) F(acouK) F(didiK) F(dipoK) F(c_k_modulation_k)
    F(inhoR) F(dVddi_long) F(didiR) F(modulation_k) F(Z_u_R)
#undef F
    return NULL;
}
int *fo_relaxor(fo_ctx *o) { return o->relaxor; }
unsigned char *fo_sum_p(fo_ctx *o) { return o->sum_p; }
double *fo_strain(fo_ctx *o) { return o->strain; }
double *fo_homogeneous(fo_ctx *o) { return o->homogeneous; }
double *fo_nose(fo_ctx *o) { return o->nose; }

/* ------------------------------------------------------------------------- */
/* Set-up tables                                                              */
/* ------------------------------------------------------------------------- */
/* decay_function_module.f:13-33 */
double fo_B_decay(double r, double rr, double kappa)
{
    return erfc(kappa * r) / (r * rr) + M_2_SQRTPI_D * kappa * exp(-(kappa * kappa) * rr) / rr;
}
double fo_C_decay(double r, double rr, double kappa)
{
    double Crr = 3 * erfc(kappa * r) / (r * rr)
               + M_2_SQRTPI_D * kappa * exp(-(kappa * kappa) * rr) * (2 * (kappa * kappa) + 3 / rr);
    return Crr / rr;
}

/* fft_3d_3x3upper_r2c.F:9-36: FFT the 6 upper-triangle components of a real
 * (x,y,z,3,3) array; the lower triangle is the conjugate.  The callers keep
 * only the real part, so the table stored here is Re of the transform. */
static void fft_3x3upper_real(fo_ctx *o, double *TMPreal /* N*9, (c, a, b) */,
                              double *table /* 9 per k */)
{
    size_t N = o->N, Kh = o->Kh;
    double *spec = zalloc(2 * Kh);
    for (int b = 0; b < 3; ++b)
        for (int a = 0; a <= b; ++a) {
            fo_r2c(o, TMPreal + N * (a + 3 * b), 1, 0, spec);
            for (size_t k = 0; k < Kh; ++k) {
                table[9 * k + a + 3 * b] = spec[2 * k];
                table[9 * k + b + 3 * a] = spec[2 * k];   /* dble(conjg(.)) */
            }
        }
    free(spec);
}

/* dipole_dipole_long_range (dipole-dipole-long-range.F:14-152) */
static void dipole_dipole_long_range(fo_ctx *o)
{
    const fo_params *p = &o->p;
    const int Lx = p->Lx, Ly = p->Ly, Lz = p->Lz, Kx = o->Kx;
    const size_t N = o->N;
    const double kappa = o->kappa;
    double four_pi_over_unit_cell_volume = 4 * M_PI_D / (p->a0 * p->a0 * p->a0);
    double two_pi_over_a = 2 * M_PI_D / p->a0;
    double four_kappa2 = 4 * (kappa * kappa);
    double Lx_a0 = Lx * p->a0, Ly_a0 = Ly * p->a0, Lz_a0 = Lz * p->a0;
    int max_x = 24 / Lx, max_y = 24 / Ly, max_z = 24 / Lz;
    double *T = zalloc(9 * N);   /* TMPreal(ix,iy,iz,a,b) -> T[c + N*(a+3b)] */

    for (int a = 0; a < 3; ++a)
        T[0 + N * (a + 3 * a)] = -2 * M_2_SQRTPI_D * (kappa * kappa * kappa) / 3;

    /* real-space sum: image order jz, jy, jx (outer) as the reference, so each
     * cell accumulates its images in the same order; cells are independent. */
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Lz; ++iz) {
        for (int jz = -1 - max_z; jz <= max_z; ++jz) {
            double n3 = jz * Lz_a0;
            for (int jy = -1 - max_y; jy <= max_y; ++jy) {
                double n2 = jy * Ly_a0;
                for (int jx = -1 - max_x; jx <= max_x; ++jx) {
                    double n1 = jx * Lx_a0;
                    for (int b = 0; b < 3; ++b)
                        for (int a = 0; a < 3; ++a) {
                            double r3 = iz * p->a0;
                            for (int iy = 0; iy < Ly; ++iy) {
                                double r2 = iy * p->a0;
                                for (int ix = 0; ix < Lx; ++ix) {
                                    if (jx == 0 && jy == 0 && jz == 0 && ix == 0 && iy == 0 && iz == 0)
                                        continue;
                                    double r1 = ix * p->a0;
                                    double rn[3] = {r1 + n1, r2 + n2, r3 + n3};
                                    double rnrn = rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2];
                                    double rn_norm = sqrt(rnrn);
                                    size_t c = ix + (size_t)Lx * (iy + (size_t)Ly * iz);
                                    double *t = &T[c + N * (a + 3 * b)];
                                    if (a == b) *t = *t + fo_B_decay(rn_norm, rnrn, kappa);
                                    *t = *t - rn[a] * rn[b] * fo_C_decay(rn_norm, rnrn, kappa);
                                }
                            }
                        }
                }
            }
        }
    }
    fft_3x3upper_real(o, T, o->long_range);
    free(T);

    /* reciprocal-space sum (dipole-dipole-long-range.F:107-144) */
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Lz; ++iz) {
        double k3 = (double)iz / Lz; if (k3 > 0.5) k3 = k3 - 1.0;
        for (int iy = 0; iy < Ly; ++iy) {
            double k2 = (double)iy / Ly; if (k2 > 0.5) k2 = k2 - 1.0;
            for (int ix = 0; ix <= Lx / 2; ++ix) {
                double k1 = (double)ix / Lx;
                size_t k = ix + (size_t)Kx * (iy + (size_t)Ly * iz);
                for (int b = 0; b < 3; ++b)
                    for (int a = 0; a < 3; ++a) {
                        double acc = o->long_range[9 * k + a + 3 * b];
                        for (int nz = -max_z; nz <= max_z; ++nz) {
                            double kG3 = two_pi_over_a * (k3 + nz);
                            for (int ny = -max_y; ny <= max_y; ++ny) {
                                double kG2 = two_pi_over_a * (k2 + ny);
                                for (int nx = -max_x; nx <= max_x; ++nx) {
                                    if (ix == 0 && iy == 0 && iz == 0 && nx == 0 && ny == 0 && nz == 0)
                                        continue;
                                    double kG1 = two_pi_over_a * (k1 + nx);
                                    double kG[3] = {kG1, kG2, kG3};
                                    double kGkG = kG1 * kG1 + kG2 * kG2 + kG3 * kG3;
                                    acc = acc + four_pi_over_unit_cell_volume * kG[a] * kG[b]
                                                  * exp(-kGkG / four_kappa2) / kGkG;
                                }
                            }
                        }
                        o->long_range[9 * k + a + 3 * b] = acc;
                    }
            }
        }
    }
    for (size_t i = 0; i < 9 * (size_t)o->Kh; ++i)
        o->long_range[i] = o->long_range[i] * (HARTREE_BOHR / p->epsilon_inf);
}

/* dipole_dipole_short_range (dipole-dipole-short-range.F:10-135) */
static void dipole_dipole_short_range(fo_ctx *o)
{
    const fo_params *p = &o->p;
    const int Lx = p->Lx, Ly = p->Ly, Lz = p->Lz;
    const size_t N = o->N;
    double *T = zalloc(9 * N);
    for (int a = 0; a < 3; ++a) T[0 + N * (a + 3 * a)] = p->P_kappa2[0] * 2;
    for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a)
            for (int iz = -1; iz <= 1; ++iz) {
                int i3 = iz, x3 = iz >= 0 ? iz : Lz + iz;
                for (int iy = -1; iy <= 1; ++iy) {
                    int i2 = iy, x2 = iy >= 0 ? iy : Ly + iy;
                    for (int ix = -1; ix <= 1; ++ix) {
                        int i1 = ix, x1 = ix >= 0 ? ix : Lx + ix;
                        int i[3] = {i1, i2, i3};
                        int ii = ix * ix + iy * iy + iz * iz;
                        size_t c = x1 + (size_t)Lx * (x2 + (size_t)Ly * x3);
                        double *t = &T[c + N * (a + 3 * b)];
                        if (ii == 1) {
                            if (a == b) *t = (i[a] == 0) ? p->j[0] : p->j[1];
                        } else if (ii == 2) {
                            if (a == b) *t = (i[a] == 0) ? p->j[3] : p->j[2];
                            else *t = p->j[4] * i[a] * i[b];
                        } else if (ii == 3) {
                            if (a == b) *t = p->j[5];
                            else *t = p->j[6] * i[a] * i[b];
                        }
                    }
                }
            }
    fft_3x3upper_real(o, T, o->short_range);
    free(T);
}

/* k vector (elastic.F:83-89, dVddi_inho_coupling.F:33-40) */
static inline void kvec(const fo_ctx *o, int ix, int iy, int iz, double k[3])
{
    const fo_params *p = &o->p;
    k[2] = (double)iz / p->Lz; if (k[2] > 0.5) k[2] = k[2] - 1.0;
    k[2] = k[2] * M_2PI_D / p->a0;
    k[1] = (double)iy / p->Ly; if (k[1] > 0.5) k[1] = k[1] - 1.0;
    k[1] = k[1] * M_2PI_D / p->a0;
    k[0] = (double)ix / p->Lx * M_2PI_D / p->a0;
}

/* elastic (elastic.F:7-237) */
static int elastic(fo_ctx *o, const double *modulation_R)
{
    const fo_params *p = &o->p;
    const int Lx = p->Lx, Ly = p->Ly, Lz = p->Lz, Kx = o->Kx;
    double *H = o->homogeneous, *G = o->homogeneous_G;
    memset(H, 0, sizeof(o->homogeneous));
    memset(G, 0, sizeof(o->homogeneous_G));
    for (int i = 0; i < 3; ++i) H6(H, i, i) = p->B11 * o->N;
    for (int i = 3; i < 6; ++i) H6(H, i, i) = p->B44 * o->N;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (i != j) H6(H, i, j) = p->B12 * o->N;
    for (int j = 0; j < 6; ++j) {
        double s = 0.0;
        for (int l = 0; l < j; ++l) s += H6(G, j, l) * H6(G, j, l);
        H6(G, j, j) = sqrt(H6(H, j, j) - s);
        o->homogeneous_G_inv[j] = 1.0 / H6(G, j, j);
        for (int i = j + 1; i < 6; ++i) {
            double s2 = 0.0;
            for (int l = 0; l < j; ++l) s2 += H6(G, i, l) * H6(G, j, l);
            H6(G, i, j) = (H6(H, i, j) - s2) * o->homogeneous_G_inv[j];
        }
    }
    double *C2 = o->homogeneous_coupling_2;
    memset(C2, 0, sizeof(o->homogeneous_coupling_2));
    for (int i = 0; i < 3; ++i) H6(C2, i, i) = p->B1xx / 2;
    for (int i = 3; i < 6; ++i) H6(C2, i, i) = p->B4yz;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (i != j) H6(C2, i, j) = p->B1yy / 2;

    static const int ki[3][3] = {{0, 3, 2}, {3, 0, 1}, {2, 1, 0}}; /* ki(a,m) = ki[m][a] */
    for (int iz = 0; iz < Lz; ++iz)
        for (int iy = 0; iy < Ly; ++iy)
            for (int ix = 0; ix <= Lx / 2; ++ix) {
                size_t kk = ix + (size_t)Kx * (iy + (size_t)Ly * iz);
                double k[3];
                kvec(o, ix, iy, iz, k);
                double *K = &o->inhomo_K[9 * kk];
                for (int b = 0; b < 3; ++b) {
                    int o1 = (b + 2) % 3, o2 = (b + 1) % 3; /* mod(b+1,3)+1, mod(b,3)+1 (1-based) */
                    for (int a = 0; a < 3; ++a) {
                        if (a == b)
                            K[a + 3 * b] = p->B11 * k[a] * k[a] + p->B44 * k[o1] * k[o1] + p->B44 * k[o2] * k[o2];
                        else
                            K[a + 3 * b] = p->B12 * k[a] * k[b] + p->B44 * k[a] * k[b];
                    }
                }
                double *Gk = &o->inhomo_K_G[12 * kk];
                for (int i = 0; i < 12; ++i) Gk[i] = 0.0;
                if (ix != 0 || iy != 0 || iz != 0) {
                    for (int b = 0; b < 3; ++b) {
                        double s = 0.0;
                        for (int l = 0; l < b; ++l) s += Gk[b + 3 * l] * Gk[b + 3 * l];
                        Gk[b + 3 * b] = sqrt(K[b + 3 * b] - s);
                        Gk[b + 3 * 3] = 1.0 / Gk[b + 3 * b];
                        for (int a = b + 1; a < 3; ++a) {
                            double s2 = 0.0;
                            for (int l = 0; l < b; ++l) s2 += Gk[a + 3 * l] * Gk[b + 3 * l];
                            Gk[a + 3 * b] = (K[a + 3 * b] - s2) * Gk[b + 3 * 3];
                        }
                    }
                }
                double *Ck = &o->inhomo_coup2_K[18 * kk];
                for (int i = 0; i < 6; ++i)
                    for (int a = 0; a < 3; ++a) {
                        double v;
                        if (i < 3) v = (i == a) ? k[a] * p->B1xx / 2 : k[a] * p->B1yy / 2;
                        else {
                            int q = ki[i - 3][a];
                            v = (q == 0 ? 0.0 : k[q - 1]) * p->B4yz;
                        }
                        Ck[a + 3 * i] = v;
                    }
            }

    /* c_k_modulation_k (elastic.F:187-234) */
    memset(o->c_k_modulation_k, 0, sizeof(double) * 6 * o->Kh);
    if (modulation_R) {
        if (p->modulation_constant == 0.0) return 2;
        fo_r2c(o, modulation_R, 1, 0, o->modulation_k);
        for (int iz = 0; iz < Lz; ++iz)
            for (int iy = 0; iy < Ly; ++iy)
                for (int ix = 0; ix <= Lx / 2; ++ix) {
                    size_t kk = ix + (size_t)Kx * (iy + (size_t)Ly * iz);
                    double k[3];
                    kvec(o, ix, iy, iz, k);
                    double mr = o->modulation_k[2 * kk], mi = o->modulation_k[2 * kk + 1];
                    for (int a = 0; a < 3; ++a) {
                        double f = p->modulation_constant * k[a];
                        o->c_k_modulation_k[2 * (3 * kk + a)] = f * mr;
                        o->c_k_modulation_k[2 * (3 * kk + a) + 1] = f * mi;
                    }
                }
    } else if (p->modulation_constant != 0.0) {
        return 3;
    }
    return 0;
}

int fo_setup_tables(fo_ctx *o, const double *modulation_R)
{
    dipole_dipole_long_range(o);
    dipole_dipole_short_range(o);
    return elastic(o, modulation_R);
}

/* ------------------------------------------------------------------------- */
/* Force evaluation                                                           */
/* ------------------------------------------------------------------------- */
/* mirror.F:9-36 */
static void mirror(fo_ctx *o, double *f)
{
    const fo_params *p = &o->p;
    const size_t plane = (size_t)p->Lx * p->Ly;
    if (p->gap_id == 0) {
        for (int iz = 0; iz < p->Lz / 2; ++iz)
            for (size_t c = 0; c < plane; ++c) {
                double *d = &f[3 * (c + plane * (p->Lz - 1 - iz))];
                const double *s = &f[3 * (c + plane * iz)];
                d[0] = -s[0]; d[1] = -s[1]; d[2] = s[2];
            }
    } else {
        for (int iz = 1; iz <= (p->Lz + 1) / 2 - 1; ++iz)
            for (size_t c = 0; c < plane; ++c) {
                double *d = &f[3 * (c + plane * (p->Lz - iz))];
                const double *s = &f[3 * (c + plane * iz)];
                d[0] = -s[0]; d[1] = -s[1]; d[2] = s[2];
            }
    }
}

/* dipoFFT (dipoFFT.F:9-47) */
static void dipoFFT(fo_ctx *o)
{
    const size_t N = o->N;
    for (int a = 0; a < 3; ++a) fo_r2c(o, o->dipoR, 3, a, o->dipoK);
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        const double *u = &o->dipoR[3 * c];
        double *d = &o->didiR[6 * c];
        d[0] = u[0] * u[0]; d[1] = u[1] * u[1]; d[2] = u[2] * u[2];
        d[3] = u[1] * u[2]; d[4] = u[2] * u[0]; d[5] = u[0] * u[1];
    }
    for (int j = 0; j < 6; ++j) fo_r2c(o, o->didiR, 6, j, o->didiK);
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c)
        for (int a = 0; a < 3; ++a) o->Z_u_R[3 * c + a] = o->Z_starR[c] * o->dipoR[3 * c + a];
    for (int a = 0; a < 3; ++a) fo_r2c(o, o->Z_u_R, 3, a, o->Z_u_K);
}

/* acouFFT (acouFFT.F:9-26): acouK = i * FFT(acouR) */
static void acouFFT(fo_ctx *o)
{
    for (int a = 0; a < 3; ++a) fo_r2c(o, o->acouR, 3, a, o->acouK);
    for (size_t i = 0; i < 3 * (size_t)o->Kh; ++i) {
        double re = o->acouK[2 * i], im = o->acouK[2 * i + 1];
        o->acouK[2 * i] = -im;
        o->acouK[2 * i + 1] = re;
    }
}

/* dVddi_dipole_dipole (dVddi-dipole-dipole.F:8-103) */
static void dVddi_dipole_dipole(fo_ctx *o, fo_energy *e)
{
    const size_t N = o->N, Kh = o->Kh;
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < Kh; ++k)
        for (int a = 0; a < 3; ++a) {
            double re = 0.0, im = 0.0;
            for (int b = 0; b < 3; ++b) {
                double s = o->short_range[9 * k + a + 3 * b];
                re = re + s * o->dipoK[2 * (3 * k + b)];
                im = im + s * o->dipoK[2 * (3 * k + b) + 1];
            }
            o->dipoT[2 * (3 * k + a)] = re; o->dipoT[2 * (3 * k + a) + 1] = im;
        }
    for (int a = 0; a < 3; ++a) fo_c2r(o, o->dipoT, 3, a, o->dVddi);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < 3 * N; ++i) o->dVddi[i] = o->dVddi[i] * o->Ninv;

#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < Kh; ++k)
        for (int a = 0; a < 3; ++a) {
            double re = 0.0, im = 0.0;
            for (int b = 0; b < 3; ++b) {
                double s = o->long_range[9 * k + a + 3 * b];
                re = re + s * o->Z_u_K[2 * (3 * k + b)];
                im = im + s * o->Z_u_K[2 * (3 * k + b) + 1];
            }
            o->dipoT[2 * (3 * k + a)] = re; o->dipoT[2 * (3 * k + a) + 1] = im;
        }
    for (int a = 0; a < 3; ++a) fo_c2r(o, o->dipoT, 3, a, o->dVddi_long);
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c)
        for (int a = 0; a < 3; ++a)
            o->dVddi_long[3 * c + a] = o->Z_starR[c] * o->dVddi_long[3 * c + a] * o->Ninv;

    const int Lz = o->p.Lz;
    const size_t NP = N / Lz;
    double *esr = plane_buf(o, 0), *elr = plane_buf(o, 1);
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Lz; ++iz) {
        double ts = 0.0, tl = 0.0;
        for (size_t c = iz * NP; c < (iz + 1) * NP; ++c) {
            const double *u = &o->dipoR[3 * c], *f = &o->dVddi[3 * c], *g = &o->dVddi_long[3 * c];
            ts = ts + (u[0] * f[0] + u[1] * f[1] + u[2] * f[2]);
            tl = tl + (u[0] * g[0] + u[1] * g[1] + u[2] * g[2]);
        }
        esr[iz] = ts;
        elr[iz] = tl;
    }
    e->short_range = plane_sum(esr, Lz) / 2;
    e->long_range = plane_sum(elr, Lz) / 2;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < 3 * N; ++i) o->dVddi[i] = o->dVddi[i] + o->dVddi_long[i];
}

/* dVddi_unharmonic (dVddi-unharmonic.F:6-101) */
static void dVddi_unharmonic(fo_ctx *o, fo_energy *e)
{
    const fo_params *p = &o->p;
    const size_t N = o->N;
    double fa2g[3], tg[3], sk[3], tk2[3], fk[3];
    for (int r = 0; r < 3; ++r) {
        fa2g[r] = 4 * p->P_alpha[r] + 2 * p->P_gamma[r];
        tg[r] = 2 * p->P_gamma[r];
        sk[r] = 6 * p->P_k1[r] + 2 * p->P_k2[r];
        tk2[r] = 2 * p->P_k2[r];
        fk[r] = 4 * p->P_k2[r] - 2 * p->P_k3[r];
    }
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        int r = o->relaxor[c] - 1;
        const double *u = &o->dipoR[3 * c], *d = &o->didiR[6 * c];
        double *f = &o->dVddi[3 * c];
        for (int a = 0; a < 3; ++a) {
            int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
            f[a] = f[a] + u[a] * 2 * (p->P_kappa2[r] - p->P_kappa2[0]);
            f[a] = f[a] + u[a] * (fa2g[r] * (d[0] + d[1] + d[2]) - tg[r] * d[a]);
            double s = d[0] + d[1] + d[2];
            f[a] = f[a] + u[a] * (sk[r] * (s * s) + 8 * p->P_k4[r] * (s * s * s)
                                  - tk2[r] * (d[a] * d[a]) - fk[r] * d[a1] * d[a2]);
        }
    }
    const int Lz = p->Lz;
    const size_t NP = N / Lz;
    double *part = plane_buf(o, 0);
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Lz; ++iz) {
      double unharm0 = 0.0;
      for (size_t c = iz * NP; c < (iz + 1) * NP; ++c) {
        int r = o->relaxor[c] - 1;
        const double *d = &o->didiR[6 * c];
        double s = d[0] + d[1] + d[2];
        unharm0 = unharm0
            + (p->P_kappa2[r] - p->P_kappa2[0]) * (d[0] + d[1] + d[2])
            + p->P_alpha[r] * (s * s)
            + p->P_gamma[r] * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])
            + p->P_k1[r] * (s * s * s)
            + p->P_k2[r] * ((d[0] * d[0]) * (d[1] + d[2]) + (d[1] * d[1]) * (d[2] + d[0])
                            + (d[2] * d[2]) * (d[0] + d[1]))
            + p->P_k3[r] * d[0] * d[1] * d[2]
            + p->P_k4[r] * ((s * s) * (s * s));
      }
      part[iz] = unharm0;
    }
    e->unharmonic = plane_sum(part, Lz);
}

/* optimize_homo_strain (optimize-homo-strain.F:8-89) */
static void optimize_homo_strain(fo_ctx *o, fo_energy *e)
{
    const fo_params *p = &o->p;
    double c2uu[6];
    const double *C2 = o->homogeneous_coupling_2, *H = o->homogeneous, *G = o->homogeneous_G;
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        for (int j = 0; j < 6; ++j) s = s + H6(C2, i, j) * o->didiK[2 * j];   /* dble(didiK(:,0,0,0)) */
        c2uu[i] = s;
    }
    for (int i = 0; i < 3; ++i) c2uu[i] = c2uu[i] + o->pNa03[i];
    double *st = o->strain;
    if (p->bulk_or_film == 2) {          /* epit */
        st[0] = p->epi_strain;
        st[1] = p->epi_strain;
        st[2] = (-(H6(H, 2, 0) + H6(H, 2, 1)) * p->epi_strain - c2uu[2]) / H6(H, 2, 2);
        st[3] = -c2uu[3] / H6(H, 3, 3);
        st[4] = -c2uu[4] / H6(H, 4, 4);
        st[5] = 0.0;
    } else if (p->bulk_or_film == 3) {   /* e110 */
        st[2] = p->epi_strain;
        st[3] = 0.0; st[4] = 0.0;
        double numerator = -(H6(H, 0, 0) + H6(H, 1, 1) + 2 * H6(H, 1, 1)) * p->epi_strain
                           - (H6(H, 2, 0) + H6(H, 2, 1)) * st[2] - c2uu[0] - c2uu[1] + 2 * c2uu[5];
        double denominator = H6(H, 0, 0) + H6(H, 1, 1) + 2 * H6(H, 1, 0) + 4 * H6(H, 5, 5);
        double delta = numerator / denominator;
        st[0] = p->epi_strain + delta;
        st[1] = p->epi_strain + delta;
        st[5] = -2 * delta;
    } else {
        for (int i = 0; i < 6; ++i) st[i] = -c2uu[i];
        for (int i = 0; i < 6; ++i) {
            double tmp = 0.0;
            for (int j = 0; j < i; ++j) tmp = tmp + H6(G, i, j) * st[j];
            st[i] = (st[i] - tmp) * o->homogeneous_G_inv[i];
        }
        for (int i = 5; i >= 0; --i) {
            double tmp = 0.0;
            for (int j = i + 1; j < 6; ++j) tmp = tmp + st[j] * H6(G, j, i);
            st[i] = (st[i] - tmp) * o->homogeneous_G_inv[i];
        }
    }
    double hs = 0.0;
    for (int i = 0; i < 6; ++i) {
        double d = 0.0;
        for (int j = 0; j < 6; ++j) d = d + H6(H, i, j) * st[j];
        hs = hs + st[i] * d;
    }
    e->homo_strain = hs / 2;
    double hc = 0.0;
    for (int i = 0; i < 6; ++i) hc = hc + st[i] * c2uu[i];
    e->homo_coupling = hc;
}

/* e(j) = sum_i eta_i hc2(i,j); F += coupling (dVddi-homo-coupling.F:19-47) */
static inline void add_coupling_force(double *f, const double *u, const double *e)
{
    f[0] = f[0] + 2 * e[0] * u[0] + e[4] * u[2] + e[5] * u[1];
    f[1] = f[1] + 2 * e[1] * u[1] + e[3] * u[2] + e[5] * u[0];
    f[2] = f[2] + 2 * e[2] * u[2] + e[3] * u[1] + e[4] * u[0];
}

static void dVddi_homo_coupling(fo_ctx *o)
{
    double e[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < 6; ++j)
        for (int i = 0; i < 6; ++i)
            e[j] = e[j] + o->strain[i] * H6(o->homogeneous_coupling_2, i, j);
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < (size_t)o->N; ++c)
        add_coupling_force(&o->dVddi[3 * c], &o->dipoR[3 * c], e);
}

/* optimize_inho_strain (optimize-inho-strain.F:10-60) */
static void optimize_inho_strain(fo_ctx *o)
{
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < (size_t)o->Kh; ++k) {
        const double *C = &o->inhomo_coup2_K[18 * k], *G = &o->inhomo_K_G[12 * k];
        const double *dd = &o->didiK[12 * k], *cm = &o->c_k_modulation_k[6 * k];
        double tr[3], ti[3];
        for (int a = 0; a < 3; ++a) {
            double dr = 0.0, di = 0.0;
            for (int i = 0; i < 6; ++i) {
                dr = dr + C[a + 3 * i] * dd[2 * i];
                di = di + C[a + 3 * i] * dd[2 * i + 1];
            }
            tr[a] = -cm[2 * a] - dr;
            ti[a] = -cm[2 * a + 1] - di;
        }
        for (int a = 0; a < 3; ++a) {
            double sr = 0.0, si = 0.0;
            for (int b = 0; b < a; ++b) { sr = sr + G[a + 3 * b] * tr[b]; si = si + G[a + 3 * b] * ti[b]; }
            tr[a] = (tr[a] - sr) * G[a + 9];
            ti[a] = (ti[a] - si) * G[a + 9];
        }
        for (int a = 2; a >= 0; --a) {
            double sr = 0.0, si = 0.0;
            for (int b = a + 1; b < 3; ++b) { sr = sr + tr[b] * G[b + 3 * a]; si = si + ti[b] * G[b + 3 * a]; }
            tr[a] = (tr[a] - sr) * G[a + 9];
            ti[a] = (ti[a] - si) * G[a + 9];
        }
        for (int a = 0; a < 3; ++a) {
            o->acouK[2 * (3 * k + a)] = tr[a];
            o->acouK[2 * (3 * k + a) + 1] = ti[a];
        }
    }
}

/* dVdac (dVdac.F:11-63) */
static void dVdac(fo_ctx *o)
{
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < (size_t)o->Kh; ++k) {
        const double *C = &o->inhomo_coup2_K[18 * k], *K = &o->inhomo_K[9 * k];
        const double *dd = &o->didiK[12 * k], *cm = &o->c_k_modulation_k[6 * k];
        const double *w = &o->acouK[6 * k];
        for (int a = 0; a < 3; ++a) {
            double m1r = 0.0, m1i = 0.0, m2r = 0.0, m2i = 0.0;
            for (int i = 0; i < 6; ++i) { m1r = m1r + C[a + 3 * i] * dd[2 * i]; m1i = m1i + C[a + 3 * i] * dd[2 * i + 1]; }
            for (int b = 0; b < 3; ++b) { m2r = m2r + K[a + 3 * b] * w[2 * b]; m2i = m2i + K[a + 3 * b] * w[2 * b + 1]; }
            double vr = ((0.0 + cm[2 * a]) + m1r) + m2r;
            double vi = ((0.0 + cm[2 * a + 1]) + m1i) + m2i;
            /* times (0,-1) */
            o->dVdacK[2 * (3 * k + a)] = vi;
            o->dVdacK[2 * (3 * k + a) + 1] = -vr;
        }
    }
    for (int a = 0; a < 3; ++a) fo_c2r(o, o->dVdacK, 3, a, o->dVdacR);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < 3 * (size_t)o->N; ++i) o->dVdacR[i] = o->dVdacR[i] * o->Ninv;
}

/* Half-spectrum Parseval weights (energy_module.F:162-177): kx=0 and kx=Lx/2
 * count 1/2, interior 1; the kx loop of the reference sums the interior first. */
static void inhomogeneous_energies(fo_ctx *o, fo_energy *e)
{
    const fo_params *p = &o->p;
    const int Lx = p->Lx, Ly = p->Ly, Lz = p->Lz, Kx = o->Kx;
    double *pS = plane_buf(o, 0), *pC = plane_buf(o, 1), *pM = plane_buf(o, 2);
    const int modulation = p->modulation_constant != 0.0;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Lz; ++iz) {
        const int xs[3] = {0, 1, Lx / 2};
        double tS = 0.0, tC = 0.0, tM = 0.0;
        for (int iy = 0; iy < Ly; ++iy)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    double t[3];
                    for (int part = 0; part < 3; ++part) {
                        int x0 = part == 1 ? 1 : xs[part], x1 = part == 1 ? Lx / 2 - 1 : xs[part];
                        double s = 0.0;
                        for (int ix = x0; ix <= x1; ++ix) {
                            size_t k = ix + (size_t)Kx * (iy + (size_t)Ly * iz);
                            const double *w = &o->acouK[6 * k];
                            double K = o->inhomo_K[9 * k + a + 3 * b];
                            /* dble(conjg(wa) * K * wb) */
                            double cr = w[2 * a] * K, ci = -w[2 * a + 1] * K;
                            s = s + (cr * w[2 * b] - ci * w[2 * b + 1]);
                        }
                        t[part] = s;
                    }
                    tS = tS + t[0] / 2 + t[1] + t[2] / 2;
                }
        for (int iy = 0; iy < Ly; ++iy)
            for (int i = 0; i < 6; ++i)
                for (int a = 0; a < 3; ++a) {
                    double t[3];
                    for (int part = 0; part < 3; ++part) {
                        int x0 = part == 1 ? 1 : xs[part], x1 = part == 1 ? Lx / 2 - 1 : xs[part];
                        double s = 0.0;
                        for (int ix = x0; ix <= x1; ++ix) {
                            size_t k = ix + (size_t)Kx * (iy + (size_t)Ly * iz);
                            const double *w = &o->acouK[6 * k], *dd = &o->didiK[12 * k];
                            double C = o->inhomo_coup2_K[18 * k + a + 3 * i];
                            double cr = w[2 * a] * C, ci = -w[2 * a + 1] * C;
                            s = s + (cr * dd[2 * i] - ci * dd[2 * i + 1]);
                        }
                        t[part] = s;
                    }
                    tC = tC + t[0] / 2 + t[1] + t[2] / 2;
                }
        if (modulation)
            for (int iy = 0; iy < Ly; ++iy)
                for (int a = 0; a < 3; ++a) {
                    double t[3];
                    for (int part = 0; part < 3; ++part) {
                        int x0 = part == 1 ? 1 : xs[part], x1 = part == 1 ? Lx / 2 - 1 : xs[part];
                        double s = 0.0;
                        for (int ix = x0; ix <= x1; ++ix) {
                            size_t k = ix + (size_t)Kx * (iy + (size_t)Ly * iz);
                            const double *w = &o->acouK[6 * k], *cm = &o->c_k_modulation_k[6 * k];
                            s = s + (w[2 * a] * cm[2 * a] + w[2 * a + 1] * cm[2 * a + 1]);
                        }
                        t[part] = s;
                    }
                    tM = tM + t[0] / 2 + t[1] + t[2] / 2;
                }
        pS[iz] = tS;
        pC[iz] = tC;
        pM[iz] = tM;
    }
    e->inho_strain = plane_sum(pS, Lz) * o->Ninv;
    e->inho_coupling = plane_sum(pC, Lz) * 2 * o->Ninv;
    e->inho_modulation = modulation ? plane_sum(pM, Lz) * 2 * o->Ninv : 0.0;
}

/* dVddi_inho_coupling (dVddi_inho_coupling.F:9-84) */
static void dVddi_inho_coupling(fo_ctx *o)
{
    const fo_params *p = &o->p;
    const int Lx = p->Lx, Ly = p->Ly, Lz = p->Lz, Kx = o->Kx;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Lz; ++iz)
        for (int iy = 0; iy < Ly; ++iy)
            for (int ix = 0; ix <= Lx / 2; ++ix) {
                size_t k = ix + (size_t)Kx * (iy + (size_t)Ly * iz);
                double kv[3];
                kvec(o, ix, iy, iz, kv);
                const double *w = &o->acouK[6 * k];
                double *h = &o->inhoK[12 * k];
                for (int ri = 0; ri < 2; ++ri) {
                    h[0 + ri] = kv[0] * w[0 + ri];
                    h[2 + ri] = kv[1] * w[2 + ri];
                    h[4 + ri] = kv[2] * w[4 + ri];
                    h[6 + ri] = kv[2] * w[2 + ri] + kv[1] * w[4 + ri];
                    h[8 + ri] = kv[0] * w[4 + ri] + kv[2] * w[0 + ri];
                    h[10 + ri] = kv[1] * w[0 + ri] + kv[0] * w[2 + ri];
                }
            }
    for (int j = 0; j < 6; ++j) fo_c2r(o, o->inhoK, 6, j, o->inhoR);
    const double *C2 = o->homogeneous_coupling_2;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < (size_t)o->N; ++c) {
        const double *h = &o->inhoR[6 * c];
        double e[6];
        for (int i = 0; i < 6; ++i)
            e[i] = h[0] * o->Ninv * H6(C2, 0, i) + h[1] * o->Ninv * H6(C2, 1, i)
                 + h[2] * o->Ninv * H6(C2, 2, i) + h[3] * o->Ninv * H6(C2, 3, i)
                 + h[4] * o->Ninv * H6(C2, 4, i) + h[5] * o->Ninv * H6(C2, 5, i);
        add_coupling_force(&o->dVddi[3 * c], &o->dipoR[3 * c], e);
    }
}

/* mirror_force_mirror (mirror_force_mirror.F:6-89) */
void fo_mirror_force_mirror(fo_ctx *o, fo_energy *e)
{
    const fo_params *p = &o->p;
    const int film = (p->bulk_or_film == 1 || p->bulk_or_film == 2);
    if (film) mirror(o, o->dipoR);
    dipoFFT(o);
    if (p->mass_acou[0] > 0.0) acouFFT(o);
    dVddi_dipole_dipole(o, e);
    dVddi_unharmonic(o, e);
    if (p->E_field_always_zero) {
        e->dipole_E_field = 0.0;
    } else {
#pragma omp parallel for schedule(static)
        for (size_t c = 0; c < (size_t)o->N; ++c)
            for (int a = 0; a < 3; ++a)
                o->dVddi[3 * c + a] = o->dVddi[3 * c + a] - o->Z_starR[c] * o->E_field[3 * c + a];
        const int Lz = p->Lz;
        const size_t NP = (size_t)o->N / Lz;
        double *part = plane_buf(o, 0);
#pragma omp parallel for schedule(static)
        for (int iz = 0; iz < Lz; ++iz) {
            double t = 0.0;
            for (size_t c = iz * NP; c < (iz + 1) * NP; ++c) {
                const double *u = &o->dipoR[3 * c], *E = &o->E_field[3 * c];
                t = t + o->Z_starR[c] * (u[0] * E[0] + u[1] * E[1] + u[2] * E[2]);
            }
            part[iz] = t;
        }
        e->dipole_E_field = -plane_sum(part, Lz);
    }
    optimize_homo_strain(o, e);
    dVddi_homo_coupling(o);
    if (p->mass_acou[0] > 0.0) dVdac(o);
    else optimize_inho_strain(o);
    inhomogeneous_energies(o, e);
    dVddi_inho_coupling(o);
    if (film) mirror(o, o->dVddi);
}

/* ------------------------------------------------------------------------- */
/* Integrators                                                                */
/* ------------------------------------------------------------------------- */
double fo_kinetic_energy_dipo(fo_ctx *o)
{
    const int Lz = o->p.Lz;
    const size_t NP = (size_t)o->N / Lz;
    double *part = plane_buf(o, 2);
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Lz; ++iz) {
        double t = 0.0;
        for (size_t c = iz * NP; c < (iz + 1) * NP; ++c)
            for (int a = 0; a < 3; ++a)
                if (o->sum_p[c]) t = t + o->dipoP[3 * c + a] * o->dipoP[3 * c + a] * o->mass_dipo_inv[c];
        part[iz] = t;
    }
    return plane_sum(part, Lz) / o->n_sum_p * o->N / 2;
}

double fo_kinetic_energy_acou(fo_ctx *o)
{
    const int Lz = o->p.Lz;
    const size_t NP = (size_t)o->N / Lz;
    double *part = plane_buf(o, 2);
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Lz; ++iz) {
        double t = 0.0;
        for (size_t c = iz * NP; c < (iz + 1) * NP; ++c)
            for (int a = 0; a < 3; ++a)
                t = t + o->acouP[3 * c + a] * o->acouP[3 * c + a] * o->mass_acou_inv[c];
        part[iz] = t;
    }
    return plane_sum(part, Lz) / 2;
}

double fo_sum_potential_energies(const fo_energy *e)
{
    return e->short_range + e->long_range + e->dipole_E_field + e->unharmonic + e->homo_strain
         + e->homo_coupling + e->inho_strain + e->inho_coupling + e->inho_modulation;
}

void fo_nose_poincare_hamiltonian(fo_ctx *o, fo_energy *e)
{
    double s = o->nose[0], pi = o->nose[1], H0 = o->nose[2];
    double H = e->dipo_kinetic / (s * s) + fo_sum_potential_energies(e)
             + pi * pi / 2 / o->p.Q_Nose + (double)(3 * o->N) * o->temperature_eV * log(s) - H0;
    e->H_Nose_Poincare = H * s;
}

/* nose_poincare (nose-poincare.F:10-78) */
void fo_nose_poincare(fo_ctx *o, fo_energy *e)
{
    const fo_params *p = &o->p;
    const size_t N = o->N;
    double *s = &o->nose[0], *pi = &o->nose[1];
    double dt2s = o->dt_2 * *s;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < 3 * N; ++i) o->dipoP[i] = o->dipoP[i] - dt2s * o->dVddi[i];
    e->dipo_kinetic = fo_kinetic_energy_dipo(o);
    double CB = (double)o->n_freedom * o->temperature_eV * (1.0 + log(*s))
              - e->dipo_kinetic / (*s * *s) + fo_sum_potential_energies(e) - o->nose[2];
    CB = o->dt_2 * CB - *pi;
    *pi = -2 * CB / (1.0 + sqrt(1.0 - CB * p->dt / p->Q_Nose));
    double s_old = *s;
    *s = *s * (1.0 + o->dt_2 * *pi / p->Q_Nose) / (1.0 - o->dt_2 * *pi / p->Q_Nose);
    double dt2_s_s = o->dt_2 * (1 / *s + 1 / s_old);
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c)
        for (int a = 0; a < 3; ++a)
            o->dipoR[3 * c + a] = o->dipoR[3 * c + a] + dt2_s_s * o->dipoP[3 * c + a] * o->mass_dipo_inv[c];
    fo_mirror_force_mirror(o, e);
    fo_nose_poincare_hamiltonian(o, e);
    double delta_H = e->H_Nose_Poincare / *s;
    *pi = *pi + o->dt_2 * (2 * e->dipo_kinetic / (*s * *s) - (double)o->n_freedom * o->temperature_eV)
        - o->dt_2 * delta_H;
    dt2s = o->dt_2 * *s;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < 3 * N; ++i) o->dipoP[i] = o->dipoP[i] - dt2s * o->dVddi[i];
    e->dipo_kinetic = fo_kinetic_energy_dipo(o);
    fo_nose_poincare_hamiltonian(o, e);
}

/* leapfrog (leapfrog.F:9-157); heat-source layers (delta_e) not restated. */
/* Heat-source and sink layers at Lz/4 and 3Lz/4 (leapfrog.F:38-73). */
static void heat_source_layers(fo_ctx *o)
{
    const fo_params *p = &o->p;
    const int Lx = p->Lx, Ly = p->Ly, Lz = p->Lz;
    const size_t LxLy = (size_t)Lx * Ly;
    const int acou = p->mass_acou[0] > 0.0;
    const double Lx_Ly_delta_e = (double)(Lx * Ly) * o->delta_e;
    const double Lx_Ly_delta_e_acou = (double)(Lx * Ly) * o->delta_e_acou;
    const int Lz_1_4 = Lz / 4, Lz_3_4 = Lz_1_4 * 3;
    if (o->delta_e == 0.0 && o->delta_e_acou == 0.0) return;
    for (int iz = 0; iz < Lz; ++iz) {                                   /* :44-58 */
        double el = 0.0;
        for (size_t c = iz * LxLy; c < (iz + 1) * LxLy; ++c)
            for (int a = 0; a < 3; ++a) el = el + o->dipoP[3 * c + a] * o->dipoP[3 * c + a] * o->mass_dipo_inv[c];
        o->e_layer_dipo[iz] = el / 2 * Lx * Ly / o->n_sum_layer[iz];
    }
    for (int iz = 0; iz < Lz; ++iz) {                                   /* :59-73 */
        double ea = 0.0;
        if (acou)
            for (size_t c = iz * LxLy; c < (iz + 1) * LxLy; ++c)
                for (int a = 0; a < 3; ++a) ea = ea + o->acouP[3 * c + a] * o->acouP[3 * c + a] * o->mass_acou_inv[c];
        o->e_layer_acou[iz] = ea / 2;
    }
    if (o->delta_e != 0.0) {                                            /* :75-78 */
        const double f1 = sqrt((o->e_layer_dipo[Lz_1_4] + Lx_Ly_delta_e) / o->e_layer_dipo[Lz_1_4]);
        const double f3 = sqrt((o->e_layer_dipo[Lz_3_4] - Lx_Ly_delta_e) / o->e_layer_dipo[Lz_3_4]);
        for (size_t c = 0; c < LxLy; ++c)
            for (int a = 0; a < 3; ++a) {
                o->dipoP[3 * (Lz_1_4 * LxLy + c) + a] *= f1;
                o->dipoP[3 * (Lz_3_4 * LxLy + c) + a] *= f3;
            }
    }
    if (o->delta_e_acou != 0.0) {                                       /* :80-83 */
        const double f1 = sqrt((o->e_layer_acou[Lz_1_4] + Lx_Ly_delta_e_acou) / o->e_layer_acou[Lz_1_4]);
        const double f3 = sqrt((o->e_layer_acou[Lz_3_4] - Lx_Ly_delta_e_acou) / o->e_layer_acou[Lz_3_4]);
        for (size_t c = 0; c < LxLy; ++c)
            for (int a = 0; a < 3; ++a) {
                o->acouP[3 * (Lz_1_4 * LxLy + c) + a] *= f1;
                o->acouP[3 * (Lz_3_4 * LxLy + c) + a] *= f3;
            }
    }
}

void fo_set_heat_source(fo_ctx *o, double delta_e, double delta_e_acou)
{
    o->delta_e = delta_e;
    o->delta_e_acou = delta_e_acou;
}

double *fo_layers(fo_ctx *o, int which) { return which == 0 ? o->e_layer_dipo : o->e_layer_acou; }

void fo_leapfrog(fo_ctx *o, fo_energy *e, int i_step)
{
    const fo_params *p = &o->p;
    const size_t N = o->N;
    const int acou = p->mass_acou[0] > 0.0;
    heat_source_layers(o);
    if (i_step == 0) {
        fo_mirror_force_mirror(o, e);
        for (size_t i = 0; i < 3 * N; ++i) o->dipoP[i] = o->dipoP[i] - o->dt_2 * o->dVddi[i];
        if (acou)
            for (size_t i = 0; i < 3 * N; ++i) o->acouP[i] = o->acouP[i] - o->dt_2 * o->dVdacR[i];
    } else {
        for (size_t c = 0; c < N; ++c)
            for (int a = 0; a < 3; ++a)
                o->dipoR[3 * c + a] = o->dipoR[3 * c + a] + p->dt * o->dipoP[3 * c + a] * o->mass_dipo_inv[c];
        if (acou)
            for (size_t c = 0; c < N; ++c)
                for (int a = 0; a < 3; ++a)
                    o->acouR[3 * c + a] = o->acouR[3 * c + a] + p->dt * o->acouP[3 * c + a] * o->mass_acou_inv[c];
        fo_mirror_force_mirror(o, e);
        for (size_t i = 0; i < 3 * N; ++i) o->dipoP[i] = o->dipoP[i] - p->dt * o->dVddi[i];
        if (acou)
            for (size_t i = 0; i < 3 * N; ++i) o->acouP[i] = o->acouP[i] - p->dt * o->dVdacR[i];
    }
    e->dipo_kinetic = fo_kinetic_energy_dipo(o);
    if (acou) e->acou_kinetic = fo_kinetic_energy_acou(o);
}

/* velocity_scaling (velocity_scaling.F:9-103) */
void fo_velocity_scaling(fo_ctx *o, fo_energy *e)
{
    const fo_params *p = &o->p;
    const size_t N = o->N;
    const int acou = p->mass_acou[0] > 0.0;
    for (size_t c = 0; c < N; ++c)
        for (int a = 0; a < 3; ++a)
            o->dipoR[3 * c + a] = o->dipoR[3 * c + a] + p->dt * o->dipoP[3 * c + a] * o->mass_dipo_inv[c];
    if (acou)
        for (size_t c = 0; c < N; ++c)
            for (int a = 0; a < 3; ++a)
                o->acouR[3 * c + a] = o->acouR[3 * c + a] + p->dt * o->acouP[3 * c + a] * o->mass_acou_inv[c];
    fo_mirror_force_mirror(o, e);
    for (size_t c = 0; c < N; ++c)
        for (int a = 0; a < 3; ++a)
            o->dipoP[3 * c + a] = o->sum_p[c] ? o->dipoP[3 * c + a] - p->dt * o->dVddi[3 * c + a] : 0.0;
    e->dipo_kinetic = fo_kinetic_energy_dipo(o);
    double scale = sqrt(o->target_kinetic_energy / e->dipo_kinetic);
    for (size_t i = 0; i < 3 * N; ++i) o->dipoP[i] = o->dipoP[i] * scale;
    if (acou) {
        for (size_t i = 0; i < 3 * N; ++i) o->acouP[i] = o->acouP[i] - p->dt * o->dVdacR[i];
        e->acou_kinetic = fo_kinetic_energy_acou(o);
        scale = sqrt(o->target_kinetic_energy / e->acou_kinetic);
        for (size_t i = 0; i < 3 * N; ++i) o->acouP[i] = o->acouP[i] * scale;
    }
}

/* initialize-dipoR.F:59-69 then molecular-dynamics.F:44-55 */
void fo_md_prologue(fo_ctx *o, fo_energy *e)
{
    memset(e, 0, sizeof(*e));
    fo_mirror_force_mirror(o, e);
    o->nose[0] = 1.0;
    o->nose[1] = 0.0;
    e->acou_kinetic = fo_kinetic_energy_acou(o);
    e->dipo_kinetic = fo_kinetic_energy_dipo(o);
    e->H_Nose_Poincare = 0.0;
    e->total = e->dipo_kinetic / (o->nose[0] * o->nose[0]) + e->acou_kinetic + fo_sum_potential_energies(e);
    o->nose[2] = 0.0;
    fo_nose_poincare_hamiltonian(o, e);
    o->nose[2] = e->H_Nose_Poincare / o->nose[0];
    fo_nose_poincare_hamiltonian(o, e);
}

/* ------------------------------------------------------------------------- */
/* Marsaglia-Tsang uni64 (marsaglia_tsang_uni64_module.f:41-130)              */
/* ------------------------------------------------------------------------- */
static struct {
    int i, j;
    double c, times;
    double u[97];
    int sw;
    double t, uu;
} rng = {0, 0, 0.0, 0.0, {0}, 1, 0.0, 0.0};

/* Default Fortran integers are 32 bits: 6969*x wraps around for large seeds
 * and mod() keeps the sign of the dividend, as C's % does. */
static inline int32_t mul_wrap32(int32_t a, int32_t b)
{
    return (int32_t)((uint32_t)a * (uint32_t)b);
}

static void fillu(double *u, int seed1, int seed2)
{
    int32_t x = seed1, y = seed2;
    for (int i = 0; i < 97; ++i) {
        double s = 0.0, t = 0.5;
        for (int j = 0; j < 53; ++j) {
            x = mul_wrap32(6969, x) % 65543;
            y = mul_wrap32(8888, y) % 65579;
            if (((x ^ y) & 32) > 0) s = s + t;
            t = 0.5 * t;
        }
        u[i] = s;
    }
}

double fo_uni64_seed(int seed1, int seed2)
{
    fillu(rng.u, seed1, seed2);
    rng.i = 97; rng.j = 33; rng.c = 0.0;
    rng.times = rng.times + 1.0;
    return rng.times;
}

double fo_uni64(void)
{
    const double r = 9007199254740881.0 / 9007199254740992.0;
    const double d = 362436069876.0 / 9007199254740992.0;
    double x = rng.u[rng.i - 1] - rng.u[rng.j - 1];
    if (x < 0.0) x = x + 1.0;
    rng.u[rng.i - 1] = x;
    rng.i = rng.i - 1; if (rng.i == 0) rng.i = 97;
    rng.j = rng.j - 1; if (rng.j == 0) rng.j = 97;
    rng.c = rng.c - d;
    if (rng.c < 0.0) rng.c = rng.c + r;
    x = x - rng.c;
    return x < 0.0 ? x + 1.0 : x;
}

double fo_normal_dist_seed(int seed1, int seed2)
{
    rng.sw = 1;
    return fo_uni64_seed(seed1, seed2);
}

double fo_normal_dist(double mu, double sigma)
{
    double N01;
    if (rng.sw) {
        rng.sw = 0;
        rng.t = sqrt(-2 * log(1.0 - fo_uni64()));
        rng.uu = M_2PI_D * fo_uni64();
        N01 = rng.t * cos(rng.uu);
    } else {
        rng.sw = 1;
        N01 = rng.t * sin(rng.uu);
    }
    return N01 * sigma + mu;
}

/* initialize-dipoR.F:40-58: alpha outermost, then z, y, x. */
void fo_init_normal(fo_ctx *o, const double avg[3], const double dev[3], int seed1, int seed2,
                    const double k_stripe[3])
{
    const fo_params *p = &o->p;
    (void)fo_normal_dist_seed(seed1, seed2);
    for (int a = 0; a < 3; ++a)
        for (int iz = 0; iz < p->Lz; ++iz)
            for (int iy = 0; iy < p->Ly; ++iy)
                for (int ix = 0; ix < p->Lx; ++ix) {
                    size_t c = ix + (size_t)p->Lx * (iy + (size_t)p->Ly * iz);
                    double v = fo_normal_dist(avg[a], dev[a]);
                    if (cos(M_2PI_D * (k_stripe[0] * ix + k_stripe[1] * iy + k_stripe[2] * iz + 1.0e-4)) < 0.0)
                        v = -v;
                    o->dipoR[3 * c + a] = v;
                }
    memset(o->dipoP, 0, sizeof(double) * 3 * (size_t)o->N);
    memset(o->acouR, 0, sizeof(double) * 3 * (size_t)o->N);
}

/* n consecutive uni64() draws into out (marsaglia_tsang_uni64_module.f:99-103) */
void fo_uni64_fill(double *out, long n)
{
    for (long i = 0; i < n; ++i) out[i] = fo_uni64();
}

/* ------------------------------------------------------------------------- */
/* add_to_Average (average_module.F:57-163).                                 */
/* ------------------------------------------------------------------------- */
void fo_add_to_average(fo_ctx *o, const fo_energy *e, fo_average *a, int64_t *dist1, int32_t *dist3)
{
    const int N = o->N, max_k = 99, W = 2 * max_k + 1;
    const double step_inv = 1.0 / 0.005;                  /* average_module.F:13-14 */
    const double *ee = &e->acou_kinetic;
    double *ae = &a->energy.acou_kinetic;
    for (int i = 0; i < 13; ++i) ae[i] = ae[i] + ee[i];   /* :76 */
    a->e2 = a->e2 + e->total * e->total;                   /* :77 */
    a->dipo_kinetic_true = a->dipo_kinetic_true + e->dipo_kinetic / (o->nose[0] * o->nose[0]);   /* :79 */
    for (int i = 0; i < 6; ++i) a->strain[i] = a->strain[i] + o->strain[i];                    /* :80 */
    double ut[3] = {0, 0, 0}, uut[3] = {0, 0, 0}, pt[3] = {0, 0, 0}, ppt[3] = {0, 0, 0};
    for (int c = 0; c < N; ++c) {                          /* :89-100, z outer, x inner */
        if (!o->sum_p[c]) continue;
        for (int q = 0; q < 3; ++q) {
            ut[q] = ut[q] + o->dipoR[3 * c + q];
            uut[q] = uut[q] + o->dipoR[3 * c + q] * o->dipoR[3 * c + q];
        }
    }
    for (int q = 0; q < 3; ++q) { a->u[q] = a->u[q] + ut[q]; a->uu[q] = a->uu[q] + ut[q] * ut[q]; }  /* :101-104 */
    a->uu[3] = a->uu[3] + ut[1] * ut[2];
    a->uu[4] = a->uu[4] + ut[2] * ut[0];
    a->uu[5] = a->uu[5] + ut[0] * ut[1];
    for (int c = 0; c < N; ++c) {                          /* :111-121 */
        if (!o->sum_p[c]) continue;
        for (int q = 0; q < 3; ++q) {
            pt[q] = pt[q] + o->Z_u_R[3 * c + q];
            ppt[q] = ppt[q] + o->Z_u_R[3 * c + q] * o->Z_u_R[3 * c + q];
        }
    }
    for (int q = 0; q < 3; ++q) { a->p[q] = a->p[q] + pt[q]; a->pp[q] = a->pp[q] + pt[q] * pt[q]; }  /* :122-125 */
    a->pp[3] = a->pp[3] + pt[1] * pt[2];
    a->pp[4] = a->pp[4] + pt[2] * pt[0];
    a->pp[5] = a->pp[5] + pt[0] * pt[1];
    if (dist1 || dist3) {                                  /* :130-160 */
        for (int c = 0; c < N; ++c) {
            if (!o->sum_p[c]) continue;
            long k[3];
            int in3 = 1;
            for (int q = 0; q < 3; ++q) {
                k[q] = lround(o->dipoR[3 * c + q] * step_inv);   /* nint */
                const int in = labs(k[q]) <= max_k;
                if (in && dist1) dist1[q + 3 * (k[q] + max_k)] += 1;
                in3 = in3 && in;
            }
            if (in3 && dist3) dist3[(k[0] + max_k) + W * ((k[1] + max_k) + W * (k[2] + max_k))] += 1;
        }
    }
    for (int q = 0; q < 3; ++q) { a->u_tmp[q] = ut[q]; a->uu_tmp[q] = uut[q]; a->p_tmp[q] = pt[q]; a->pp_tmp[q] = ppt[q]; }
    a->n_add += 1;
}

/* wave() of wave_module.F:8-67 and the linear field sweep of
 * hysteresis_loop.F:47-49 (type 5; the amplitude before n_thermalize),
 * integer and real operations in the reference's order.  Types: 1
 * triangular_cos, 2 triangular_sin, 3 ramping_on, 4 ramping_off, 5 the
 * hysteresis sweep; anything else returns the amplitude (wave's default). */
void fo_wave(int type, const double amp[3], long period, long n_thermalize, long n_average, long i_step,
             double out[3])
{
    if (amp[0] == 0.0 && amp[1] == 0.0 && amp[2] == 0.0) {          /* :19-24 */
        for (int q = 0; q < 3; ++q) out[q] = amp[q];
        return;
    }
    if (type == 1 || type == 2) {                                   /* :27-39 */
        long i = i_step - n_thermalize;
        if (i < 0) i = 0;
        long tmp;
        if (type == 1) tmp = labs(i % period - period / 2) - period / 4;
        else tmp = labs((i + 3 * period / 4) % period - period / 2) - period / 4;
        for (int q = 0; q < 3; ++q) out[q] = amp[q] * (double)tmp / ((double)period / 4.0);
    } else if (type == 3) {                                         /* :40-50 */
        for (int q = 0; q < 3; ++q) out[q] = i_step < period ? amp[q] * (double)i_step / (double)period : amp[q];
    } else if (type == 4) {                                         /* :51-61 */
        for (int q = 0; q < 3; ++q)
            out[q] = i_step < period ? amp[q] * (double)(period - i_step) / (double)period : 0.0;
    } else if (type == 5) {                                         /* hysteresis_loop.F:47-49 */
        for (int q = 0; q < 3; ++q)
            out[q] = i_step >= n_thermalize
                         ? amp[q] - 2 * amp[q] * (double)(i_step - n_thermalize) / (double)n_average
                         : amp[q];
    } else {
        for (int q = 0; q < 3; ++q) out[q] = amp[q];
    }
}
