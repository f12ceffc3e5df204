/*
 * feram_oracle.h -- CPU FP64 restatement of feram's per-step force-and-integrate
 * path (the checker).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle that the HIP product is
 * checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker (or as the timed CPU
 * baseline).  Nothing in the product (feram_amd/) links, loads or calls it.
 *
 * Every routine follows the reference Fortran in operation order; the citation
 * next to each declaration names the file:line it restates (paths relative to
 * the reference's src/ directory).
 *
 * Layout (all unpadded, x fastest):
 *   cell index          c = x + Lx*(y + Ly*z)
 *   3-component fields  f[3*c + a]          (component fastest, as the reference)
 *   half-spectrum index k = kx + Kx*(ky + Ly*kz), Kx = Lx/2+1
 *   3x3 tables          t[9*k + 3*b + a]    (Fortran column-major (a,b) order)
 *   spectra             s[2*(ncomp*k + j) + {0:re,1:im}]
 */
#ifndef FERAM_ORACLE_H
#define FERAM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors the user-given part of Param_type (param_module.F:8-111). */
typedef struct fo_params {
    int Lx, Ly, Lz;
    int method;            /* 0 'md' (Nose-Poincare), 1 'vs', 2 'lf' */
    int bulk_or_film;      /* 0 bulk, 1 film, 2 epit, 3 e110 */
    int gap_id;
    int E_field_always_zero;
    int pad_;
    double GPa[3];
    double kelvin;
    double mass_dipo[3];   /* [eV A^-2 ps^2] = amu / 9648.53377163899 */
    double mass_acou[3];   /* <= 0 : acoustic displacement optimised */
    double Q_Nose;
    double a0;
    double epi_strain;
    double dt;
    double external_E_field[3];
    double P_kappa2[3], P_alpha[3], P_gamma[3];
    double P_k1[3], P_k2[3], P_k3[3], P_k4[3];
    double j[7];
    double B11, B12, B44;
    double B1xx, B1yy, B4yz;
    double modulation_constant;
    double Z_star[3];
    double epsilon_inf;
    double kappa;          /* <= 0 : derived as in make_Param */
    double gap_dipole_u[3];
} fo_params;

/* Energy_type (energy_module.F:7-23), same member order. */
typedef struct fo_energy {
    double acou_kinetic, dipo_kinetic, short_range, long_range, dipole_E_field,
           unharmonic, homo_strain, homo_coupling, inho_strain, inho_coupling,
           inho_modulation, total, H_Nose_Poincare;
} fo_energy;

typedef struct fo_ctx fo_ctx;

/* alloc_Coord + make_Param (coord_module.F:67-432, param_module.F:452-480) */
fo_ctx *fo_create(const fo_params *p);
void    fo_destroy(fo_ctx *o);
/* dipole_dipole (dipole-dipole.F:6-73) and elastic (elastic.F:7-237);
 * modulation_R may be NULL (no .modulation file). */
int     fo_setup_tables(fo_ctx *o, const double *modulation_R);

/* Derived parameters (param_module.F:452-480). */
double  fo_kappa(const fo_ctx *o);
double  fo_pNa03(const fo_ctx *o, int i);
double  fo_target_kinetic_energy(const fo_ctx *o);
double  fo_temperature_eV(const fo_ctx *o);

/* Field access: name in {"dipoR","dipoP","dVddi","acouR","acouP","dVdacR",
 * "E_field","Z_starR","mass_dipo_inv","mass_acou_inv","long_range",
 * "short_range","inhomo_K","inhomo_K_G","inhomo_coup2_K","acouK","didiK",
 * "dipoK","c_k_modulation_k","inhoR","dVddi_long"} ; returns NULL if unknown. */
double *fo_field(fo_ctx *o, const char *name);
int    *fo_relaxor(fo_ctx *o);
unsigned char *fo_sum_p(fo_ctx *o);
void    fo_recount_sum_p(fo_ctx *o);          /* read_defects_u.F:71-77 */
double *fo_strain(fo_ctx *o);                /* c%strain(6) */
double *fo_homogeneous(fo_ctx *o);           /* c%homogeneous(6,6) col-major */
double *fo_nose(fo_ctx *o);                  /* {s_Nose, pi_Nose, H0} */

/* Hot path. */
void fo_mirror_force_mirror(fo_ctx *o, fo_energy *e);      /* mirror_force_mirror.F:6-89 */
void fo_nose_poincare(fo_ctx *o, fo_energy *e);            /* nose-poincare.F:10-78 */
void fo_leapfrog(fo_ctx *o, fo_energy *e, int i_step);     /* leapfrog.F:9-157 */
/* Heat-source / sink layers of leapfrog (leapfrog.F:38-83): delta_e and
 * delta_e_acou [eV/unitcell]; fo_layers(o, 0|1) are e_layer_dipo /
 * e_layer_acou (Lz values) of the last step. */
void fo_set_heat_source(fo_ctx *o, double delta_e, double delta_e_acou);
double *fo_layers(fo_ctx *o, int which);
void fo_velocity_scaling(fo_ctx *o, fo_energy *e);         /* velocity_scaling.F:9-103 */
double fo_kinetic_energy_dipo(fo_ctx *o);                  /* kinetic_energy_dipo.F:1-29 */
double fo_kinetic_energy_acou(fo_ctx *o);                  /* kinetic_energy_acou.F:1-28 */
double fo_sum_potential_energies(const fo_energy *e);      /* energy_module.F:73-87 */

/* Average_type sums (average_module.F:17-30) and add_to_Average
 * (average_module.F:57-163).  dist1 is distribution(3,-99:99) with the
 * component fastest, dist3 distribution3d(-99:99,-99:99,-99:99) with kx
 * fastest; either may be NULL. */
typedef struct fo_average {
    fo_energy energy;
    double e2, acou_kinetic_true, dipo_kinetic_true;
    double strain[6];
    double u[3], uu[6], p[3], pp[6];
    double u_tmp[3], uu_tmp[3], p_tmp[3], pp_tmp[3];
    int64_t n_add;
} fo_average;
void fo_add_to_average(fo_ctx *o, const fo_energy *e, fo_average *a, int64_t *dist1, int32_t *dist3);
void fo_nose_poincare_hamiltonian(fo_ctx *o, fo_energy *e);/* energy_module.F:128-148 */
/* initialize_dipoR tail (initialize-dipoR.F:59-69) followed by the H0 set-up
 * of molecular_dynamics (molecular-dynamics.F:44-55). */
void fo_md_prologue(fo_ctx *o, fo_energy *e);

/* Seeded normal initialisation of u (initialize-dipoR.F:40-58). */
void fo_init_normal(fo_ctx *o, const double avg[3], const double dev[3],
                    int seed1, int seed2, const double k_stripe[3]);

/* Marsaglia-Tsang uni64 and normal_dist (marsaglia_tsang_uni64_module.f). */
double fo_uni64_seed(int seed1, int seed2);
double fo_uni64(void);
void   fo_uni64_fill(double *out, long n);
double fo_normal_dist_seed(int seed1, int seed2);
double fo_normal_dist(double mu, double sigma);

/* FFT building blocks with FFTW's conventions (forward e^{-i}, unnormalised).
 * r2c: real field f[ncomp*c + comp] -> half spectrum s[2*(ncomp*k+comp)+ri].
 * c2r: ignores Im at kx=0 and kx=Lx/2 after the y/z inverse passes. */
void fo_r2c(fo_ctx *o, const double *in, int ncomp, int comp, double *out);
void fo_c2r(fo_ctx *o, const double *in, int ncomp, int comp, double *out);
/* 1-D complex DFT of length n (in place, interleaved re/im), sign -1 or +1. */
void fo_fft1d(double *x, int n, int sign);

/* Uniform field of step i_step: wave() (wave_module.F:8-67; types 1
 * triangular_cos, 2 triangular_sin, 3 ramping_on, 4 ramping_off) or the
 * hysteresis-loop sweep (type 5, hysteresis_loop.F:47-49). */
void fo_wave(int type, const double amp[3], long period, long n_thermalize, long n_average, long i_step,
             double out[3]);

/* Erfc-based Ewald decay functions (decay_function_module.f:13-33). */
double fo_B_decay(double r, double rr, double kappa);
double fo_C_decay(double r, double rr, double kappa);

#ifdef __cplusplus
}
#endif
#endif
