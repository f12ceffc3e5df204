/*
 * feram_hip.h -- C ABI of the MI355X (gfx950) per-step force-and-integrate
 * path of feram.
 *
 * The entry points replace, at link time, the Fortran subroutines that the
 * reference's time-loop drivers call once per MD step.  Paths below are
 * relative to the reference's src/ directory.
 *
 *   fhip_force            <- mirror_force_mirror(p,c,e)   mirror_force_mirror.F:6-89
 *                            (also called from initialize_dipoR, initialize-dipoR.F:62)
 *   fhip_nose_poincare    <- nose_poincare(p,c,e)         nose-poincare.F:10-78,
 *                            called at molecular-dynamics.F:113-116
 *   fhip_leapfrog         <- leapfrog(p,c,e,i_step)       leapfrog.F:9-157,
 *                            called at molecular-dynamics.F:119-120
 *   fhip_velocity_scaling <- velocity_scaling(p,c,e)      velocity_scaling.F:9-103,
 *                            called at molecular-dynamics.F:117-118, hysteresis_loop.F:69
 *   fhip_md_prologue      <- initialize-dipoR.F:59-69 followed by the H0 set-up of
 *                            molecular-dynamics.F:44-55
 *
 * Host arrays keep the reference's padded layout (coord_module.F:91-93,134-141):
 * 3-component fields are (3, 0:2(Lx/2+1)-1, 0:Ly+padding_y-1, 0:Lz-1) real*8
 * with the component fastest; per-cell scalars drop the leading 3; the
 * long-range table is long_range(3,3, 0:Lx/2, 0:Ly+padding_y-1, 0:Lz-1).
 * Padding entries are ignored on input and left untouched on output.
 *
 * Every function returns FHIP_OK (0) or a negative FHIP_ERR_* code;
 * fhip_last_error() gives the message.  One context per process and device;
 * calls on one context must not overlap (the reference calls from the master
 * thread, outside OpenMP regions).
 */
#ifndef FERAM_HIP_H
#define FERAM_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
    FHIP_OK = 0,
    FHIP_ERR_ARG = -1,          /* bad argument or size */
    FHIP_ERR_UNSUPPORTED = -2,  /* configuration outside what the library runs */
    FHIP_ERR_HIP = -3,          /* HIP runtime error (message has the detail) */
    FHIP_ERR_ALLOC = -4,        /* device memory allocation failed */
    FHIP_ERR_STATE = -5         /* call out of order */
};

/* Methods and geometries (param_module.F:36, :43). */
enum { FHIP_MD = 0, FHIP_VS = 1, FHIP_LF = 2 };
enum { FHIP_BULK = 0, FHIP_FILM = 1, FHIP_EPIT = 2, FHIP_E110 = 3 };
/* Arithmetic of spectra and local forces; thermostat and energy sums are FP64 always. */
enum { FHIP_FP64 = 0, FHIP_FP32 = 1 };

/* The parameters of Param_type (param_module.F:8-111) the step reads. Masses
 * are in eV A^-2 ps^2 (amu / 9648.53377163899); mass_acou <= 0 means the
 * acoustic displacement is optimised each step (param_module.F:135). */
typedef struct fhip_params {
    int32_t Lx, Ly, Lz, padding_y;
    int32_t method, bulk_or_film, precision, E_field_always_zero;
    double dt, kelvin, Q_Nose, a0, epi_strain;
    double GPa[3];
    double mass_dipo[3], mass_acou[3];
    double P_kappa2[3], P_alpha[3], P_gamma[3], P_k1[3], P_k2[3], P_k3[3], P_k4[3];
    double j[7];
    double B11, B12, B44, B1xx, B1yy, B4yz;
    double modulation_constant;
    double Z_star[3];
    double epsilon_inf;
    double kappa;                 /* <= 0: sqrt(pi)/a0/N^(1/6) (param_module.F:462-464) */
    double external_E_field[3];
    int32_t gap_id;               /* film/epit dead layers 0, 1, 2 (param_module.F:268-270, 470-478) */
    int32_t reserved;             /* 0 */
} fhip_params;

/* Energy_type (energy_module.F:7-23), same member order, supercell totals in eV. */
typedef struct fhip_energy {
    double acou_kinetic, dipo_kinetic, short_range, long_range, dipole_E_field,
           unharmonic, homo_strain, homo_coupling, inho_strain, inho_coupling,
           inho_modulation, total, H_Nose_Poincare;
} fhip_energy;

/* 3-component fields of Coord_type (coord_module.F:14-64). */
enum {
    FHIP_DIPO_R = 0,   /* c%dipoR   u                                   */
    FHIP_DIPO_P = 1,   /* c%dipoP   p                                   */
    FHIP_DVDDI  = 2,   /* c%dVddi   dV/du (read only)                   */
    FHIP_ACOU_R = 3,   /* c%acouR   w                                   */
    FHIP_ACOU_P = 4,   /* c%acouP   p_w                                 */
    FHIP_DVDAC  = 5,   /* dV/dw in the dVdac_ptr buffer (read only)      */
    FHIP_E_LOCAL = 6,  /* c%E_field_local; the uniform part is set with fhip_set_E_field_add */
    FHIP_ACOU_R_OPT = 7, /* w reconstructed from the optimised w(k) as write_system does
                           (coord_module.F:457-469) (read only)           */
    FHIP_DIPO_R_AVG = 8  /* c%dipoRavg, the sum of u over the averaging steps
                            (molecular-dynamics.F:144-148)                */
};

typedef struct fhip_ctx fhip_ctx;

/* alloc_Coord + dipole_dipole + elastic (coord_module.F:67-432,
 * dipole-dipole.F:6-73, elastic.F:7-237): device buffers, FFT twiddles, the
 * Ewald long-range table (built on the device) and the elastic constants.
 * State starts at u = p = w = p_w = 0, s = 1, pi = 0, H0 = 0. */
int fhip_create(const fhip_params *p, fhip_ctx **ctx);
void fhip_destroy(fhip_ctx *ctx);
const char *fhip_last_error(const fhip_ctx *ctx);

/* Replace the long-range table with the host's c%long_range (already scaled
 * by 14.399644/epsilon_inf, dipole-dipole-long-range.F:146-150). */
int fhip_set_long_range(fhip_ctx *ctx, const double *long_range);
/* Copy the device table out as c%long_range (symmetric: both triangles are
 * written), e.g. for the dispersion output of dipole-dipole.F:24-70. */
int fhip_get_long_range(fhip_ctx *ctx, double *long_range);
/* c%modulation_R (elastic.F:187-200); requires modulation_constant != 0. */
int fhip_set_modulation(fhip_ctx *ctx, const double *modulation_R);
/* Per-cell data set by read_relaxor / read_defects_u (read_relaxor.F:24-33,
 * read_defects_u.F:34-52): relaxor species 1..3 (NULL: all 1), sum_p as 0/1
 * (NULL: all counted), mass_dipo_inv (NULL: 1/mass_dipo(species)); a cell's
 * mass_dipo_inv must be 1/mass_dipo(species) or 0 (frozen defect).  For film
 * and epit the rules of read_defects_u.F:52-69 are applied on top (also at
 * fhip_create): layers (Lz+1)/2 .. Lz-1 are not counted, and the dead layers
 * (gap_id 1: iz = 0; 2: iz = 0 and Lz/2) are frozen and not counted; their u
 * (gap_dipole_u) is the caller's, as fhip_put leaves it. */
int fhip_set_cells(fhip_ctx *ctx, const int32_t *relaxor, const int32_t *sum_p,
                   const double *mass_dipo_inv);

int fhip_put(fhip_ctx *ctx, int field, const double *host);
int fhip_get(fhip_ctx *ctx, int field, double *host);
/* Unpadded structure-of-arrays variants: host[a*N + x + Lx*(y + Ly*z)]. */
int fhip_put_soa(fhip_ctx *ctx, int field, const double *host);
int fhip_get_soa(fhip_ctx *ctx, int field, double *host);

/* Thermostat state c%s_Nose, c%pi_Nose, c%H0 (coord_module.F:53). */
int fhip_get_nose(fhip_ctx *ctx, double *s, double *pi, double *H0);
int fhip_set_nose(fhip_ctx *ctx, double s, double pi, double H0);
/* c%strain(6) of the last force evaluation. */
int fhip_get_strain(fhip_ctx *ctx, double strain[6]);
/* Uniform part of E(R) = E_local(R) + E_add (molecular-dynamics.F:66-76).
 * film/epit: fhip_create starts from initialize_E_field's mirrored field
 * (initialize_E_field.F:68-71); this call, like the per-step updates of the
 * reference, leaves the field unmirrored. */
int fhip_set_E_field_add(fhip_ctx *ctx, const double E_add[3]);
/* Time-dependent uniform field, evaluated on the device at the start of every
 * step of fhip_nose_poincare / fhip_leapfrog / fhip_velocity_scaling (and of
 * batches): E_add(i_step) = wave(E_wave_type, i_step, n_thermalize,
 * n_E_wave_period, external_E_field) as molecular-dynamics.F:66-76 calls
 * wave_module.F:8-67, or the linear sweep of the hysteresis loop
 * (hysteresis_loop.F:47-49; amplitude before n_thermalize).  i_step is the
 * index of the next step (molecular_dynamics counts from 1, hysteresis_loop
 * from 0); every step advances it.  FHIP_E_CONSTANT holds E_add = amplitude
 * (as fhip_set_E_field_add).  film/epit: the field is mirrored until the
 * schedule first rewrites it (step >= 1 for the waves, step >= n_thermalize
 * for the sweep), unmirrored afterwards, as the reference's loops leave it. */
enum { FHIP_E_CONSTANT = 0, FHIP_E_TRIANGULAR_COS = 1, FHIP_E_TRIANGULAR_SIN = 2, FHIP_E_RAMPING_ON = 3,
       FHIP_E_RAMPING_OFF = 4, FHIP_E_HYSTERESIS = 5 };
int fhip_set_E_schedule(fhip_ctx *ctx, int type, const double amplitude[3], int period, int n_thermalize,
                        int n_average, int64_t i_step);
/* The index of the next step and the E_add it will use. */
int fhip_get_step(fhip_ctx *ctx, int64_t *i_step, double E_add[3]);
/* Heat-source and sink layers of leapfrog (leapfrog.F:38-83; param_module.F:88-89
 * delta_e, delta_e_acou in eV/unitcell): at the start of every leapfrog step
 * the kinetic energies of the z layers are formed (e_layer_dipo scaled by
 * Lx Ly / n_sum_layer) and the momenta of layer Lz/4 scaled up and of layer
 * 3Lz/4 down by sqrt((e +- Lx Ly delta_e)/e).  Leapfrog contexts only. */
int fhip_set_heat_source(fhip_ctx *ctx, double delta_e, double delta_e_acou);
/* c%e_layer_dipo, c%e_layer_acou (Lz values each, either may be NULL) of the
 * last step, or with averaged != 0 their sums over the averaging steps
 * (avg%e_layer_*, average_module.F:82-85; divide by n_add for the .T_layer
 * output of average_module.F:225-238). */
int fhip_get_layer_energies(fhip_ctx *ctx, double *e_dipo, double *e_acou, int averaged);
/* Thermostat / velocity-scaling temperature p%kelvin, as a new run of a
 * temperature scan reads it (param_module.F:224-226,468); the next
 * fhip_md_prologue starts the run. */
int fhip_set_kelvin(fhip_ctx *ctx, double kelvin);

int fhip_force(fhip_ctx *ctx, fhip_energy *e);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the force evaluation of fhip_force (mirror_force_mirror)
 * without a context, on caller-owned device arrays.  It calls no hipMalloc,
 * hipFree or synchronisation: it only enqueues kernels on `stream` (a
 * hipStream_t; NULL is the default stream).  Every intermediate (spectra,
 * tables, partial sums, scalars) lives in a fixed 16 MiB device array of the
 * library, one per device, held until the process ends; the supercell must
 * fit it (9 FP64 spectra, the L(k) table and the cell codes: N = Lx Ly Lz up
 * to about 110592, e.g. 48^3), larger ones are rejected with
 * FHIP_ERR_UNSUPPORTED.  Calls that overlap on one device race on that array.
 *   p          FP64, bulk or e110, acoustic displacement optimised
 *              (mass_acou <= 0), no modulation; every cell species 1, counted;
 *   long_range device array, c%long_range in the layout of fhip_set_long_range
 *              (fhip_get_long_range of a context with the same parameters);
 *   u          device array u[a*N + x + Lx*(y + Ly*z)] (fhip_put_soa's layout);
 *   dVddi      device array, same layout, receives dV/du;
 *   energy     device pointer, receives the Energy_type of the evaluation
 *              (kinetic terms 0, as for a fresh context);
 *   strain     device double[6] or NULL, receives c%strain;
 *   err        host buffer of errlen bytes (or NULL) for the message of a
 *              failed call.
 * The first call on a device resolves the array's address there; make it
 * before capturing the call into a graph. */
int fhip_force_static(const fhip_params *p, const double *long_range, const double *u, double *dVddi,
                      fhip_energy *energy, double *strain, void *stream, char *err, int errlen);
int fhip_md_prologue(fhip_ctx *ctx, fhip_energy *e);
/* nsteps consecutive steps (>= 1); e receives the energies after the last one. */
int fhip_nose_poincare(fhip_ctx *ctx, int nsteps, fhip_energy *e);
int fhip_leapfrog(fhip_ctx *ctx, int i_step, int nsteps, fhip_energy *e);
int fhip_velocity_scaling(fhip_ctx *ctx, int nsteps, fhip_energy *e);

/* Binary checkpoint for exact resume.  The reference's restart is the .coord
 * text written by write_system (coord_module.F:472-483: u in f10.6, p/s and
 * the forces in e14.6) and read back by initialize-dipoR.F:22-38, which loses
 * the state beyond those digits; this keeps every double a later step call
 * reads: the thermostat and integrator scalars (s, pi, H0, the drift factor,
 * pending velocity scaling, the E-field schedule and its step index, strain,
 * energies), u, p, dV/du (and w, p_w, dV/dw with acoustic MD), the optimised
 * w(k) when materialised, the averaging sums, histograms and dipoRavg, and
 * the heat-source layer energies.  Loading into a context created with the
 * same parameters (and cells, modulation and heat source set as before) and
 * stepping gives bitwise what the saving context's next calls give.  The
 * buffer is host memory of fhip_checkpoint_bytes() bytes; the format is the
 * library's (a header with the supercell, method, precision and geometry,
 * checked on load, then the raw sections). */
int64_t fhip_checkpoint_bytes(fhip_ctx *ctx);
int fhip_checkpoint_save(fhip_ctx *ctx, void *buf, int64_t bytes);
int fhip_checkpoint_load(fhip_ctx *ctx, const void *buf, int64_t bytes);

/* Average_type (average_module.F:17-30) as the sums add_to_Average builds
 * (average_module.F:57-128), before divide_and_write_Average divides them. */
typedef struct fhip_average {
    fhip_energy energy;           /* sum of e (:76) */
    double e2;                    /* sum of e%total**2 (:77) */
    double acou_kinetic_true;     /* stays 0: its accumulation is commented out in the reference (:78) */
    double dipo_kinetic_true;     /* sum of e%dipo_kinetic / s_Nose**2 (:79) */
    double strain[6];             /* sum of c%strain (:80) */
    double u[3], uu[6];           /* sum of U and of U_a U_b (xx yy zz yz zx xy), U = sum of u over
                                     the sum_p cells (:87-107) */
    double p[3], pp[6];           /* the same for P = sum of Z_u_R (:109-128) */
    double u_tmp[3], uu_tmp[3];   /* U and the sum of u**2 of the last call (the detail lines :165-176) */
    double p_tmp[3], pp_tmp[3];   /* P and the sum of Z_u_R**2 of the last call */
    int64_t n_add;                /* number of add_to_Average calls */
} fhip_average;

/* reset_Average (average_module.F:32-55).  Options: FHIP_AVG_HISTOGRAMS also
 * clears and fills distribution(3,-99:99) and distribution3d(-99:99,-99:99,-99:99);
 * FHIP_AVG_DIPO_R_AVG also clears c%dipoRavg (field FHIP_DIPO_R_AVG) and adds
 * u to it at every add, as molecular-dynamics.F:144-148 does. */
enum { FHIP_AVG_HISTOGRAMS = 1, FHIP_AVG_DIPO_R_AVG = 2 };
int fhip_avg_reset(fhip_ctx *ctx, int options);
/* add_to_Average(p,c,e,avg) (average_module.F:57-163) with the state and
 * energies of the last step or force evaluation, on the device. */
int fhip_avg_add(fhip_ctx *ctx);
/* enable != 0: fhip_nose_poincare / fhip_leapfrog / fhip_velocity_scaling call
 * add_to_Average after every step, as molecular-dynamics.F:141-148 does in
 * the averaging phase, without leaving the device. */
int fhip_avg_auto(fhip_ctx *ctx, int enable);
/* The sums; distribution (3*199 int64, component fastest) and distribution3d
 * (199^3 int32, kx fastest) may be NULL. */
int fhip_avg_get(fhip_ctx *ctx, fhip_average *out, int64_t *distribution, int32_t *distribution3d);

/* Replica batches.  The reference runs one supercell per process and sweeps
 * temperature or field by running the time loop of molecular_dynamics
 * (molecular-dynamics.F:98-150) once per configuration; a batch steps many
 * independent contexts (replicas) together on one device, each exactly as its
 * own fhip_nose_poincare / fhip_leapfrog / fhip_velocity_scaling call would
 * (same kernels, same arithmetic, bitwise the same state).  Replicas with the
 * same edges, precision and acoustic setting form a size class whose stages
 * run as one launch each; the steps between the first and the last of a call
 * replay a captured HIP graph.
 *
 * Every context of a batch has the same method and edges of at most 64; a
 * context belongs to at most one batch and must outlive it.  Per-context calls
 * (put/get, averages, E_add, ...) remain valid between batch calls; their
 * effects are picked up by the next fhip_batch_step.  Averaging settings
 * (fhip_avg_auto, histograms) must agree within a size class. */
typedef struct fhip_batch fhip_batch;
int fhip_batch_create(fhip_ctx *const *ctxs, int n, fhip_batch **batch);
void fhip_batch_destroy(fhip_batch *batch);
const char *fhip_batch_last_error(const fhip_batch *batch);
/* nsteps (>= 1) steps of every replica with its context's method; e (NULL or
 * n entries, in the order of ctxs) receives each replica's energies after the
 * last step, as the per-context step call would return them. */
int fhip_batch_step(fhip_batch *batch, int nsteps, fhip_energy *e);
/* enable = 0: launch every step directly instead of replaying the graph. */
int fhip_batch_graph(fhip_batch *batch, int enable);
/* No counterpart in the reference (its tile shapes are FFTW's business).
 * The first fhip_batch_step of a size class times two tile shapes for each of
 * two knobs on the device -- rows per x-transform workgroup and lines per y
 * tile, full or half -- and keeps the faster (default on; enable = 0 keeps
 * the per-context shapes).  Plans are cached per (edges, precision, acoustic,
 * replicas) for the process.  fhip_batch_plan reports class cls: shape =
 * (Lx, Ly, Lz, x rows, y lines), us = measured us per launch (x full, x half,
 * y full, y half); returns 1 once the class is tuned, 0 before. */
int fhip_batch_tune(fhip_batch *batch, int enable);
int fhip_batch_plan(const fhip_batch *batch, int cls, int32_t shape[5], double us[4]);
int fhip_batch_size(const fhip_batch *batch);
int fhip_batch_classes(const fhip_batch *batch);

/* Learned warm start (no counterpart in the reference, which starts every run
 * from init_dipo_avg / init_dipo_dev or a .restart file, initialize-dipoR.F:22-58).
 * A 3-128-128-13 perceptron with GELU hidden layers, BF16 MFMA arithmetic with
 * FP32 accumulation and FP32 master weights, trained with Adam on the device.
 * Inputs and targets are normalised by the caller; the meaning of the 3 inputs
 * and 13 outputs is the caller's (feram_amd.warmstart: temperature, E_z and
 * pressure -> sorted |<u>|, per-component variance of u, homogeneous strain and
 * the potential energy per cell).
 * Parameters (fhip_mlp_nparams() floats): W1[128][3], b1[128], W2[128][128],
 * b2[128], W3[13][128], b3[13], row-major W[out][in]. */
typedef struct fhip_mlp fhip_mlp;
enum { FHIP_MLP_IN = 3, FHIP_MLP_HIDDEN = 128, FHIP_MLP_OUT = 13, FHIP_MLP_MAX_SAMPLES = 65536 };
int fhip_mlp_nparams(void);
/* Glorot-uniform weights and zero biases from seed. */
int fhip_mlp_create(uint64_t seed, fhip_mlp **mlp);
void fhip_mlp_destroy(fhip_mlp *mlp);
const char *fhip_mlp_last_error(const fhip_mlp *mlp);
int fhip_mlp_set_params(fhip_mlp *mlp, const float *params);   /* also resets Adam's moments */
int fhip_mlp_get_params(fhip_mlp *mlp, float *params);
/* y[n][13] = network(x[n][3]) (host arrays). */
int fhip_mlp_forward(fhip_mlp *mlp, int n, const float *x, float *y);
/* loss = sum_i sum_o w_o (net(x_i)_o - y_io)^2 / n and its gradient with
 * respect to every parameter (w may be NULL: all 1). */
int fhip_mlp_gradient(fhip_mlp *mlp, int n, const float *x, const float *y, const float *w,
                      float *loss, float *grad);
/* iters full-batch Adam steps (beta 0.9, 0.999, eps 1e-8) at learning rate lr;
 * loss_history (NULL or iters floats) receives the loss before each step. */
int fhip_mlp_train(fhip_mlp *mlp, int n, const float *x, const float *y, const float *w, int iters,
                   float lr, float *loss_history);

/* Measurement: when enabled, HIP events bracket every kernel launch on the
 * context's stream and accumulate per-kernel time.  fhip_kernel_stats fills
 * up to n entries (name, total ms, launches) and returns the entry count. */
int fhip_profile(fhip_ctx *ctx, int enable);
/* No counterpart in the reference (its steps are loops of subroutine calls):
 * multi-step calls replay one captured step as a hipGraph for all but the last
 * step (default on; off while profiling).  enable = 0 launches every kernel
 * directly; the arithmetic is the same either way. */
int fhip_use_graph(fhip_ctx *ctx, int enable);
int fhip_kernel_stats(fhip_ctx *ctx, int n, const char **names, double *total_ms, int64_t *launches);
int fhip_synchronize(fhip_ctx *ctx);
/* No counterpart in the reference (measurement): the attainable HBM rate of
 * the current device, a copy of `bytes` (a whole number of 16-byte words)
 * between two device buffers, timed with HIP events over reps launches after
 * two of warm-up; gbs = (bytes read + bytes written) / time per copy.  The
 * two buffers are allocated and freed by the call. */
int fhip_copy_rate(int64_t bytes, int reps, double *gbs, double *us_per_copy);
/* Bytes of device memory the context holds. */
int64_t fhip_device_bytes(const fhip_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
