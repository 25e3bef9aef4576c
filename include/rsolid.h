/*
 * rsolid.h -- C ABI of the MI355X-native 2D rapid-solidification time-stepper
 * (coupled phase field psi/phi, CALPHAD solute c, coarse-grid temperature T).
 *
 * The reference has no library boundary: its "interface" is the set of CUDA
 * kernels that main() launches (src/main.hpp:54-85, fun.cu:431-514), the
 * run-state struct `constants` (fun.cu:197-241, src/main.hpp:10-53) and the
 * parameter set (src/constants.hpp:15-196 plus the -D flags of run.sh:75-90).
 * This header replaces that surface with a coarse-grained, POD-only C ABI:
 *
 *   rs_params     <- constants.hpp names + fun.cu model macros      (inputs)
 *   rs_constants  <- struct constants (fun.cu:197-241)              (run state)
 *   rs_create/... <- main()'s cudaMalloc / set_parameters / set_initial
 *   rs_step       <- one pass of the loop body fun.cu:448-509
 *                    (pull_back, compute_phi, compute_psi, compute_c_CALPHAD,
 *                     integrate_phi, compute_temperature x r,
 *                     interpolate_temperature, swaps, l_offset_Vp update)
 *
 * Conventions: every call returns an int status (RS_OK == 0, negative on
 * error) and never throws; rs_last_error() gives the message. Host arrays
 * passed in are only read/written during the call and never retained.
 * Field layout everywhere is the reference's: row-major pos(i,j) = j*Nx + i,
 * x (i) = growth direction = contiguous axis, ghost cells included
 * (fun.cu:176-179). One context per host thread; each context owns its own
 * HIP stream and device buffers.
 */
#ifndef RSOLID_H
#define RSOLID_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RS_ABI_VERSION 2

/* ---- status codes ------------------------------------------------------ */
enum {
    RS_OK = 0,
    RS_ERR_ARG = -1,      /* bad argument (null pointer, bad enum, bad size) */
    RS_ERR_HIP = -2,      /* a HIP runtime call failed                       */
    RS_ERR_DOMAIN = -3,   /* physical/geometric domain violated (T window)   */
    RS_ERR_NOMEM = -4,    /* device allocation failed                        */
    RS_ERR_STATE = -5,    /* call not valid in the context's current state   */
    RS_ERR_NONFINITE = -6 /* finite-value check found a NaN or Inf (opt-in)   */
};

/* ---- precision / thermodynamics paths ---------------------------------- */
enum {
    RS_PREC_FP64 = 0,     /* plain FP64 path: psi, c, T in FP64              */
    RS_PREC_MIXED = 1,    /* psi/phi and T in FP32; c, flux divergence,
                             mass and ln-terms in FP64                        */
    RS_PREC_SURROGATE = 2 /* RS_PREC_MIXED storage with the learned CALPHAD
                             surrogate fused into the step kernel; needs
                             rs_surrogate_train or rs_surrogate_set first    */
};

/* ---- fields for rs_upload / rs_download -------------------------------- */
enum {
    RS_FIELD_C = 0,       /* Ti mole fraction c, Nx*Ny                       */
    RS_FIELD_PSI = 1,     /* psi (unclamped below -1), Nx*Ny                 */
    RS_FIELD_PHI = 2,     /* phi = clamp(psi,-1,1), Nx*Ny (download only)    */
    RS_FIELD_T = 3,       /* coarse temperature, NxT_all*NyT_all             */
    RS_FIELD_TFINE = 4    /* fine temperature interpolated as the stepper
                             sees it, Nx*Ny (download only)                  */
};

/*
 * Model and material parameters. Field names follow src/constants.hpp and
 * the fun.cu macros they replace (cited per field). Presets filling this
 * struct live in the host library (rs_params_reference, rs_params_config).
 */
typedef struct rs_params {
    /* grid: Nx_all, Ny_all (run.sh:6-7 -> -D Nx_all/Ny_all, fun.cu:38-39) */
    int32_t Nx, Ny;
    int32_t i_tip_target;      /* fun.cu:40 (Nx-268)                         */
    int32_t S;                 /* SS, W/W0 (fun.cu:63)                        */
    int32_t ratio_dxT_dx;      /* fun.cu:109 (10)                             */
    int32_t ifTFC;             /* fun.cu:25: 1 coupled coarse T, 0 frozen T  */
    int32_t if_Vp_ramp;        /* fun.cu:30                                   */
    int32_t dt_cap;            /* fun.cu:1989-1991 cap dt at dt_coef*dx^2/4  */
    int32_t num_fields_output;      /* fun.cu:50 */
    int32_t num_tip_output;         /* fun.cu:51 */
    int32_t num_check_point_output; /* fun.cu:52 */
    int32_t precision;         /* RS_PREC_*                                   */
    int32_t reserved0;

    double W0;                 /* nm (fun.cu:64)                              */
    double Gamma;              /* K*nm (fun.cu:65)                            */
    double muk0;               /* nm/ns/K (fun.cu:68)                         */
    double Tmelt;              /* K (fun.cu:72)                               */
    double deltaT0;            /* K (fun.cu:73)                               */
    double Dl;                 /* nm^2/ns (fun.cu:74)                         */
    double DT;                 /* nm^2/ns thermal diffusivity (fun.cu:111)    */
    double GG;                 /* K/nm (-D GG, run.sh:13)                     */
    double Vp;                 /* m/s (-D Vp, run.sh:12)                      */
    double Vp_ramp_rate;       /* m/s per us (fun.cu:1977, 0.0107587)         */
    double eps1;               /* capillary anisotropy (-D eps1)              */
    double epk1;               /* kinetic anisotropy (-D epk1)                */
    double cinf;               /* far-field c, Ti fraction (fun.cu:62)        */
    double latentHeat;         /* J/mol (fun.cu:87)                           */
    double num_moles;          /* mol/m^3 (fun.cu:92)                         */
    double v0;                 /* m^3/mol (fun.cu:91)                         */
    double cp;                 /* J/mol/K (fun.cu:94)                         */
    double RR;                 /* J/mol/K (fun.cu:103)                        */
    double dx;                 /* W units (fun.cu:165-174)                    */
    double AA;                 /* q(psi) parameter (fun.cu:165-174)           */
    double TCOLD;              /* K (fun.cu:108)                              */
    double L0_bcc, L0_liq;     /* J/mol Redlich-Kister L0 (fun.cu:1668-1669)  */
    double L1_bcc, L1_liq;     /* J/mol Redlich-Kister L1 (0 = reference)     */
    double dt_coef;            /* 0.6 (fun.cu:1989)                           */
    double dtT_coef;           /* 0.9 (fun.cu:2029)                           */
    double length_sample_um;   /* T sample length, um (fun.cu:2037, 50)       */
    double lap_w_axis;         /* 2/3 (fun.cu:1051)                           */
    double lap_w_diag;         /* 1/3 (fun.cu:1052); 1,0 = 5-point            */
    double q_solid;            /* D_s/D_l floor of q(psi); 0 = reference      */
    double total_time;         /* ns (-D total_time, run.sh:18)               */
    double delta_initial;      /* -1 (fun.cu:2013)                            */
    double phi_interface;      /* 0 (fun.cu:126)                              */
    double c_liq0, c_sol0;     /* rs_init_config compositions (Ti fraction)   */
    /* Extension absent from the reference, off when at_coef == 0: the
     * antitrapping solute current J = a_t W (1 - k) c_L (dphi/dt) n
     * (n: unit normal into the liquid; Karma, PRL 87 (2001) 115701) times the
     * solute-trapping factor 1/(1 + |V|/V_D), V = (dphi/dt)/|grad phi| the
     * local interface speed, which turns the partition into Aziz's
     * continuous-growth k(V) = (k + V/V_D)/(1 + V/V_D). */
    double at_coef;            /* a_t; 1/(2 sqrt 2) for phi in [-1, 1]       */
    double V_D;                /* diffusive speed, m/s; 0 = no trapping term */
    double k_part;             /* equilibrium partition c_S / c_L of c       */
    /* Euler angles of the grain, degrees (fun.cu:129-131; 0 in the reference,
     * applied by set_orientation at initialisation, fun.cu:1017); any non-zero
     * angle selects the general-orientation anisotropy kernels */
    double alpha0, beta0, gamma0;
} rs_params;

/* Run state: field-for-field the reference's struct constants
 * (fun.cu:197-241), plus i_left_end_interp (the i_left_end used by the last
 * interpolate_temperature, i.e. the T_fine the next step reads). */
typedef struct rs_constants {
    int64_t step;
    int64_t initial_step;
    int64_t total_step;
    int32_t step_fields;
    int32_t step_tip;
    int32_t step_check_point;
    int32_t step_pull_back;
    int32_t if_begin_tip;
    int32_t if_begin_field;
    int32_t i_offset;
    int32_t i_offset_history_initial;
    int32_t i_offset_history;
    int32_t i_ignore;
    double dt;
    double x_tip_history;
    double delta_initial;
    double delta;
    double l_offset_Vp_initial;
    double l_offset_Vp;
    int32_t i_left_end_initial;
    int32_t i_left_end;
    double dtT;
    int32_t ratio_dt_dtT;
    int32_t NxT;
    int32_t NyT;
    int32_t NxT_all;
    int32_t NyT_all;
    int32_t i_left_end_interp;
    double Tcold;
    double Thot;
    double length_sample;
} rs_constants;

/* Tip record, one row of tip.txt (src/main.cu:2671-2709). */
typedef struct rs_tip {
    int64_t step;
    double t_us;       /* (2)  */
    double delta;      /* (3)  */
    double T_tip;      /* (4)  K */
    double V_tip;      /* (5)  m/s */
    double Vp;         /* (6)  m/s */
    double R;          /* (7)  nm */
    int32_t i, j;      /* (8),(9) */
    double x, y;       /* (10),(11) */
} rs_tip_row;

typedef struct rs_ctx rs_ctx;

/* ---- parameters (host only, no device work) ---------------------------- */

/* Reference preset: fun.cu macro values (fun.cu:20-174) with the run.sh -D
 * flags (run.sh:6-23, 75-90) for the given grid. Replaces set_parameters'
 * inputs. */
int rs_params_reference(rs_params* p, int32_t Nx, int32_t Ny);

/* Ti-10at%Nb configuration preset (frozen or coupled T) on an N x N grid.
 * tfc = 0: frozen T(x,t) = T0 + G(x - V t); tfc = 1: coupled coarse-grid heat
 * equation with thermal diffusivity DT = alpha_ratio * Dl. */
int rs_params_config(rs_params* p, int32_t N, int32_t tfc, double alpha_ratio);

/* Restatement of set_parameters (fun.cu:1984-2058): derives dt, cadences,
 * coarse-T geometry, Tcold/Thot and i_left_end from the parameters. */
int rs_derive(const rs_params* p, rs_constants* k);

/* sizeof(rs_params), sizeof(rs_constants), sizeof(rs_tip_row) and
 * RS_ABI_VERSION, for binding-side layout checks. No counterpart in the
 * reference (it has no ABI). */
int rs_abi_info(int32_t* sizes4);

/* ---- context ------------------------------------------------------------ */
int rs_create(const rs_params* p, rs_ctx** out);
int rs_destroy(rs_ctx* ctx);

/* Context facts with no counterpart in the reference (diagnostics):
 * out[0] = largest error (J/mol) of the mixed path's SGTE polynomial table
 *          against the FP64 routine on its window (0 on the FP64 path),
 * out[1] = Tref, out[2] = window low (K), out[3] = window high (K),
 * out[4] = band height (rows per wavefront), out[5] = waves per block,
 * out[6] = step kernel of the last rs_step launch: -1 the FP64 kernel (or
 *          none yet), else the FP32-field kernel's thermodynamics variant
 *          (0 exact table, 1 BF16 network, 2 split-BF16, 3 FP16 with rows
 *          lacking a trusted cell skipping the network, 4 packed FP16
 *          likewise, 7 FP16 and 8 packed FP16 evaluated in every row),
 * out[7] = 1 if that launch used the identity-orientation kernel,
 * out[8] = how many launches in a row, up to the last, used that kernel. */
int rs_context_info(rs_ctx* ctx, double* out, int32_t n);
/* never NULL; with ctx NULL: why the calling thread's last rs_create or
 * rs_solute_mass_noalloc failed (which rule its arguments broke, or the
 * device error) */
const char* rs_last_error(const rs_ctx* ctx);
int rs_device_name(char* buf, int32_t len);

/* Initial conditions: initial + initial_temperature + interpolate_temperature
 * (fun.cu:983-1028, 559-571). rs_init_config draws one front-position jitter
 * in [-1,1) cells per row from splitmix64(seed). */
int rs_init_reference(rs_ctx* ctx);
int rs_init_config(rs_ctx* ctx, uint64_t seed);

int rs_upload(rs_ctx* ctx, int32_t field, const double* host, int64_t count);
int rs_download(rs_ctx* ctx, int32_t field, double* host, int64_t count);
int rs_get_constants(rs_ctx* ctx, rs_constants* k);
int rs_set_constants(rs_ctx* ctx, const rs_constants* k);

/* Runs nsteps passes of the loop body fun.cu:448-509 starting at the
 * context's current step: pull-back when step % step_pull_back == 0, the
 * fused phi/psi/c update, latent heat, coarse T sub-steps, swaps and the
 * l_offset_Vp update. Work is enqueued on the context stream; the call
 * returns after the last pull-back decision (which needs the tip column).
 * Surrogate contexts also wait for the stream once after their first 64
 * surrogate steps and then once per 1024, to read the row counts that pick
 * the step kernel variant (unless RS_SURR_ALWAYS fixed it at rs_create). */
int rs_step(rs_ctx* ctx, int64_t nsteps);

/* Pieces of the loop body, for drivers that interleave their own output. */
int rs_tip_column(rs_ctx* ctx, int32_t* i_tip);            /* get_itip_all_grain */
int rs_pull_back(rs_ctx* ctx, int32_t i_offset);            /* set_pull_back+pull_back */
int rs_solute_mass(rs_ctx* ctx, double* mass);              /* sum of c, interior */
/* Solute ledger (no counterpart in the reference, which has no mass
 * metric). Off at rs_create: the first call must pass reset != 0, which
 * starts it at the current fields (a call with reset == 0 before that
 * returns RS_ERR_STATE); once on it stays on, and the step kernels record
 * the boundary-face fluxes (the FP32-field kernels then run a variant that
 * records them, about 4% slower). It accounts for every change of the
 * interior solute mass since the last call with reset != 0, the context's
 * initialisation or the last upload of c. out (up to 6 values): [0] interior mass now, [1] interior mass
 * at the reset, [2] summed mass change of the pull-backs (the shifted-out
 * columns and the injected cinf, fun.cu:1479-1486), [3] solute gained through
 * the two x boundary faces (x = 1/2 and Nx - 3/2), [4] through the two y
 * boundary faces (the ghost cells copy c and psi from their images but see
 * their own T, fun.cu:1220-1230, 1585-1606, so these faces carry a flux;
 * with frozen T only the x faces do), [5] unaccounted change
 * [0] - [1] - [2] - [3] - [4]: the interior faces cancel, so this is the
 * rounding of the c updates plus what the c >= 1e-6 clamp (fun.cu:1464)
 * added, if it bound anywhere. Runs advanced by rs_batch_step are
 * accounted too. */
int rs_solute_ledger(rs_ctx* ctx, double* out, int32_t n, int32_t reset);
/* The same reduction without a context and without any allocation. No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops). c: device
 * pointer to an Nx x Ny FP64 field (row-major, ghosts included); mass:
 * device pointer receiving the interior sum, bitwise the value
 * rs_solute_mass gives for the same field; stream: a hipStream_t (NULL =
 * the null stream). Enqueues two kernels and returns without waiting. The
 * per-row partial sums live in a static device array of 65536 doubles
 * (512 KiB per GPU, held until the process ends): Ny > 65536 returns
 * RS_ERR_ARG. Every call on one GPU shares that array, so calls that
 * overlap on a GPU race; the first call on a device looks up the array's
 * address there, so a capturing caller makes one call before capturing. */
int rs_solute_mass_noalloc(const double* c, int32_t Nx, int32_t Ny, double* mass, void* stream);
int rs_tip(rs_ctx* ctx, rs_tip_row* tip);                       /* set_tip_output */
int rs_sync(rs_ctx* ctx);

/* Kernel timing with HIP events on the context stream: after
 * rs_timing_enable(ctx, every) (every >= 1; 0 switches timing off) the
 * kernels of every every-th step are bracketed by event pairs, and
 * rs_timing returns their average duration (ms) and count since the enable.
 * which: 0 = fused phi/psi/c step kernel (with the antitrapping pass when
 * on), 1 = coarse T kernel(s). */
int rs_timing_enable(rs_ctx* ctx, int32_t every);
int rs_timing(rs_ctx* ctx, int32_t which, double* avg_ms, int64_t* count);

/* Finite-value sentinel (no counterpart in the reference, where a NaN or
 * Inf propagates silently through every later step): after
 * rs_set_finite_check(ctx, every) with every >= 1, rs_step checks psi, c and
 * (coupled T) the coarse T after every every-th step and stops at the first
 * check that finds a non-finite value, returning RS_ERR_NONFINITE;
 * rs_last_error names the step, the field, the number of such cells and the
 * first one in row-major order. The step's results are kept. Each check
 * reads the fields once and waits for the stream. 0 (default) turns it off.
 * Runs advanced by rs_batch_step are not checked. */
int rs_set_finite_check(rs_ctx* ctx, int32_t every);

/* Measurement helper with no counterpart in the reference: launches a plain
 * FP64 copy kernel (n doubles read, n written, n <= Nx*Ny) reps times on
 * the context stream, from the current c buffer into the other one (which
 * the next step overwrites; the state is kept), so profiler byte counters
 * can be calibrated against a known traffic. Writes the average kernel
 * time (HIP events). */
int rs_calibrate_copy(rs_ctx* ctx, int64_t n, int32_t reps, double* avg_ms);

/* ---- on-disk formats of the reference's driver (host only) --------------
 * The files main() writes, so the reference's post-processing and restart
 * read this build's output unchanged. */
/* save_field (fun.cu:1768-1790): "%g\t" per value, one line per row, every
 * di-th column and dj-th row */
int rs_write_field_txt(const char* path, const double* data, int32_t Nx, int32_t Ny, int32_t di,
                       int32_t dj);
/* one tip.txt row (src/main.cu:2671-2709), appended; the two header lines
 * first when header != 0 */
int rs_write_tip_row(const char* path, const rs_tip_row* t, int32_t header);
/* offset.txt record of a pull-back (fun.cu:710-716), appended */
int rs_append_offset(const char* path, int64_t step, int32_t i_offset, int32_t i_offset_history,
                     int32_t step_pull_back);
/* c_history_temp.txt records of the columns a pull-back shifts out
 * (fun.cu:719-740); cols is Ny x ncols row-major (rs_download_columns);
 * *i_ignore is the reference's skip counter, updated */
int rs_append_history(const char* path, const double* cols, int32_t ncols, int32_t Ny,
                      int32_t i_offset_history, int32_t di, int32_t dj, int32_t* i_ignore);
/* raw FP64 checkpoints (save_check_point / load_check_point,
 * fun.cu:1848-1916): n doubles in memory order */
int rs_write_checkpoint(const char* path, const double* data, int64_t n);
int rs_read_checkpoint(const char* path, double* data, int64_t n);
/* parameters.txt (save_parameters fun.cu:2060-2202, same lines and order);
 * rs_read_parameters fills the restart fields load_parameters
 * (fun.cu:2204-2388) takes from it into k (derived constants on entry) */
int rs_write_parameters(const char* path, const rs_params* p, const rs_constants* k, int32_t i_tip);
int rs_read_parameters(const char* path, const rs_params* p, rs_constants* k);
/* columns [i0, i0+ncols) of RS_FIELD_C or RS_FIELD_PSI into host
 * (Ny x ncols, row-major): the columns a pull-back shifts out, without a
 * whole-field copy */
int rs_download_columns(rs_ctx* ctx, int32_t field, int32_t i0, int32_t ncols, double* host);
/* Exact state file (no counterpart in the reference, whose restart goes
 * through %g text): run state, parities and the device fields in their
 * stored precision; loading into a context created from the same
 * parameters continues the run bit for bit. */
int rs_save_state(rs_ctx* ctx, const char* path);
int rs_load_state(rs_ctx* ctx, const char* path);

/* ---- learned CALPHAD surrogate ------------------------------------------
 * No counterpart in the reference, which evaluates Gl/Gs/Glc/Gsc
 * (fun.cu:1671-1720) exactly at every use. A 2 -> 64 -> 64 -> 4 MLP maps
 * (T, x_Nb = 1 - c) to (f_L, f_S, mu_L, mu_S) without their ideal-mixing
 * terms, which stay exact in FP64. It is trained on the GPU (BF16 MFMA,
 * FP32 accumulate, Adam on FP32 master weights) from samples the exact
 * routine generates, and evaluated inside the RS_PREC_SURROGATE step
 * kernel (MFMA in the activation format act_format). A trust table of
 * 64 x 64 bins over the normalised (T, x_Nb) domain, built by comparing
 * the network (in that format) with the exact routine on a 513 x 513
 * grid, marks the bins whose nondimensional residual
 *   max(|d(f_L - f_S)| / (2 hh v0), |d mu_L| / (h0 v0), |d mu_S| / (h0 v0))
 * stays within residual_tol; cells outside the domain or in an untrusted
 * bin fall back to the exact routine. */
typedef struct rs_surrogate_opts {
    int64_t n_uniform;     /* uniform samples over the domain (2^20)          */
    int64_t n_near;        /* samples within +-0.05 x_Nb of the liquidus and
                              solidus compositions (2^18)                     */
    uint64_t seed;
    int32_t epochs;        /* passes over the sample set                      */
    int32_t batch;         /* samples per optimiser step, a multiple of 64    */
    double T_lo, T_hi;     /* training window, K; T_lo >= T_hi derives it from
                              the run's temperature field                    */
    double lr, lr_final;   /* Adam step size, cosine-annealed to lr_final     */
    double residual_tol;   /* trust bound on the residual above (1e-4)        */
    int32_t n_curve;       /* points of the liquidus/solidus tables           */
    int32_t act_format;    /* inference format of the hidden activations,
                              RS_ACT_* (RS_ACT_FP16_PACKED)                   */
} rs_surrogate_opts;

/* Hidden activations are computed in FP32 and rounded once where they enter
 * the next product (except RS_ACT_FP16_PACKED). Training always runs
 * RS_ACT_BF16 (for RS_ACT_BF16 inference) or RS_ACT_BF16_SPLIT (the others);
 * weights are BF16 values in all. */
enum {
    RS_ACT_BF16 = 0,       /* one BF16 part (8 significant bits), BF16 MFMA   */
    RS_ACT_BF16_SPLIT = 1, /* two BF16 parts hi + lo (~16 bits), 2 MFMAs      */
    RS_ACT_FP16 = 2,       /* FP16 (11 bits), FP16 MFMA for layers 2 and 3    */
    RS_ACT_FP16_PACKED = 3 /* as RS_ACT_FP16 with the activation itself in
                              packed FP16 arithmetic (input, clamp-FMA and
                              product each rounded to FP16): 3-7% faster
                              step. Its rounding noise varies between the
                              validation nodes by up to ~25% of
                              residual_tol, so its trust table is built at
                              0.8 residual_tol (the default format)        */
};

#define RS_SURR_NPARAM 4612  /* W1[64][2] b1[64] W2[64][64] b2[64] W3[4][64] b3[4] */
#define RS_SURR_NMETA 14     /* t_mid, t_half, y_mean[4], y_std[4], T_lo, T_hi, tol,
                                act_format */
#define RS_SURR_NREPORT 12

int rs_surrogate_defaults(rs_surrogate_opts* o);
/* Generates the samples, trains, and builds the trust table. report (up to
 * RS_SURR_NREPORT): [0] final mean squared error (normalised outputs, whole
 * sample set), [1] fraction of trusted bins, [2..5] largest |error| (J/mol)
 * of y0 = f_L - f_S, y1 = f_S - ideal, y2 = mu_L - ideal, y3 = mu_S - mu_L
 * over the validation grid, [6] largest residual over trusted bins,
 * [7] training time (s, HIP events), [8] optimiser steps, [9] T_lo,
 * [10] T_hi, [11] mean squared error before training. */
int rs_surrogate_train(rs_ctx* ctx, const rs_surrogate_opts* o, double* report, int32_t n);
/* Current master weights (FP32, RS_SURR_NPARAM) and meta (RS_SURR_NMETA). */
int rs_surrogate_get(rs_ctx* ctx, float* params, int32_t np, double* meta, int32_t nm);
/* Installs weights and meta, repacks the weight fragments and rebuilds the
 * trust table. */
int rs_surrogate_set(rs_ctx* ctx, const float* params, int32_t np, const double* meta, int32_t nm);
/* Gradient of the mean squared error over n rows (n a multiple of 64,
 * n <= 65536) of normalised inputs X (n x 2) and targets Yn (n x 4) at the
 * current weights, by the training kernels; grad in the parameter layout. */
int rs_surrogate_grad(rs_ctx* ctx, const float* X, const float* Yn, int32_t n, float* grad,
                      double* loss);
/* Network evaluation at n (T, c) pairs (Tc: n x 2): out (n x 5) = f_L, f_S,
 * mu_L, mu_S (J/mol, ideal terms added exactly) and 1.0 where the point is
 * in the domain and its bin trusted, else 0.0. */
int rs_surrogate_eval(rs_ctx* ctx, const double* Tc, int64_t n, double* out);
/* The network alone: the average time (ms, HIP events) of one evaluation of
 * the current network in its inference format over n device-resident inputs
 * (64 <= n <= 2^26) spread over the domain, reps timed launches after one
 * warm-up, with nothing of the stencil around it -- what a fused step pays
 * at least for n cells. Measurement only; no reference counterpart. */
int rs_surrogate_time(rs_ctx* ctx, int64_t n, int32_t reps, double* avg_ms);
/* Counts of the fused kernel since the last reset (out: 5 values): out[0]
 * cell evaluations by the exact routine, out[1] all evaluations, out[2]
 * their ratio, out[3] wavefront rows (62 cells) with no trusted cell, which
 * need not evaluate the network, out[4] all wavefront rows. reset != 0
 * clears the counters afterwards. */
int rs_surrogate_stats(rs_ctx* ctx, double* out, int32_t reset);

/* ---- batched ensemble ----------------------------------------------------
 * No counterpart in the reference (one run per job). Many independent runs
 * -- any grid sizes, undercoolings and pulling speeds, one precision path
 * (RS_PREC_MIXED or RS_PREC_SURROGATE) -- advance together: each step is one
 * persistent launch of the fused step kernel over the work list of all runs
 * (run, column tile, band group), the antitrapping pass of the runs that
 * have it on, one coarse-T launch per sub-step (the largest ratio_dt_dtT of
 * the coupled runs; each run takes its own number) and a state-advance
 * launch, and the step sequence is captured once in a HIP graph and
 * replayed. Pull-backs (fun.cu:448) split the replay into segments and run
 * on the host between them. Results are bitwise those of the runs stepped
 * one by one with rs_step. */
typedef struct rs_batch rs_batch;

int rs_batch_create(const rs_params* params, int32_t n, rs_batch** out);
int rs_batch_destroy(rs_batch* b);
const char* rs_batch_last_error(const rs_batch* b);   /* never NULL */
/* The context of run i (owned by the batch): init, upload/download, tip,
 * mass, constants as for a stand-alone context. */
int rs_batch_run(rs_batch* b, int32_t i, rs_ctx** ctx);
int rs_batch_step(rs_batch* b, int64_t nsteps);
/* Trains one surrogate on the union of the runs' temperature windows and
 * installs it in every run (report as rs_surrogate_train). */
int rs_batch_surrogate_train(rs_batch* b, const rs_surrogate_opts* o, double* report, int32_t n);
/* Autotuner: band height per grid (from 10, 20, 30, 40, 60, 80 rows) and the
 * persistent grid size (1, 2, 4 x resident blocks), each timed over `steps`
 * graph replays on a snapshot of the state, which is restored. out: [0] best
 * ms per step, [1] grid multiplier, then (Nx, band height) per distinct
 * grid. Choices are cached per (precision, grid) for later batches. */
int rs_batch_tune(rs_batch* b, int32_t steps, double* out, int32_t n);
/* out: [0] runs, [1] cells (sum of Nx*Ny), [2] work items, [3] persistent
 * blocks, [4] waves per block, [5] steps per captured graph, [6] ms of the
 * last rs_batch_step (HIP events), [7] its steps, [8] grid multiplier,
 * [9] the step kernel's thermodynamics variant (0 exact table, 1-4 the
 * surrogate formats, 8 packed FP16 with the network in every row). */
int rs_batch_info(rs_batch* b, double* out, int32_t n);

#ifdef __cplusplus
}
#endif
#endif /* RSOLID_H */
