/*
 * rs_oracle.h -- CPU FP64 restatement of the reference time-stepper.
 *
 * TEST INFRASTRUCTURE ONLY. This is the correctness oracle for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker (or as the timed CPU
 * baseline). Nothing in the product path calls it.
 *
 * Every function restates one piece of the reference (amakthel/
 * 2d-rapid-solidification, file fun.cu) and cites the lines it follows. The
 * per-kernel structure is kept deliberately unfused (one loop nest per
 * reference kernel, separate phi and psi arrays, stored T_fine) so that it
 * checks the fused GPU kernels independently.
 *
 * Pinning status (what checks each function against something other than
 * itself) is listed in the tests and in the oracle section of the design
 * notes; functions without such a check are marked "parity unpinned" there.
 */
#ifndef RS_ORACLE_H
#define RS_ORACLE_H

#include <stdint.h>
#include "../include/rsolid.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_state or_state;

/* Parameters, derived constants (fun.cu:1984-2058). */
int or_derive(const rs_params* p, rs_constants* k);
double or_tau0(const rs_params* p);

/* CALPHAD (fun.cu:1628-1720). c is the Ti mole fraction. */
double or_G_Ti_BCC(double T);
double or_G_Ti_Liq(double T);
double or_G_Nb_BCC(double T);
double or_G_Nb_Liq(double T);
double or_Gl(const rs_params* p, double T, double c);
double or_Gs(const rs_params* p, double T, double c);
double or_Glc(const rs_params* p, double T, double c);
double or_Gsc(const rs_params* p, double T, double c);

/* set_orientation (fun.cu:750-771): the 9-entry rotation of Euler angles
 * (degrees) with the float-precision PI of fun.cu:59. */
void or_orientation(double* o, double alpha, double beta, double gamma);

/* get_anisotropy (fun.cu:1030-1139) on a 3x3 phi neighbourhood given as
 * ph[3*dj + di] with di,dj in {0,1,2} (centre = ph[4]); orientation is the
 * 9-entry rotation (fun.cu:750-771). Returns the anisotropic term, writes
 * a_s and a_k. */
double or_anisotropy(const rs_params* p, const double* ph, const double* orient,
                     double* as, double* ak);

/* Whole-run state. */
or_state* or_create(const rs_params* p);
void or_destroy(or_state* s);
int or_init_reference(or_state* s);
int or_init_config(or_state* s, uint64_t seed);
/* field ids as RS_FIELD_*; RS_FIELD_PHI is settable here (the oracle keeps
 * phi and psi separately like the reference). */
int or_get(or_state* s, int field, double* out, int64_t n);
int or_put(or_state* s, int field, const double* in, int64_t n);
int or_get_constants(or_state* s, rs_constants* k);
int or_set_constants(or_state* s, const rs_constants* k);

/* One pass of the loop body fun.cu:448-509 (no output). */
int or_step(or_state* s, int64_t nsteps);
int or_tip_column(or_state* s);                    /* get_itip_all_grain */
int or_pull_back(or_state* s, int i_offset);       /* set_pull_back + pull_back */
double or_solute_mass(or_state* s);                /* FP64 sum of c, interior */

/* Tip fit (fun.cu:2628-2954) on a host psi array Nx*Ny. Writes
 * i, j, x, y, R (R in nm, clamped like fun.cu:2927-2931). Returns the
 * fit_parabola status. */
int or_precise_tip(const rs_params* p, const double* psi, int32_t* ti,
                   int32_t* tj, double* tx, double* ty, double* tR);
/* set_tip_output (fun.cu:3024-3072): fills an rs_tip row and advances
 * x_tip_history like the reference. */
int or_tip(or_state* s, rs_tip_row* tip);

/* Common tangent of the liquid and bcc Gibbs energies at temperature T:
 * returns 0 and writes (c_liquid, c_solid) on convergence. */
int or_common_tangent(const rs_params* p, double T, double* cl, double* cs);
/* antitrapping face flux of the trapping extension (rs_params.at_coef):
 * cells P, Q along the face normal, old/new phi, old c, the four tangential
 * neighbours (t1 + t2) - (t3 + t4) across the face; vdn = V_D in W/tau0 */
double or_at_flux(const rs_params* p, double vdn, double dt, double pP, double pQ, double nP,
                  double nQ, double cP, double cQ, double t1, double t2, double t3, double t4);

/* interpolate_temperature (fun.cu:1580-1608) of the current coarse T into
 * T_fine with the given i_left_end (read back as RS_FIELD_TFINE). */
int or_interpolate_at(or_state* s, int i_left_end);
/* integrate_phi (fun.cu:1494-1535): latent-heat block sums 0.5 * sum of
 * (psi_next - psi) for the given Nx*Ny arrays, written to out
 * (NxT_all*NyT_all). Replaces the state's psi and psi_next. */
int or_integrate_phi_at(or_state* s, const double* psi, const double* psi_next, int i_left_end,
                        double* out);

/* splitmix64 stream used by the seeded configuration initialiser. */
uint64_t or_splitmix64(uint64_t* state);

int or_set_threads(int n);

/* solute ledger: out7 = mass, mass at the reset, summed pull-back change,
 * x and y boundary face gains, clamped cells, unaccounted change;
 * reset != 0 restarts it at the current fields */
int or_solute_ledger(or_state* s, double* out7, int reset);

#ifdef __cplusplus
}
#endif
#endif
