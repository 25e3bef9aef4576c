/*
 * tp_oracle.h -- CPU FP64 restatement of the 3D three-phase peritectic
 * Ti-Ag stepper (multi/peritectic_multiGPU/TiAg.cu) on one z-slab.
 *
 * TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, as the checker.
 */
#ifndef TP_ORACLE_H
#define TP_ORACLE_H

#include <stdint.h>

#include "../include/rs_peritectic.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tpo_state tpo_state;

/* g, gp1, gp2 (TiAg.cu:71-73), computeMu for one cell (TiAg.cu:446-456),
 * and the bulk terms fphi_1..3 of diffuse (TiAg.cu:498-514) */
double tpo_g(double pi, double pj, double pk);
double tpo_gp1(double pi, double pj, double pk);
double tpo_gp2(double pi, double pj, double pk);
double tpo_mu(const tp_params* p, double p1, double p2, double c);
void tpo_bulk_force(const tp_params* p, double p1, double p2, double muu, double* out3);

tpo_state* tpo_create(const tp_params* p);
void tpo_destroy(tpo_state* s);
int tpo_init_reference(tpo_state* s);                 /* initial(), TiAg.cu:411-444 */
int tpo_get(tpo_state* s, int field, double* out, int64_t n);
int tpo_put(tpo_state* s, int field, const double* in, int64_t n);
int tpo_step(tpo_state* s, int64_t nsteps);            /* TiAg.cu:288-345 */
int64_t tpo_get_step(tpo_state* s);
int tpo_report(tpo_state* s, int kc, double* out2);    /* TiAg.cu:229-262 */

#ifdef __cplusplus
}
#endif
#endif
