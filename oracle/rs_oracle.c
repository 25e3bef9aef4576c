/*
 * rs_oracle.c -- CPU FP64 restatement of the reference time-stepper.
 *
 * TEST INFRASTRUCTURE ONLY (see rs_oracle.h). Compiled with
 * -ffp-contract=off so every expression rounds exactly as written; the
 * expression order of each formula follows the reference line it cites so
 * that the restatement and the reference round alike wherever libm agrees.
 *
 * Reference = amakthel/2d-rapid-solidification, fun.cu (source of record).
 */
#include "rs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_REF 3.1415927 /* fun.cu:59 (float-precision literal, kept) */

struct or_state {
    rs_params p;
    rs_constants k;
    int64_t n;      /* Nx*Ny */
    int64_t nT;     /* NxT_all*NyT_all */
    double *c, *phi, *psi, *c_next, *phi_next, *psi_next, *T_fine;
    double *T, *T_next, *dphidt;
    double orient[9];
    /* solute ledger: per-row x-edge face gains (left Ny, right Ny) and
     * clamp counts, the summed pull-back mass change, the reference mass */
    double* led;
    int* led_clamp;
    double led_pull, led_m0;
};

int or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* derived physical constants (fun.cu:59-105)                          */
/* ------------------------------------------------------------------ */
double or_tau0(const rs_params* p) {
    double sw = p->S * p->W0; /* pow2(SS*W0)/(Gamma*muk0), fun.cu:69 */
    return (sw * sw) / (p->Gamma * p->muk0);
}
static double or_a1(void) { return 2.0 * sqrt(2.0) / 3.0; } /* fun.cu:58 */
static double or_hh(const rs_params* p) {                   /* fun.cu:97-104 */
    double lh2 = p->latentHeat * p->num_moles;
    double st = p->Gamma * 1e-9 * lh2 / p->Tmelt;
    return st / (or_a1() * p->S * p->W0 * 1.0e-9);
}
static double or_h0(const rs_params* p) { return p->RR * p->Tmelt / p->v0; } /* fun.cu:105 */
static double or_Lcp(const rs_params* p) { return p->latentHeat / p->cp; }    /* fun.cu:95 */

/* Vp_ramp (fun.cu:1974-1981), t in ns */
static double or_vp(const rs_params* p, double t) {
    if (p->if_Vp_ramp) {
        t = t * 0.001;
        return p->Vp + p->Vp_ramp_rate * t;
    }
    return p->Vp;
}

/* set_parameters (fun.cu:1984-2058) */
int or_derive(const rs_params* p, rs_constants* k) {
    if (!p || !k) return RS_ERR_ARG;
    memset(k, 0, sizeof(*k));
    double tau0 = or_tau0(p);
    double sw = p->S * p->W0;
    double DD = p->Dl * tau0 / (sw * sw);
    k->initial_step = 0;
    k->step = 0;
    k->dt = p->dt_coef * p->dx * p->dx / (4.0 * DD);
    if (p->dt_cap && k->dt > p->dt_coef * p->dx * p->dx / 4.0)
        k->dt = p->dt_coef * p->dx * p->dx / 4.0;
    k->i_ignore = 0;
    k->i_offset = 0;
    k->i_offset_history_initial = 0;
    k->i_offset_history = 0;
    k->x_tip_history = p->i_tip_target * p->dx;
    k->if_begin_tip = 1;
    k->if_begin_field = 1;
    double dt = k->dt;
    k->total_step = (int64_t)(p->total_time / (tau0 * dt));
    k->step_fields = (int32_t)(p->total_time / p->num_fields_output / (tau0 * dt));
    k->step_tip = (int32_t)(p->total_time / p->num_tip_output / (tau0 * dt));
    k->step_check_point =
        (int32_t)(p->total_time / p->num_check_point_output / (tau0 * dt));
    if (p->if_Vp_ramp)
        k->step_pull_back = (int32_t)(p->S * p->W0 * p->dx / (1.0 * dt * tau0));
    else
        k->step_pull_back = (int32_t)(p->S * p->W0 * p->dx / (p->Vp * dt * tau0));
    if (k->step_pull_back < 1) k->step_pull_back = 1;
    if (k->step_fields < 1) k->step_fields = 1;
    if (k->step_tip < 1) k->step_tip = 1;
    if (k->step_check_point < 1) k->step_check_point = 1;
    k->delta_initial = p->delta_initial;
    k->delta = k->delta_initial;
    k->l_offset_Vp_initial = 0;
    k->l_offset_Vp = 0;

    /* coarse temperature grid (fun.cu:2025-2053) */
    int r = p->ratio_dxT_dx;
    double swr = p->S * p->W0 * r;
    double DDT = p->DT * tau0 / (swr * swr);
    double dtT = p->dtT_coef * p->dx * p->dx / (4.0 * DDT);
    k->ratio_dt_dtT = 1;
    if (dtT < k->dt) k->ratio_dt_dtT = (int32_t)ceil(k->dt / dtT);
    k->dtT = k->dt / k->ratio_dt_dtT;
    double cell = p->dx * p->S * p->W0 * r;
    k->length_sample = ceil(p->length_sample_um * 1000 / cell) * cell;
    k->NxT = (int32_t)(k->length_sample / cell + 2);
    k->NxT_all = (int32_t)(ceil(1.0 * k->NxT / 64) * 64);
    k->NyT = ((p->Ny - 2) / r) + 2;
    k->NyT_all = (int32_t)(ceil(1.0 * k->NyT / 4) * 4);
    k->Tcold = p->TCOLD;
    k->Thot = k->Tcold + p->GG * k->length_sample;
    k->i_left_end_initial =
        (int32_t)((p->Tmelt - p->deltaT0 - k->Tcold) / (p->GG * p->dx * p->S * p->W0) -
                  p->i_tip_target);
    k->i_left_end = k->i_left_end_initial;
    k->i_left_end_interp = k->i_left_end;
    return RS_OK;
}

/* ------------------------------------------------------------------ */
/* CALPHAD (fun.cu:1628-1720): SGTE unaries, regular solution in c=x_Ti */
/* ------------------------------------------------------------------ */
double or_G_Ti_BCC(double T) { /* fun.cu:1628-1636 */
    if (T >= 1155.0 && T < 1941.0)
        return 6667.385 + 105.366379 * T - 22.3771 * T * log(T) + 1.21707e-3 * (T * T) -
               0.84534e-6 * ((T * T) * T) - 2002750.0 / T;
    else if (T >= 1941.0)
        return 26483.26 - 182.426471 * T + 19.0900905 * T * log(T) -
               22.00832e-3 * (T * T) + 1.228863e-6 * ((T * T) * T) + 1400501.0 / T;
    return NAN;
}
double or_G_Ti_Liq(double T) { /* fun.cu:1638-1646 */
    if (T >= 1300.0 && T < 1941.0)
        return 369519.198 - 2554.0225 * T + 342.059267 * T * log(T) -
               163.409355e-3 * (T * T) + 12.457117e-6 * ((T * T) * T) - 67034516.0 / T;
    else if (T >= 1941.0 && T < 4000.0)
        return -19887.066 + 298.7367 * T - 46.29 * T * log(T);
    return NAN;
}
double or_G_Nb_BCC(double T) { /* fun.cu:1648-1656 */
    if (T >= 298.15 && T < 2750.0)
        return -8519.353 + 142.045475 * T - 26.4711 * T * log(T) + 0.203475e-3 * (T * T) -
               0.35012e-6 * ((T * T) * T) + 93399.0 / T;
    else if (T >= 2750.0) {
        double t3 = (T * T) * T;
        return -37669.3 + 271.720843 * T - 41.77 * T * log(T) + 1528.238e29 / ((t3 * t3) * t3);
    }
    return NAN;
}
double or_G_Nb_Liq(double T) { /* fun.cu:1658-1666 */
    if (T >= 298.15 && T < 2750.0) {
        double t2 = T * T;
        return 21262.202 + 131.229057 * T - 26.4711 * T * log(T) + 0.203475e-3 * t2 -
               0.35012e-6 * (t2 * T) + 93399.0 / T - 306.098e-25 * (((t2 * t2) * t2) * T);
    } else if (T >= 2750.0)
        return -7499.398 + 260.756148 * T - 41.77 * T * log(T);
    return NAN;
}

/* ideal-mixing part RT[(1-c)ln(1-c) + c ln c] with the c->0 / c->1
 * branches of fun.cu:1673-1680 */
static double or_ideal(const rs_params* p, double T, double c) {
    double onemc = 1.0 - c;
    if (c < 1.e-8) return p->RR * T * (onemc * log(onemc));
    if (onemc < 1.e-8) return p->RR * T * (c * log(c));
    return p->RR * T * (onemc * log(onemc) + c * log(c));
}
/* Redlich-Kister excess c(1-c)[L0 + L1(1-2c)]; L1 = 0 is fun.cu:1682 */
static double or_excess(double c, double L0, double L1) {
    double onemc = 1.0 - c;
    return c * onemc * (L0 + L1 * (1.0 - 2.0 * c));
}
static double or_excess_dc(double c, double L0, double L1) {
    /* d/dc of or_excess: L0(1-2c) + L1(1-6c+6c^2); L1 = 0 is fun.cu:1694 */
    return L0 * (1.0 - 2.0 * c) + L1 * (1.0 - 6.0 * c + 6.0 * (c * c));
}
double or_Gl(const rs_params* p, double T, double c) { /* fun.cu:1671-1684 */
    double onemc = 1.0 - c;
    return or_ideal(p, T, c) + onemc * or_G_Nb_Liq(T) + c * or_G_Ti_Liq(T) +
           or_excess(c, p->L0_liq, p->L1_liq);
}
double or_Gs(const rs_params* p, double T, double c) { /* fun.cu:1686-1699 */
    double onemc = 1.0 - c;
    return or_ideal(p, T, c) + onemc * or_G_Nb_BCC(T) + c * or_G_Ti_BCC(T) +
           or_excess(c, p->L0_bcc, p->L1_bcc);
}
static double or_logratio(const rs_params* p, double T, double c) {
    double onemc = 1.0 - c;
    double lp = 0.;
    if (c > 1.e-8 && onemc > 1.e-8) lp = p->RR * T * log(c / onemc);
    return lp;
}
double or_Glc(const rs_params* p, double T, double c) { /* fun.cu:1701-1710 */
    double tl = or_excess_dc(c, p->L0_liq, p->L1_liq);
    return or_G_Ti_Liq(T) - or_G_Nb_Liq(T) + tl + or_logratio(p, T, c);
}
double or_Gsc(const rs_params* p, double T, double c) { /* fun.cu:1712-1720 */
    double tl = or_excess_dc(c, p->L0_bcc, p->L1_bcc);
    return or_G_Ti_BCC(T) - or_G_Nb_BCC(T) + tl + or_logratio(p, T, c);
}

/* ------------------------------------------------------------------ */
/* interpolation polynomials (fun.cu:189-195)                           */
/* ------------------------------------------------------------------ */
static inline double or_qq(const rs_params* p, double ps) {
    double om = 1.0 - ps;
    double q = 0.5 * p->AA * om - 0.25 * (p->AA - 1) * (om * om);
    return q * (1.0 - p->q_solid) + p->q_solid;
}
static inline double or_gg(double ps) {
    double p2 = ps * ps;
    return 1.875 * ps - 1.25 * (p2 * ps) + 0.375 * ((p2 * p2) * ps);
}
static inline double or_dgg(double ps) {
    double a = 1.0 - ps * ps;
    return 1.875 * (a * a);
}

/* ------------------------------------------------------------------ */
/* anisotropy (fun.cu:1030-1139, if_1_eps branch)                        */
/* ------------------------------------------------------------------ */
double or_anisotropy(const rs_params* p, const double* ph, const double* o, double* as,
                     double* ak) {
    const double dx = p->dx;
    double pmm = ph[0], p0m = ph[1], p1m = ph[2];
    double pm0 = ph[3], p00 = ph[4], p10 = ph[5];
    double pm1 = ph[6], p01 = ph[7], p11 = ph[8];

    double gx = (p10 - pm0) / (2.0 * dx);
    double gy = (p01 - p0m) / (2.0 * dx);
    double ag = sqrt(gx * gx + gy * gy);
    double gxx = (p10 - 2.0 * p00 + pm0) / (dx * dx);
    double gyy = (p01 - 2.0 * p00 + p0m) / (dx * dx);
    double gxy = (p11 - pm1 - p1m + pmm) / (4. * dx * dx);
    double lap = p->lap_w_axis * (gxx + gyy);
    lap += p->lap_w_diag * (p11 + pm1 + p1m + pmm - 4.0 * p00) / (2.0 * dx * dx);

    double thx, thy, ct, st;
    if (ag > 1.0e-15) {
        double ag2 = ag * ag;
        thx = (gx * gxy - gy * gxx) / ag2;
        thy = (gx * gyy - gy * gxy) / ag2;
        ct = gx / ag;
        st = gy / ag;
    } else {
        thx = thy = ct = st = 0.0;
    }
    /* rotated normal n' = R (cos, sin) and its theta derivatives */
    double n[3], d1n[3];
    for (int r = 0; r < 3; r++) {
        double a = o[3 * r + 0], b = o[3 * r + 1];
        n[r] = a * ct + b * st;
        d1n[r] = -a * st + b * ct;
    }
    double np4 = 0.0, d1 = 0.0, d2 = 0.0;
    for (int r = 0; r < 3; r++) {
        double n2 = n[r] * n[r];
        double n3 = n2 * n[r];
        np4 = (r == 0) ? (n2 * n2) : np4 + (n2 * n2);
        d1 = (r == 0) ? n3 * d1n[r] : d1 + n3 * d1n[r];
        double nd = n[r] * d1n[r];
        d2 += 3.0 * (nd * nd) + n3 * (-n[r]);
    }
    d1 = 4.0 * d1;
    d2 *= 4;
    *ak = 1.0 - 3.0 * p->epk1 + 4.0 * p->epk1 * np4;
    *as = 1.0 - 3.0 * p->eps1 + 4.0 * p->eps1 * np4;
    double a = *as;
    double da1 = 4.0 * p->eps1 * d1;
    double da2 = 4.0 * p->eps1 * d2;
    return (a * a) * lap + 2.0 * a * da1 * (thx * gx + thy * gy) +
           (da2 * a + da1 * da1) * (thy * gx - thx * gy);
}

/* ------------------------------------------------------------------ */
/* state                                                               */
/* ------------------------------------------------------------------ */
or_state* or_create(const rs_params* p) {
    if (!p || p->Nx < 4 || p->Ny < 4) return NULL;
    or_state* s = (or_state*)calloc(1, sizeof(or_state));
    if (!s) return NULL;
    s->p = *p;
    or_derive(p, &s->k);
    s->n = (int64_t)p->Nx * p->Ny;
    s->nT = (int64_t)s->k.NxT_all * s->k.NyT_all;
    double** f[] = {&s->c, &s->phi, &s->psi, &s->c_next, &s->phi_next, &s->psi_next, &s->T_fine};
    for (int q = 0; q < 7; q++) {
        *f[q] = (double*)calloc((size_t)s->n, sizeof(double));
        if (!*f[q]) { or_destroy(s); return NULL; }
    }
    /* one spare coarse row: the bilinear read of row jT+1 reaches row NyT_all
     * (with weight 0) when NyT is a multiple of 4, e.g. Ny = 1024 */
    double** g[] = {&s->T, &s->T_next, &s->dphidt};
    for (int q = 0; q < 3; q++) {
        *g[q] = (double*)calloc((size_t)(s->nT + s->k.NxT_all), sizeof(double));
        if (!*g[q]) { or_destroy(s); return NULL; }
    }
    for (int q = 0; q < 9; q++) s->orient[q] = (q % 4 == 0) ? 1.0 : 0.0;
    s->led = (double*)calloc((size_t)2 * (p->Ny + p->Nx), sizeof(double));
    s->led_clamp = (int*)calloc((size_t)p->Ny, sizeof(int));
    if (!s->led || !s->led_clamp) { or_destroy(s); return NULL; }
    return s;
}

void or_destroy(or_state* s) {
    if (!s) return;
    free(s->c); free(s->phi); free(s->psi); free(s->c_next); free(s->phi_next);
    free(s->psi_next); free(s->T_fine); free(s->T); free(s->T_next); free(s->dphidt);
    free(s->led); free(s->led_clamp);
    free(s);
}

#define POS(s, i, j) ((int64_t)(j) * (s)->p.Nx + (i))
#define PST(s, i, j) ((int64_t)(j) * (s)->k.NxT_all + (i))

/* set_orientation (fun.cu:750-771) */
void or_orientation(double* o, double al, double be, double ga) {
    al = al * PI_REF / 180.0;
    be = be * PI_REF / 180.0;
    ga = ga * PI_REF / 180.0;
    o[0] = cos(be) * cos(ga);
    o[1] = sin(al) * sin(be) * cos(ga) - cos(al) * sin(ga);
    o[2] = cos(al) * sin(be) * cos(ga) + sin(al) * sin(ga);
    o[3] = cos(be) * sin(ga);
    o[4] = sin(al) * sin(be) * sin(ga) + cos(al) * cos(ga);
    o[5] = cos(al) * sin(be) * sin(ga) - sin(al) * cos(ga);
    o[6] = -sin(be);
    o[7] = sin(al) * cos(be);
    o[8] = cos(al) * cos(be);
}

/* interpolate_temperature (fun.cu:1580-1608) into T_fine */
static void or_interpolate(or_state* s, int i_left_end) {
    const int Nx = s->p.Nx, Ny = s->p.Ny, r = s->p.ratio_dxT_dx;
    const double dxT = s->p.dx * r;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) {
            int iT = (i - 1 + i_left_end) / r;
            int jT = (int)((j - 1 + dxT) / r);
            double x = (i - 1.0 + i_left_end) / r - iT;
            double y = (j - 1.0 + dxT) / r - jT;
            double f00 = s->T[PST(s, iT, jT)];
            double f10 = s->T[PST(s, iT + 1, jT)];
            double f01 = s->T[PST(s, iT, jT + 1)];
            double f11 = s->T[PST(s, iT + 1, jT + 1)];
            double a10 = f10 - f00, a01 = f01 - f00, a11 = f11 - f10 - f01 + f00;
            s->T_fine[POS(s, i, j)] = f00 + a10 * x + a01 * y + a11 * x * y;
        }
    }
    s->k.i_left_end_interp = i_left_end;
}

/* initial_temperature (fun.cu:1023-1028) over the padded coarse grid and
 * the spare row, written into both ping-pong buffers */
static void or_init_T(or_state* s) {
    const int r = s->p.ratio_dxT_dx;
    for (int jT = 0; jT <= s->k.NyT_all; jT++)
        for (int iT = 0; iT < s->k.NxT_all; iT++) {
            double v = s->k.Tcold + (iT - 1) * s->p.S * s->p.W0 * s->p.dx * r * s->p.GG;
            s->T[PST(s, iT, jT)] = v;
            s->T_next[PST(s, iT, jT)] = v;
        }
}

static int or_check_window(or_state* s, int i_left_end) {
    if (!s->p.ifTFC) return RS_OK;
    const int r = s->p.ratio_dxT_dx;
    if (i_left_end - 1 < 0) return RS_ERR_DOMAIN;
    if ((s->p.Nx - 2 + i_left_end) / r + 1 > s->k.NxT - 1) return RS_ERR_DOMAIN;
    return RS_OK;
}

static void or_ledger_reset(or_state* s) {
    memset(s->led, 0, sizeof(double) * 2 * (size_t)(s->p.Ny + s->p.Nx));
    memset(s->led_clamp, 0, sizeof(int) * (size_t)s->p.Ny);
    s->led_pull = 0.0;
    s->led_m0 = or_solute_mass(s);
}

/* initial (fun.cu:983-1021) */
int or_init_reference(or_state* s) {
    const int Nx = s->p.Nx, Ny = s->p.Ny;
    const double dx = s->p.dx;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Ny; j++) {
        double itip = s->p.i_tip_target + 0.5 * sin(2.0 * PI_REF * j / Ny);
        for (int i = 0; i < Nx; i++) {
            double ph = -tanh((i - itip) * dx / sqrt(2.0));
            double cc = 0.3184 * tanh(((i - itip) * dx - 0.08) / 0.67) + 0.6816;
            s->phi[POS(s, i, j)] = ph;
            s->c[POS(s, i, j)] = cc * s->p.cinf;
            s->psi[POS(s, i, j)] = 0.0 + ph;
        }
    }
    or_orientation(s->orient, s->p.alpha0, s->p.beta0, s->p.gamma0);
    or_ledger_reset(s);
    if (s->p.ifTFC) {
        or_init_T(s);
        or_interpolate(s, s->k.i_left_end);
        return or_check_window(s, s->k.i_left_end);
    }
    return RS_OK;
}

uint64_t or_splitmix64(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* seeded planar front: x0_j = i_tip_target + U[-1,1) per row j (interior
 * rows drawn in order j = 1..Ny-2; ghost rows are their periodic images) */
int or_init_config(or_state* s, uint64_t seed) {
    const int Nx = s->p.Nx, Ny = s->p.Ny;
    const double dx = s->p.dx;
    double* x0 = (double*)malloc(sizeof(double) * Ny);
    if (!x0) return RS_ERR_NOMEM;
    uint64_t st = seed;
    for (int j = 1; j <= Ny - 2; j++) {
        double u = (double)(or_splitmix64(&st) >> 11) * 0x1.0p-53;
        x0[j] = s->p.i_tip_target + (2.0 * u - 1.0);
    }
    x0[0] = x0[Ny - 2];
    x0[Ny - 1] = x0[1];
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) {
            double ph = -tanh((i - x0[j]) * dx / sqrt(2.0));
            double c = 0.5 * (1.0 + ph) * s->p.c_sol0 + 0.5 * (1.0 - ph) * s->p.c_liq0;
            s->phi[POS(s, i, j)] = ph;
            s->psi[POS(s, i, j)] = ph;
            s->c[POS(s, i, j)] = c;
        }
    }
    free(x0);
    or_orientation(s->orient, s->p.alpha0, s->p.beta0, s->p.gamma0);
    or_ledger_reset(s);
    if (s->p.ifTFC) {
        or_init_T(s);
        or_interpolate(s, s->k.i_left_end);
        return or_check_window(s, s->k.i_left_end);
    }
    return RS_OK;
}

static double* or_field(or_state* s, int field, int64_t* n) {
    switch (field) {
    case RS_FIELD_C: *n = s->n; return s->c;
    case RS_FIELD_PSI: *n = s->n; return s->psi;
    case RS_FIELD_PHI: *n = s->n; return s->phi;
    case RS_FIELD_T: *n = s->nT; return s->T;
    case RS_FIELD_TFINE: *n = s->n; return s->T_fine;
    }
    return NULL;
}
int or_get(or_state* s, int field, double* out, int64_t n) {
    int64_t m;
    double* f = or_field(s, field, &m);
    if (!f || n != m || !out) return RS_ERR_ARG;
    memcpy(out, f, sizeof(double) * (size_t)n);
    return RS_OK;
}
int or_put(or_state* s, int field, const double* in, int64_t n) {
    int64_t m;
    double* f = or_field(s, field, &m);
    if (!f || n != m || !in) return RS_ERR_ARG;
    memcpy(f, in, sizeof(double) * (size_t)n);
    if (field == RS_FIELD_T) memcpy(s->T_next, in, sizeof(double) * (size_t)n);
    if (field == RS_FIELD_C) or_ledger_reset(s);
    return RS_OK;
}
int or_get_constants(or_state* s, rs_constants* k) { *k = s->k; return RS_OK; }
int or_set_constants(or_state* s, const rs_constants* k) { s->k = *k; return RS_OK; }

/* get_delta (fun.cu:1610-1619), frozen-temperature mode */
static inline double or_delta(const or_state* s, int i) {
    const rs_params* p = &s->p;
    return s->k.delta_initial +
           p->GG * (i - p->i_tip_target + s->k.i_offset_history - s->k.i_offset_history_initial) *
               p->dx * p->S * p->W0 / p->deltaT0 -
           p->GG * (s->k.l_offset_Vp - s->k.l_offset_Vp_initial) * p->S * p->W0 / p->deltaT0;
}
static inline double or_T_at(const or_state* s, int i, int j) {
    if (s->p.ifTFC) return s->T_fine[POS(s, i, j)];
    return or_delta(s, i) * s->p.deltaT0 + s->p.Tmelt;
}

/* ghost copies as compute_boundary_individual (fun.cu:1220-1230) leaves
 * them when every write is visible: x no-flux copy, y periodic copy */
static void or_ghosts(or_state* s, double* f) {
    const int Nx = s->p.Nx, Ny = s->p.Ny;
    for (int j = 1; j <= Ny - 2; j++) {
        f[POS(s, 0, j)] = f[POS(s, 1, j)];
        f[POS(s, Nx - 1, j)] = f[POS(s, Nx - 2, j)];
    }
    for (int i = 0; i < Nx; i++) {
        f[POS(s, i, 0)] = f[POS(s, i, Ny - 2)];
        f[POS(s, i, Ny - 1)] = f[POS(s, i, 1)];
    }
}

/* compute_phi (fun.cu:1143-1218) */
static void or_compute_phi(or_state* s) {
    const rs_params* p = &s->p;
    const int Nx = p->Nx, Ny = p->Ny;
    const double hh = or_hh(p);
    const double dt = s->k.dt;
#pragma omp parallel for schedule(static)
    for (int j = 1; j < Ny - 1; j++) {
        double ph[9];
        for (int i = 1; i < Nx - 1; i++) {
            for (int b = 0; b < 3; b++)
                for (int a = 0; a < 3; a++) ph[3 * b + a] = s->phi[POS(s, i + a - 1, j + b - 1)];
            double c00 = s->c[POS(s, i, j)];
            double phi00 = ph[4];
            double as, ak;
            double an = or_anisotropy(p, ph, s->orient, &as, &ak);
            double term = phi00 - (phi00 * phi00) * phi00 + an;
            double T00 = or_T_at(s, i, j);
            term += (or_Gl(p, T00, c00) - or_Gs(p, T00, c00)) * or_dgg(phi00) / (2.0 * hh * p->v0);
            /* tau = 1 + 1e4*(..)*sum_phi with sum_phi = 0 for one orientation */
            const double tau = 1.0;
            s->phi_next[POS(s, i, j)] = phi00 + dt * ak / ((as * as) * tau) * term;
        }
    }
}

/* compute_psi (fun.cu:1232-1273), one orientation */
static void or_compute_psi(or_state* s) {
    const int Nx = s->p.Nx, Ny = s->p.Ny;
#pragma omp parallel for schedule(static)
    for (int j = 1; j < Ny - 1; j++) {
        for (int i = 1; i < Nx - 1; i++) {
            int64_t q = POS(s, i, j);
            double ps = 0.0 + s->phi_next[q];
            if (ps > 1.0) {
                double ph = s->phi_next[q];
                double n_above = (ph > -1) ? 1.0 : 0.0;
                ph -= (ps - 1.0) * ((ph > -1) ? 1.0 : 0.0) / n_above;
                s->phi_next[q] = ph;
                ps = 0.0 + ph;
            }
            s->psi_next[q] = ps;
            double ph = s->phi_next[q];
            ph = fmax(ph, -1.0);
            ph = fmin(ph, 1.0);
            s->phi_next[q] = ph;
        }
    }
    or_ghosts(s, s->psi_next);
    or_ghosts(s, s->phi_next);
}

/* compute_c_CALPHAD (fun.cu:1365-1467) */
static void or_compute_c(or_state* s) {
    const rs_params* p = &s->p;
    const int Nx = p->Nx, Ny = p->Ny;
    const double denom = 1.0 / (or_h0(p) * p->v0);
    const double dt = s->k.dt;
#pragma omp parallel for schedule(static)
    for (int j = 1; j < Ny - 1; j++) {
        for (int i = 1; i < Nx - 1; i++) {
            double cc[9], pp[9];
            for (int b = 0; b < 3; b++)
                for (int a = 0; a < 3; a++) {
                    cc[3 * b + a] = s->c[POS(s, i + a - 1, j + b - 1)];
                    pp[3 * b + a] = s->psi[POS(s, i + a - 1, j + b - 1)];
                }
            /* index: 0 mm, 1 0m, 2 1m, 3 m0, 4 00, 5 10, 6 m1, 7 01, 8 11 */
            double T00 = or_T_at(s, i, j), T10 = or_T_at(s, i + 1, j), Tm0 = or_T_at(s, i - 1, j);
            double T01, T0m;
            if (p->ifTFC) {
                T01 = or_T_at(s, i, j + 1);
                T0m = or_T_at(s, i, j - 1);
            } else {
                T01 = T00; /* fun.cu:1419-1420 */
                T0m = T00;
            }
            double Sv[9];
            for (int q = 0; q < 9; q++) Sv[q] = or_qq(p, pp[q]) * cc[q] * (1.0 - cc[q]);
            double S00 = Sv[4], S10 = Sv[5], S01 = Sv[7], Sm0 = Sv[3], S0m = Sv[1];
            double S11 = Sv[8], Sm1 = Sv[6], S1m = Sv[2], Smm = Sv[0];
            double Shh = 0.25 * (S11 + S10 + S01 + S00);
            double Shn = 0.25 * (S10 + S00 + S1m + S0m);
            double Snh = 0.25 * (Sm1 + S01 + Sm0 + S00);
            double Snn = 0.25 * (Sm0 + Smm + S0m + S00);
            double g00 = or_gg(pp[4]), g10 = or_gg(pp[5]), g01 = or_gg(pp[7]);
            double gm0 = or_gg(pp[3]), g0m = or_gg(pp[1]);
            double f00 = 0.5 * (1.0 + g00) * or_Gsc(p, T00, cc[4]) +
                         0.5 * (1.0 - g00) * or_Glc(p, T00, cc[4]);
            double f10 = 0.5 * (1.0 + g10) * or_Gsc(p, T10, cc[5]) +
                         0.5 * (1.0 - g10) * or_Glc(p, T10, cc[5]);
            double f01 = 0.5 * (1.0 + g01) * or_Gsc(p, T01, cc[7]) +
                         0.5 * (1.0 - g01) * or_Glc(p, T01, cc[7]);
            double fm0 = 0.5 * (1.0 + gm0) * or_Gsc(p, Tm0, cc[3]) +
                         0.5 * (1.0 - gm0) * or_Glc(p, Tm0, cc[3]);
            double f0m = 0.5 * (1.0 + g0m) * or_Gsc(p, T0m, cc[1]) +
                         0.5 * (1.0 - g0m) * or_Glc(p, T0m, cc[1]);
            double F10 = 0.25 * (S10 + S00 + Shh + Shn) * (f10 - f00) * denom;
            double F01 = 0.25 * (S01 + S00 + Shh + Snh) * (f01 - f00) * denom;
            double Fm0 = 0.25 * (S00 + Sm0 + Snh + Snn) * (fm0 - f00) * denom;
            double F0m = 0.25 * (S00 + S0m + Shn + Snn) * (f0m - f00) * denom;
            double dc = dt * p->Dl / (p->Gamma * p->muk0) * (F10 + F01 + Fm0 + F0m) / (p->dx * p->dx);
            s->c_next[POS(s, i, j)] = fmax(cc[4] + dc, 1.0e-6);
            /* ledger: the flux through the x boundary faces, the clamp */
            if (i == 1) s->led[j] += dt * p->Dl / (p->Gamma * p->muk0) * Fm0 / (p->dx * p->dx);
            if (i == Nx - 2) s->led[Ny + j] += dt * p->Dl / (p->Gamma * p->muk0) * F10 / (p->dx * p->dx);
            /* the y faces: one row per thread, so the column entries of rows
             * 1 and Ny - 2 are written by one thread each */
            if (j == 1) s->led[2 * Ny + i] += dt * p->Dl / (p->Gamma * p->muk0) * F0m / (p->dx * p->dx);
            if (j == Ny - 2) s->led[2 * Ny + Nx + i] += dt * p->Dl / (p->Gamma * p->muk0) * F01 / (p->dx * p->dx);
            if (cc[4] + dc < 1.0e-6) s->led_clamp[j]++;
        }
    }
    or_ghosts(s, s->c_next);
}

/* Antitrapping current with the solute-trapping factor (an extension the
 * reference does not have; rs_params.at_coef/V_D/k_part). Face flux between
 * cells P and Q (Q = P + unit step along the face normal):
 *   J = a_t (1-k) c_L dphidt (-g_n/|g|) / (1 + |V|/V_D),
 *   g_n = (phi_Q - phi_P)/dx, g_t from the four cells beside the face,
 *   dphidt = mean over P, Q of (phi_next - phi)/dt, V = dphidt/|g|,
 *   c_L = c_f / (1 - h_f (1-k)), c_f, h_f = (1+phi_f)/2 face means,
 * and c_next -= dt (div J), then the clamp at 1e-6 and the ghost copies. */
double or_at_flux(const rs_params* p, double vdn, double dt, double pP, double pQ, double nP,
                  double nQ, double cP, double cQ, double t1, double t2, double t3, double t4) {
    /* J = a_t (1-k) c_L dphidt (-g_n/|g|) / (1 + |V|/V_D), c_L = c_f/(1 - h_f (1-k)),
     * written with the reciprocals of dx, 4 dx, dt, V_D and of |g|, and c_L
     * and the trapping factor under one division */
    const double idx = 1.0 / p->dx, i4dx = 1.0 / (4.0 * p->dx), idt = 1.0 / dt;
    const double ivdn = vdn > 0.0 ? 1.0 / vdn : 0.0;
    const double omk = 1.0 - p->k_part, coef = p->at_coef * (1.0 - p->k_part);
    const double gn = (pQ - pP) * idx;
    const double gt = ((t1 + t2) - (t3 + t4)) * i4dx;
    const double g2 = gn * gn + gt * gt;
    if (!(g2 > 1e-20)) return 0.0;
    const double ig = 1.0 / sqrt(g2);
    const double dpt = 0.5 * ((nP - pP) + (nQ - pQ)) * idt;
    const double cf = 0.5 * (cP + cQ);
    const double hf = 0.5 * (1.0 + 0.5 * (pP + pQ));
    double den = 1.0 - hf * omk;
    if (ivdn > 0.0) den = den * (1.0 + fabs(dpt * ig) * ivdn);
    return coef * (cf / den) * dpt * (-gn * ig);
}

static void or_antitrap(or_state* s) {
    const rs_params* p = &s->p;
    const int Nx = p->Nx, Ny = p->Ny;
    const double dt = s->k.dt;
    const double tau0 = or_tau0(p);
    const double vdn = p->V_D > 0.0 ? p->V_D * tau0 / (p->S * p->W0) : 0.0;
    const double* f = s->phi;
    const double* g = s->phi_next;
    const double* c = s->c;
#define F(i, j) f[POS(s, (i), (j))]
#define G(i, j) g[POS(s, (i), (j))]
#define CC(i, j) c[POS(s, (i), (j))]
#pragma omp parallel for schedule(static)
    for (int j = 1; j < Ny - 1; j++) {
        for (int i = 1; i < Nx - 1; i++) {
            /* x faces (i-1/2, j) and (i+1/2, j), y faces (i, j-1/2), (i, j+1/2) */
            const double Jxm = or_at_flux(p, vdn, dt, F(i - 1, j), F(i, j), G(i - 1, j), G(i, j),
                                          CC(i - 1, j), CC(i, j), F(i - 1, j + 1), F(i, j + 1),
                                          F(i - 1, j - 1), F(i, j - 1));
            const double Jxp = or_at_flux(p, vdn, dt, F(i, j), F(i + 1, j), G(i, j), G(i + 1, j),
                                          CC(i, j), CC(i + 1, j), F(i, j + 1), F(i + 1, j + 1),
                                          F(i, j - 1), F(i + 1, j - 1));
            const double Jym = or_at_flux(p, vdn, dt, F(i, j - 1), F(i, j), G(i, j - 1), G(i, j),
                                          CC(i, j - 1), CC(i, j), F(i + 1, j - 1), F(i + 1, j),
                                          F(i - 1, j - 1), F(i - 1, j));
            const double Jyp = or_at_flux(p, vdn, dt, F(i, j), F(i, j + 1), G(i, j), G(i, j + 1),
                                          CC(i, j), CC(i, j + 1), F(i + 1, j), F(i + 1, j + 1),
                                          F(i - 1, j), F(i - 1, j + 1));
            const double dc = -dt * ((Jxp - Jxm) + (Jyp - Jym)) * (1.0 / p->dx);
            s->c_next[POS(s, i, j)] = fmax(s->c_next[POS(s, i, j)] + dc, 1.0e-6);
        }
    }
#undef F
#undef G
#undef CC
    or_ghosts(s, s->c_next);
}

/* integrate_phi (fun.cu:1494-1535) */
static void or_integrate_phi(or_state* s, int i_left_end) {
    const int Nx = s->p.Nx, r = s->p.ratio_dxT_dx;
    const int NyT = s->k.NyT;
    int iT1 = i_left_end / r;
    int iTn = (i_left_end + Nx - 3) / r;
#pragma omp parallel for schedule(static)
    for (int jT = 0; jT < s->k.NyT_all; jT++) {
        for (int iT = 0; iT < s->k.NxT_all; iT++) {
            double integral = 0;
            if (iT >= iT1 && iT <= iTn && jT >= 1 && jT <= NyT - 2) {
                int i0 = r * iT + 1 - i_left_end;
                int in = i0 + r - 1;
                if (i0 < 1) i0 = 1;
                if (in > Nx - 2) in = Nx - 2;
                for (int j = 1 + r * (jT - 1); j <= r * jT; j++)
                    for (int i = i0; i <= in; i++)
                        integral += s->psi_next[POS(s, i, j)] - s->psi[POS(s, i, j)];
                s->dphidt[PST(s, iT, jT)] = integral * 0.5;
            } else {
                s->dphidt[PST(s, iT, jT)] = 0;
            }
        }
    }
}

/* test entry points of the two routines above: T_fine from the current
 * coarse T with the given i_left_end, and the latent-heat block sums of
 * (psi_next - psi) for given arrays (Nx*Ny each) into out (NxT_all*NyT_all) */
int or_interpolate_at(or_state* s, int i_left_end) {
    if (!s) return RS_ERR_ARG;
    or_interpolate(s, i_left_end);
    return RS_OK;
}

int or_integrate_phi_at(or_state* s, const double* psi, const double* psi_next, int i_left_end,
                        double* out) {
    if (!s || !psi || !psi_next || !out) return RS_ERR_ARG;
    memcpy(s->psi, psi, sizeof(double) * (size_t)s->n);
    memcpy(s->psi_next, psi_next, sizeof(double) * (size_t)s->n);
    or_integrate_phi(s, i_left_end);
    memcpy(out, s->dphidt, sizeof(double) * (size_t)s->nT);
    return RS_OK;
}

/* compute_temperature (fun.cu:1537-1578), race-free boundary semantics */
static void or_compute_temperature(or_state* s, int64_t step) {
    const rs_params* p = &s->p;
    const int NxT = s->k.NxT, NyT = s->k.NyT, r = p->ratio_dxT_dx;
    const double tau0 = or_tau0(p);
    const double dxT = p->dx * r;
    const double Lcp = or_Lcp(p);
    const double vadv = or_vp(p, step * tau0 * s->k.dt) * tau0 / (p->S * p->W0);
#pragma omp parallel for schedule(static)
    for (int jT = 1; jT <= NyT - 2; jT++) {
        for (int iT = 1; iT <= NxT - 2; iT++) {
            double T00 = s->T[PST(s, iT, jT)];
            double T10 = s->T[PST(s, iT + 1, jT)];
            double Tm0 = s->T[PST(s, iT - 1, jT)];
            double T01 = s->T[PST(s, iT, jT + 1)];
            double T0m = s->T[PST(s, iT, jT - 1)];
            double lap = (T10 + Tm0 + T01 + T0m - 4.0 * T00) / (dxT * dxT);
            double dTdt = vadv * (T10 - Tm0) / (2.0 * dxT) + lap * p->DT / (p->Gamma * p->muk0);
            s->T_next[PST(s, iT, jT)] =
                T00 + s->k.dtT * dTdt + Lcp * s->dphidt[PST(s, iT, jT)] / (s->k.ratio_dt_dtT * (r * r));
        }
    }
    for (int jT = 0; jT < NyT; jT++) {
        s->T_next[PST(s, 0, jT)] = s->k.Tcold;
        s->T_next[PST(s, NxT - 1, jT)] = s->k.Thot;
    }
    for (int iT = 0; iT < NxT; iT++) {
        s->T_next[PST(s, iT, 0)] = s->T_next[PST(s, iT, NyT - 2)];
        s->T_next[PST(s, iT, NyT - 1)] = s->T_next[PST(s, iT, 1)];
    }
    double* t = s->T;
    s->T = s->T_next;
    s->T_next = t;
}

/* get_itip_all_grain (fun.cu:1745-1754) */
int or_tip_column(or_state* s) {
    const int Nx = s->p.Nx, Ny = s->p.Ny;
    const double pin = s->p.phi_interface;
    int i_tip = 1;
    for (int j = 0; j < Ny; j++)
        for (int i = i_tip; i < Nx; i++)
            if (s->psi[POS(s, i, j)] > pin) i_tip = i;
    return i_tip;
}

/* set_pull_back state update (fun.cu:697-704) + pull_back (fun.cu:1469-1492) */
int or_pull_back(or_state* s, int i_offset) {
    const int Nx = s->p.Nx, Ny = s->p.Ny;
    const double m_before = or_solute_mass(s);
    s->k.i_offset = i_offset;
    s->k.i_offset_history += i_offset;
    if (s->p.ifTFC)
        s->k.i_left_end = s->k.i_left_end_initial + s->k.i_offset_history -
                          (int)(s->k.l_offset_Vp / s->p.dx);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) {
            int in = i + i_offset;
            if (in < 1) in = 1;
            if (in > Nx - 2) in = Nx - 2;
            s->c_next[POS(s, i, j)] = (in < Nx - 2) ? s->c[POS(s, in, j)] : s->p.cinf;
            s->phi_next[POS(s, i, j)] = s->phi[POS(s, in, j)];
            s->psi_next[POS(s, i, j)] = s->psi[POS(s, in, j)];
        }
    }
    double* t;
    t = s->c; s->c = s->c_next; s->c_next = t;
    t = s->phi; s->phi = s->phi_next; s->phi_next = t;
    t = s->psi; s->psi = s->psi_next; s->psi_next = t;
    s->led_pull += or_solute_mass(s) - m_before;
    return or_check_window(s, s->k.i_left_end);
}

int or_step(or_state* s, int64_t nsteps) {
    const rs_params* p = &s->p;
    const double tau0 = or_tau0(p);
    for (int64_t n = 0; n < nsteps; n++) {
        int64_t step = s->k.step;
        if (step % s->k.step_pull_back == 0) {
            int it = or_tip_column(s);
            int rc = or_pull_back(s, it - p->i_tip_target);
            if (rc) return rc;
        }
        or_compute_phi(s);
        or_compute_psi(s);
        or_compute_c(s);
        if (p->at_coef != 0.0) or_antitrap(s);
        if (p->ifTFC) {
            or_integrate_phi(s, s->k.i_left_end);
            for (int q = 0; q < s->k.ratio_dt_dtT; q++) or_compute_temperature(s, step);
            or_interpolate(s, s->k.i_left_end);
        }
        double* t;
        t = s->c; s->c = s->c_next; s->c_next = t;
        t = s->phi; s->phi = s->phi_next; s->phi_next = t;
        t = s->psi; s->psi = s->psi_next; s->psi_next = t;
        s->k.l_offset_Vp += or_vp(p, step * tau0 * s->k.dt) * s->k.dt * tau0 / (p->S * p->W0);
        s->k.step = step + 1;
    }
    return RS_OK;
}

double or_solute_mass(or_state* s) {
    const int Nx = s->p.Nx, Ny = s->p.Ny;
    double tot = 0.0;
    for (int j = 1; j < Ny - 1; j++) {
        double row = 0.0;
        for (int i = 1; i < Nx - 1; i++) row += s->c[POS(s, i, j)];
        tot += row;
    }
    return tot;
}

/* ------------------------------------------------------------------ */
/* tip fit (fun.cu:2628-2954)                                           */
/* ------------------------------------------------------------------ */
static int or_wrap(int x, int a, int b) { /* deal_with_boundary, periodic (fun.cu:2628) */
    int lo = a < b ? a : b, hi = a < b ? b : a;
    if (x <= lo - 1) x += (hi - lo + 1);
    else if (x >= hi + 1) x -= (hi - lo + 1);
    return x;
}
/* Gaussian elimination with partial pivoting through a row permutation
 * (fun.cu:2648-2688) */
static void or_solve5(double A[5][5], double b[5], double x[5]) {
    int pr[5] = {0, 1, 2, 3, 4};
    for (int k = 0; k < 4; k++) {
        double best = fabs(A[pr[k]][k]);
        int arg = k;
        for (int i = k + 1; i < 5; i++)
            if (fabs(A[pr[i]][k]) > best) { best = fabs(A[pr[i]][k]); arg = i; }
        int t = pr[k]; pr[k] = pr[arg]; pr[arg] = t;
        for (int i = k + 1; i < 5; i++) {
            double f = A[pr[i]][k] / A[pr[k]][k];
            for (int c = k; c < 5; c++) A[pr[i]][c] -= f * A[pr[k]][c];
            b[pr[i]] -= f * b[pr[k]];
        }
    }
    for (int i = 4; i >= 0; i--) {
        double acc = 0.0;
        for (int c = i + 1; c < 5; c++) acc += A[pr[i]][c] * x[c];
        x[i] = (b[pr[i]] - acc) / A[pr[i]][i];
    }
}
/* least-squares quartic x(y), evaluated at y = phi_interface (fun.cu:2691-2738) */
static double or_precise_x(const double* xs, const double* ys, int n, double yq) {
    double A[5][5] = {{0}}, b[5] = {0}, cf[5] = {0};
    for (int q = 0; q < n; q++) {
        double yp[9];
        yp[0] = 1.0;
        for (int k = 1; k <= 8; k++) yp[k] = yp[k - 1] * ys[q];
        for (int r = 0; r < 5; r++) {
            b[r] += xs[q] * yp[r];
            for (int c = 0; c < 5; c++) A[r][c] += yp[r + c];
        }
    }
    or_solve5(A, b, cf);
    return cf[0] + cf[1] * yq + cf[2] * yq * yq + cf[3] * yq * yq * yq + cf[4] * yq * yq * yq * yq;
}
/* row-swapping elimination (fun.cu:2740-2788) */
static int or_solve3(double A[3][3], double b[3], double x[3]) {
    for (int k = 0; k < 3; k++) {
        int pv = k;
        double mv = fabs(A[k][k]);
        for (int i = k + 1; i < 3; i++)
            if (fabs(A[i][k]) > mv) { mv = fabs(A[i][k]); pv = i; }
        if (mv < 1e-14) return -1;
        if (pv != k) {
            for (int c = 0; c < 3; c++) { double t = A[k][c]; A[k][c] = A[pv][c]; A[pv][c] = t; }
            double t = b[k]; b[k] = b[pv]; b[pv] = t;
        }
        double piv = A[k][k];
        for (int i = k + 1; i < 3; i++) {
            double f = A[i][k] / piv;
            A[i][k] = 0.0;
            for (int c = k + 1; c < 3; c++) A[i][c] -= f * A[k][c];
            b[i] -= f * b[k];
        }
    }
    for (int i = 2; i >= 0; i--) {
        double acc = b[i];
        for (int c = i + 1; c < 3; c++) acc -= A[i][c] * x[c];
        x[i] = acc / A[i][i];
    }
    return 0;
}
/* x(y) = x0 + (y-y0)^2/(2R) via a least-squares quadratic (fun.cu:2809-2865) */
static int or_parabola(const double* xs, const double* ys, int n, double* x0, double* y0,
                       double* R) {
    if (n < 3) return -1;
    double M[3][3] = {{0}}, v[3] = {0}, a[3] = {0};
    for (int q = 0; q < n; q++) {
        double f[3] = {1.0, ys[q], ys[q] * ys[q]};
        for (int r = 0; r < 3; r++) {
            for (int c = 0; c < 3; c++) M[r][c] += f[r] * f[c];
            v[r] += f[r] * xs[q];
        }
    }
    if (or_solve3(M, v, a) != 0) return -2;
    if (fabs(a[2]) < 1e-14) return -3;
    double Rl = 1.0 / (2.0 * a[2]);
    double yl = -a[1] / (2.0 * a[2]);
    *x0 = a[0] - (yl * yl) / (2.0 * Rl);
    *y0 = yl;
    *R = Rl;
    return 0;
}
/* get_tip_int_at_j (fun.cu:2867-2878) */
static int or_last_above(const rs_params* p, const double* psi, int i_start, int j) {
    int it = 0;
    for (int i = i_start; i < p->Nx - 1; i++)
        if (psi[(int64_t)j * p->Nx + i] > p->phi_interface) it = i;
    return it;
}

int or_precise_tip(const rs_params* p, const double* psi, int32_t* ti, int32_t* tj, double* tx,
                   double* ty, double* tR) {
    const int Nx = p->Nx, Ny = p->Ny;
    /* get_tip_int (fun.cu:2880-2893): running maximum of the sub-cell
     * estimate, scanning each row from the current integer tip */
    double itip = 0.0;
    int bi = 0, bj = 0;
    for (int j = 0; j < Ny; j++) {
        int i0 = or_last_above(p, psi, (int)itip, j);
        double a = psi[(int64_t)j * Nx + i0], b = psi[(int64_t)j * Nx + i0 + 1];
        double est = (b - a) * (p->phi_interface - a) + i0;
        if (est > itip) { itip = est; bi = (int)itip; bj = j; }
    }
    double xd[5], yd[5];
    for (int q = 0; q < 5; q++) {
        int j0 = or_wrap(bj - 2 + q, Ny - 2, 1);
        int i_start = bi - 5;
        int i0 = or_last_above(p, psi, i_start < 0 ? 0 : i_start, j0);
        double id[5], pd[5];
        for (int m = 0; m < 5; m++) {
            int i = i0 - 2 + m;
            if (i < 0) i = 0;
            if (i > Nx - 1) i = Nx - 1;
            id[m] = i0 - 2 + m;
            pd[m] = psi[(int64_t)j0 * Nx + i];
        }
        xd[q] = or_precise_x(id, pd, 5, p->phi_interface);
        yd[q] = j0;
    }
    double x0 = NAN, y0 = NAN, R = NAN;
    int st = or_parabola(xd, yd, 5, &x0, &y0, &R);
    *ti = bi;
    *tj = bj;
    *tx = x0;
    *ty = y0;
    double Rn = fabs(R) * p->dx * p->W0 * p->S;
    if (Rn > 1e5) Rn = 1e5;
    if (Rn < 1.0) Rn = 1.0;
    *tR = Rn;
    return st;
}

/* set_tip_output (fun.cu:3024-3072) with the tip.txt row of
 * src/main.cu:2671-2709 */
int or_tip(or_state* s, rs_tip_row* t) {
    const rs_params* p = &s->p;
    const double tau0 = or_tau0(p);
    const double dt = s->k.dt;
    int32_t ti, tj;
    double tx, ty, tR;
    int st = or_precise_tip(p, s->psi, &ti, &tj, &tx, &ty, &tR);
    int64_t step = s->k.step;
    double xnow = (s->k.i_offset_history + tx) * p->dx;
    double V = (xnow - s->k.x_tip_history) * p->S * p->W0 / (s->k.step_tip * tau0 * dt);
    s->k.x_tip_history = xnow;
    if (V < 0.0) V = 0.01;
    if (V > 10.0) V = 10.0;
    double Ttip, delta;
    if (p->ifTFC) {
        or_interpolate(s, s->k.i_left_end);
        Ttip = s->T_fine[POS(s, ti, 0)];
        delta = (Ttip - p->Tmelt) / p->deltaT0;
    } else {
        delta = or_delta(s, ti);
        Ttip = delta * p->deltaT0 + p->Tmelt;
    }
    s->k.delta = delta;
    if (s->k.if_begin_tip == 1) {
        s->k.if_begin_tip = 0;
        V = NAN;
    }
    t->step = step;
    t->t_us = step * dt * tau0 * 1e-3;
    t->delta = delta;
    t->T_tip = Ttip;
    t->V_tip = V;
    t->Vp = or_vp(p, step * tau0 * dt);
    t->R = tR;
    t->i = ti;
    t->j = tj;
    t->x = tx;
    t->y = ty;
    return st;
}

/* ------------------------------------------------------------------ */
/* common tangent (used to pin the CALPHAD routine and to pick the
 * configuration's solidus composition)                                */
/* ------------------------------------------------------------------ */
static double or_dmu(const rs_params* p, double T, double c, double L0, double L1) {
    return p->RR * T / (c * (1.0 - c)) - 2.0 * L0 + L1 * (-6.0 + 12.0 * c);
}
int or_common_tangent(const rs_params* p, double T, double* cl, double* cs) {
    double a = *cl, b = *cs; /* initial guesses from the caller */
    for (int it = 0; it < 200; it++) {
        double mul = or_Glc(p, T, a), mus = or_Gsc(p, T, b);
        double fl = or_Gl(p, T, a), fs = or_Gs(p, T, b);
        double F1 = mul - mus;
        double F2 = fl - fs - mul * (a - b);
        double dl = or_dmu(p, T, a, p->L0_liq, p->L1_liq);
        double ds = or_dmu(p, T, b, p->L0_bcc, p->L1_bcc);
        double J11 = dl, J12 = -ds, J21 = -dl * (a - b), J22 = mul - mus;
        double det = J11 * J22 - J12 * J21;
        if (det == 0.0 || !isfinite(det)) return -1;
        double da = (F1 * J22 - J12 * F2) / det;
        double db = (J11 * F2 - J21 * F1) / det;
        double lam = 1.0;
        while ((a - lam * da <= 1e-9 || a - lam * da >= 1 - 1e-9 || b - lam * db <= 1e-9 ||
                b - lam * db >= 1 - 1e-9) && lam > 1e-6)
            lam *= 0.5;
        a -= lam * da;
        b -= lam * db;
        if (fabs(da) < 1e-14 && fabs(db) < 1e-14) {
            *cl = a;
            *cs = b;
            return 0;
        }
    }
    *cl = a;
    *cs = b;
    return 1;
}

/* solute ledger (the bookkeeping rs_solute_ledger keeps on the GPU):
 * out[0] mass, [1] mass at the reset, [2] summed pull-back change,
 * [3]/[4] gains through the x / y boundary faces, [5] clamped
 * cells, [6] unaccounted change */
int or_solute_ledger(or_state* s, double* out7, int reset) {
    const int Ny = s->p.Ny, Nx = s->p.Nx;
    double X = 0.0, Y = 0.0;
    int nc = 0;
    for (int j = 1; j < Ny - 1; j++) {
        X += s->led[j] + s->led[Ny + j];
        nc += s->led_clamp[j];
    }
    for (int i = 1; i < Nx - 1; i++) Y += s->led[2 * Ny + i] + s->led[2 * Ny + Nx + i];
    const double m = or_solute_mass(s);
    out7[0] = m;
    out7[1] = s->led_m0;
    out7[2] = s->led_pull;
    out7[3] = X;
    out7[4] = Y;
    out7[5] = nc;
    out7[6] = ((m - s->led_m0) - s->led_pull) - (X + Y);
    if (reset) or_ledger_reset(s);
    return RS_OK;
}
