/*
 * tp_oracle.c -- CPU FP64 restatement of the 3D three-phase peritectic
 * Ti-Ag model (reference: multi/peritectic_multiGPU/TiAg.cu) on one z-slab.
 *
 * TEST INFRASTRUCTURE ONLY: the correctness oracle for the HIP stepper in
 * 2d-rapid-solidification_amd/csrc/tp_peritectic.hip. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Deliberately unfused, kernel by kernel as the reference runs them: mu is
 * stored (computeMu on every cell, then boundaryMu X, Y, Z), then diffuse on
 * the interior cells into the next fields, then boundary X, Y, Z on the next
 * fields, then the swap (TiAg.cu:288-345). Expressions keep the reference's
 * operation order; compiled with -ffp-contract=off.
 *
 * Pins (tests/test_peritectic_pins.py): Sum_i g_i = 1 on the simplex and the
 * derivatives gp1, gp2 against central differences of g; the bulk driving
 * forces against central differences of an independently written free
 * energy; pure-phase diffusion of c against the explicit-Euler decay of a
 * sinusoid (with and without the phase-2/3 mobilities); exact solute
 * conservation; the ghost images; the two-cap initial geometry against the
 * spherical-cap closed forms.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "tp_oracle.h"

typedef struct tpo_coef {
    double r1, r2, r3, s1, s2, s3;
} tpo_coef;

struct tpo_state {
    tp_params p;
    tpo_coef k;
    int64_t n, step;
    double *phi1, *phi2, *c, *mu, *phi1_next, *phi2_next, *c_next;
};

#define POS(s, i, j, k) ((int64_t)(k) * (s)->p.Nx * (s)->p.Ny + (int64_t)(j) * (s)->p.Nx + (i))

/* g, gp1, gp2 (TiAg.cu:71-73) */
double tpo_g(double pi, double pj, double pk) {
    return ((pi * pi) * 0.25) *
           (((15.0 * (1 - pi)) * ((1 + pi) - ((pk - pj) * (pk - pj)))) + (pi * ((9.0 * (pi * pi)) - 5.0)));
}
double tpo_gp1(double pi, double pj, double pk) {
    return (3.75 * pi) * (((((-4.0 * (pi * pi)) + (3.0 * ((pi * pi) * pi))) + 2.0) - pi) +
                          (((3.0 * pi) - 2.0) * ((pj - pk) * (pj - pk))));
}
double tpo_gp2(double pi, double pj, double pk) { return ((7.5 * (pi - 1.0)) * (pi * pi)) * (pj - pk); }

static tpo_coef coef_of(const tp_params* p) {
    const double B1 = p->B11 + p->B12 * p->T, B2 = p->B21 + p->B22 * p->T; /* TiAg.cu:56-57 */
    tpo_coef k;
    k.r1 = p->A1;          /* TiAg.cu:454, 496 */
    k.r2 = -p->A1 - p->A2;
    k.r3 = -p->A1 + p->A2;
    k.s1 = B1;             /* TiAg.cu:497 */
    k.s2 = -B1 - B2;
    k.s3 = -B1 + B2;
    return k;
}

/* computeMu for one cell (TiAg.cu:446-456) */
double tpo_mu(const tp_params* p, double p1, double p2, double c) {
    const tpo_coef k = coef_of(p);
    const double p3 = 1.0 - p1 - p2;
    return c - k.r1 * tpo_g(p1, p2, p3) - k.r2 * tpo_g(p2, p3, p1) - k.r3 * tpo_g(p3, p1, p2);
}

/* the bulk part of diffuse's muphi_i: fphi_i including the chemical
 * driving force fcphi_i, without the Laplacian (TiAg.cu:498-514) */
void tpo_bulk_force(const tp_params* p, double p1, double p2, double muu, double* out3) {
    const tpo_coef k = coef_of(p);
    const double p3 = 1.0 - p1 - p2;
    const double a1 = p->a1, a2 = p->a2, a3 = p->a3, b = p->b;
    const double fcphi1 = p->lambda * (tpo_gp1(p1, p2, p3) * (k.s1 - k.r1 * muu) +
                                       tpo_gp2(p2, p1, p3) * (k.s2 - k.r2 * muu) +
                                       tpo_gp2(p3, p1, p2) * (k.s3 - k.r3 * muu));
    const double fcphi2 = p->lambda * (tpo_gp2(p1, p2, p3) * (k.s1 - k.r1 * muu) +
                                       tpo_gp1(p2, p3, p1) * (k.s2 - k.r2 * muu) +
                                       tpo_gp2(p3, p2, p1) * (k.s3 - k.r3 * muu));
    const double fcphi3 = p->lambda * (tpo_gp2(p1, p3, p2) * (k.s1 - k.r1 * muu) +
                                       tpo_gp2(p2, p3, p1) * (k.s2 - k.r2 * muu) +
                                       tpo_gp1(p3, p1, p2) * (k.s3 - k.r3 * muu));
    out3[0] = 2.0 * p1 * (1.0 - p1) * (1.0 - 2.0 * p1) + a1 * ((p2 * p3) * (p2 * p3)) * (3.0 + 2 * b * p1) +
              2.0 * a2 * p1 * (p3 * p3) * (3 * p1 * p3 + 3.0 * p2 + b * (p2 * p2)) +
              2.0 * a3 * p1 * (p2 * p2) * (3.0 * p1 * p2 + 3.0 * p3 + b * (p3 * p3)) + fcphi1;
    out3[1] = 2.0 * p2 * (1.0 - p2) * (1.0 - 2.0 * p2) +
              2.0 * a1 * p2 * (p3 * p3) * (3 * p2 * p3 + 3.0 * p1 + b * (p1 * p1)) +
              a2 * ((p1 * p3) * (p1 * p3)) * (3.0 + 2 * b * p2) +
              2.0 * a3 * p2 * (p1 * p1) * (3.0 * p1 * p2 + 3.0 * p3 + b * (p3 * p3)) + fcphi2;
    out3[2] = 2.0 * p3 * (1.0 - p3) * (1.0 - 2.0 * p3) +
              2.0 * a1 * p3 * (p2 * p2) * (3.0 * p2 * p3 + 3.0 * p1 + b * (p1 * p1)) +
              2.0 * a2 * p3 * (p1 * p1) * (3 * p1 * p3 + 3.0 * p2 + b * (p2 * p2)) +
              a3 * ((p1 * p2) * (p1 * p2)) * (3.0 + 2 * b * p3) + fcphi3;
}

tpo_state* tpo_create(const tp_params* p) {
    if (!p || p->Nx < 4 || p->Ny < 4 || p->Nz < 4) return NULL;
    tpo_state* s = (tpo_state*)calloc(1, sizeof(tpo_state));
    if (!s) return NULL;
    s->p = *p;
    s->k = coef_of(p);
    s->n = (int64_t)p->Nx * p->Ny * p->Nz;
    double** f[] = {&s->phi1, &s->phi2, &s->c, &s->mu, &s->phi1_next, &s->phi2_next, &s->c_next};
    for (int q = 0; q < 7; q++) {
        *f[q] = (double*)calloc((size_t)s->n, sizeof(double));
        if (!*f[q]) {
            tpo_destroy(s);
            return NULL;
        }
    }
    return s;
}

void tpo_destroy(tpo_state* s) {
    if (!s) return;
    free(s->phi1); free(s->phi2); free(s->c); free(s->mu);
    free(s->phi1_next); free(s->phi2_next); free(s->c_next);
    free(s);
}

/* initial() (TiAg.cu:411-444) */
int tpo_init_reference(tpo_state* s) {
    const tp_params* p = &s->p;
    const double radius = p->R / sin(p->theta);
    const double roff = p->R / tan(p->theta);
    const double B1 = p->B11 + p->B12 * p->T, B2 = p->B21 + p->B22 * p->T;
    const double ca = (B1 + 0.5 * B2) / (p->A1 + 0.5 * p->A2) - p->A1 - p->A2; /* TiAg.cu:59 */
    const double cb = (B1 - 0.5 * B2) / (p->A1 - 0.5 * p->A2) - p->A1 + p->A2; /* TiAg.cu:60 */
    const double clb = (B1 - 0.5 * B2) / (p->A1 - 0.5 * p->A2) + p->A1;        /* TiAg.cu:62 */
    const int zcut = p->Nz - p->zcut_from_top;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < p->Nz; k++)
        for (int j = 0; j < p->Ny; j++)
            for (int i = 0; i < p->Nx; i++) {
                const int64_t ps = POS(s, i, j, k);
                int cap = 0;
                for (int q = 0; q < p->ncaps; q++) {
                    const double dx = i - p->cap_x[q], dy = j - p->cap_y[q];
                    const double dz = k - (zcut - roff);
                    if (dx * dx + dy * dy + dz * dz <= radius * radius) cap = 1;
                }
                if (k >= zcut && cap) {
                    s->phi1[ps] = 0;
                    s->phi2[ps] = 1.0;
                    s->c[ps] = ca;
                } else if (k < zcut) {
                    s->phi1[ps] = 0;
                    s->phi2[ps] = 0;
                    s->c[ps] = cb;
                } else {
                    s->phi1[ps] = 1.0;
                    s->phi2[ps] = 0;
                    s->c[ps] = clb;
                }
            }
    s->step = 0;
    return 0;
}

static double* field(tpo_state* s, int f) {
    switch (f) {
    case TP_FIELD_PHI1: return s->phi1;
    case TP_FIELD_PHI2: return s->phi2;
    case TP_FIELD_C: return s->c;
    case TP_FIELD_MU: return s->mu;
    }
    return NULL;
}

static void compute_mu(tpo_state* s);

int tpo_get(tpo_state* s, int f, double* out, int64_t n) {
    double* d = field(s, f);
    if (!d || !out || n != s->n) return -1;
    if (f == TP_FIELD_MU) compute_mu(s); /* computeMu + boundaryMu of the current fields */
    memcpy(out, d, sizeof(double) * (size_t)n);
    return 0;
}

int tpo_put(tpo_state* s, int f, const double* in, int64_t n) {
    double* d = field(s, f);
    if (!d || !in || n != s->n || f == TP_FIELD_MU) return -1;
    memcpy(d, in, sizeof(double) * (size_t)n);
    return 0;
}

/* boundaryMu{X,Y,Z} / boundary{X,Y,Z} on one array, in the reference's
 * launch order (TiAg.cu:458-483, 538-579) */
static void bx(tpo_state* s, double* f) {
    const int Nx = s->p.Nx;
    for (int k = 0; k < s->p.Nz; k++)
        for (int j = 0; j < s->p.Ny; j++) {
            f[POS(s, 0, j, k)] = f[POS(s, Nx - 2, j, k)];
            f[POS(s, Nx - 1, j, k)] = f[POS(s, 1, j, k)];
        }
}
static void by(tpo_state* s, double* f) {
    const int Ny = s->p.Ny;
    for (int k = 0; k < s->p.Nz; k++)
        for (int i = 0; i < s->p.Nx; i++) {
            f[POS(s, i, 0, k)] = f[POS(s, i, Ny - 2, k)];
            f[POS(s, i, Ny - 1, k)] = f[POS(s, i, 1, k)];
        }
}
static void bz(tpo_state* s, double* f) {
    const int Nz = s->p.Nz;
    for (int j = 0; j < s->p.Ny; j++)
        for (int i = 0; i < s->p.Nx; i++) {
            f[POS(s, i, j, 0)] = f[POS(s, i, j, 1)];
            f[POS(s, i, j, Nz - 1)] = f[POS(s, i, j, Nz - 2)];
        }
}

/* computeMu on every cell, then boundaryMu X, Y, Z */
static void compute_mu(tpo_state* s) {
    const tpo_coef k = s->k;
#pragma omp parallel for schedule(static)
    for (int64_t ps = 0; ps < s->n; ps++) {
        const double p1 = s->phi1[ps], p2 = s->phi2[ps], p3 = 1.0 - p1 - p2;
        s->mu[ps] = s->c[ps] - k.r1 * tpo_g(p1, p2, p3) - k.r2 * tpo_g(p2, p3, p1) - k.r3 * tpo_g(p3, p1, p2);
    }
    bx(s, s->mu);
    by(s, s->mu);
    bz(s, s->mu);
}

/* diffuse on the interior (TiAg.cu:486-534) */
static void diffuse(tpo_state* s) {
    const tp_params* p = &s->p;
    const double DXin = p->DXin, DYin = p->DYin, DZin = p->DZin, DT = p->DT;
    const double Da = p->Da, Db = p->Db;
    const double* phi1 = s->phi1;
    const double* phi2 = s->phi2;
    const double* mu = s->mu;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < p->Nz - 1; k++)
        for (int j = 1; j < p->Ny - 1; j++)
            for (int i = 1; i < p->Nx - 1; i++) {
                const int64_t ps = POS(s, i, j, k);
                const double p1 = phi1[ps], p2 = phi2[ps], p3 = 1.0 - p1 - p2, muu = mu[ps];
                double f[3];
                tpo_bulk_force(p, p1, p2, muu, f);
                const int64_t xp = POS(s, i + 1, j, k), xn = POS(s, i - 1, j, k);
                const int64_t yp = POS(s, i, j + 1, k), yn = POS(s, i, j - 1, k);
                const int64_t zp = POS(s, i, j, k + 1), zn = POS(s, i, j, k - 1);
                const double lap1 = (phi1[xp] + phi1[xn] - 2.0 * p1) * DXin * DXin +
                                    (phi1[yp] + phi1[yn] - 2.0 * p1) * DYin * DYin +
                                    (phi1[zp] + phi1[zn] - 2.0 * p1) * DZin * DZin;
                const double lap2 = (phi2[xp] + phi2[xn] - 2.0 * p2) * DXin * DXin +
                                    (phi2[yp] + phi2[yn] - 2.0 * p2) * DYin * DYin +
                                    (phi2[zp] + phi2[zn] - 2.0 * p2) * DZin * DZin;
                const double lap3 = -lap1 - lap2;
                const double muphi1 = f[0] - lap1, muphi2 = f[1] - lap2, muphi3 = f[2] - lap3;
                const double phi1t = -(2.0 * muphi1 - muphi2 - muphi3) * 0.333333333333333333333333333;
                const double phi2t = -(2.0 * muphi2 - muphi1 - muphi3) * 0.333333333333333333333333333;
                s->phi1_next[ps] = p1 + DT * phi1t;
                s->phi2_next[ps] = p2 + DT * phi2t;
#define TPO_H(q, d, din) \
    ((0.5 * (p1 + phi1[q]) + Da * 0.5 * (p2 + phi2[q]) + Db * 0.5 * (p3 + 1.0 - phi1[q] - phi2[q])) * (d) * (din))
                const double Hxp = TPO_H(xp, mu[xp] - muu, DXin);
                const double Hxn = TPO_H(xn, muu - mu[xn], DXin);
                const double Hyp = TPO_H(yp, mu[yp] - muu, DYin);
                const double Hyn = TPO_H(yn, muu - mu[yn], DYin);
                const double Hzp = TPO_H(zp, mu[zp] - muu, DZin);
                const double Hzn = TPO_H(zn, muu - mu[zn], DZin);
#undef TPO_H
                s->c_next[ps] = s->c[ps] + DT * p->alpha * ((Hxp - Hxn) * DXin + (Hyp - Hyn) * DYin +
                                                             (Hzp - Hzn) * DZin);
            }
}

int tpo_step(tpo_state* s, int64_t nsteps) {
    for (int64_t q = 0; q < nsteps; q++) {
        compute_mu(s);
        diffuse(s);
        double* nx[3] = {s->phi1_next, s->phi2_next, s->c_next};
        for (int f = 0; f < 3; f++) {
            bx(s, nx[f]);
            by(s, nx[f]);
            bz(s, nx[f]);
        }
        double* t;
        t = s->phi1; s->phi1 = s->phi1_next; s->phi1_next = t;
        t = s->phi2; s->phi2 = s->phi2_next; s->phi2_next = t;
        t = s->c; s->c = s->c_next; s->c_next = t;
        s->step++;
    }
    return 0;
}

int64_t tpo_get_step(tpo_state* s) { return s->step; }

/* mean of c over the interior, largest |c| over the interior of plane kc
 * (TiAg.cu:229-262); the sum is compensated (Neumaier), accurate to a few
 * ulp whatever the cell count */
int tpo_report(tpo_state* s, int kc, double* out2) {
    const tp_params* p = &s->p;
    double tot = 0.0, comp = 0.0, mx = 0.0;
    for (int k = 1; k < p->Nz - 1; k++)
        for (int j = 1; j < p->Ny - 1; j++)
            for (int i = 1; i < p->Nx - 1; i++) {
                const double v = s->c[POS(s, i, j, k)];
                const double t = tot + v;
                comp += fabs(tot) >= fabs(v) ? (tot - t) + v : (v - t) + tot;
                tot = t;
            }
    tot += comp;
    for (int j = 1; j < p->Ny - 1; j++)
        for (int i = 1; i < p->Nx - 1; i++) mx = fmax(mx, fabs(s->c[POS(s, i, j, kc)]));
    out2[0] = tot / ((double)(p->Nx - 2) * (p->Ny - 2) * (p->Nz - 2));
    out2[1] = mx;
    return 0;
}
