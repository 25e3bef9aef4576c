// rs_mixed.hpp -- the mixed-precision stepper: psi/phi and the fine-grid
// temperature in FP32; c, the solute flux divergence, the ln terms of the
// chemical potential and all time accumulation in FP64.
#pragma once

#include "rs_kernels.hpp"

namespace rs {

constexpr int kPolyDeg = 7;  // degree of the T polynomials (coefficients 1..kPolyDeg in FP32)

// Temperature dependence of the SGTE unaries on a window around the
// interface temperature, as two polynomial pieces split at the Ti
// breakpoint Tb (1941 K). Functions (index f):
//   0: U_L = G_Ti_Liq - G_Nb_Liq     1: U_S = G_Ti_BCC - G_Nb_BCC
//   2: dG_Ti = G_Ti_Liq - G_Ti_BCC   3: dG_Nb = G_Nb_Liq - G_Nb_BCC
// Each piece: value = c0[f] + sum_k c[f][k-1] t^k, t = (T' - cen) * isc,
// T' = T - Tref. c0 is kept in FP64 so that the differences of the
// chemical potential between neighbouring cells keep their digits.
struct PolyTable {
    float lo, hi;            // window in T' (outside: exact FP64 routine)
    float tb;                // breakpoint in T'
    float cen[2], isc[2];
    double c0[2][4];
    float c[2][4][kPolyDeg];
    double Tref;
};

struct StepArgsMixed {
    const float* psi;
    const double* c;
    float* psi_n;
    double* c_n;
    double* lh;
    const float* T32;        // coarse T - Tref in FP32 (NxT_all pitch)
    int Nx, Ny, band_h, nbands, lh_last_row, lh_on;
    int tfc, ratio, NxT_all, ile;
    float dxT;
    // frozen T: T' = f0 + fg * i  (FP32 in T', evaluated from FP64 on the host)
    double fta_a, fta_b;     // T'(i) = fta_a + fta_b * i
    float dt, AA, q_solid, dG_den_inv;
    float eps1, epk1, lap_w_axis, lap_w_diag, inv2dx, invdx2, inv4dx2, inv2dx2;
    float ops[6];
    double mu_den, kdc_dx2;  // 1/(h0 v0), dt Dl/(Gamma muk0)/dx^2
    double L0l, L0s, L1l, L1s, RR;
    float dL0, dL1;          // L0_liq - L0_bcc, L1_liq - L1_bcc
    PolyTable poly;
    RK rk;                   // exact routine for out-of-window cells
};

hipError_t launch_step_mixed(const StepArgsMixed& a, int wpb, hipStream_t s);
hipError_t launch_T32_copy(const double* T, float* T32, int64_t n, double Tref, hipStream_t s);
hipError_t launch_init_fields_f32(float* psi, double* c, const double* x0, int Nx, int Ny,
                                  double dx, int mode, double cinf, double c_liq, double c_sol,
                                  hipStream_t s);
hipError_t launch_pull_back_f32(const float* psi, const double* c, float* psi_n, double* c_n,
                                int Nx, int Ny, int off, double cinf, hipStream_t s);
hipError_t launch_tip_column_f32(const float* psi, int Nx, int Ny, float pin, int* rowbuf,
                                 int* out, hipStream_t s);
hipError_t launch_f32_to_f64(const float* a, double* b, int64_t n, hipStream_t s);
hipError_t launch_f64_to_f32(const double* a, float* b, int64_t n, hipStream_t s);

// host: fit the polynomial table on [T_lo, T_hi] around Tref; returns the
// largest absolute fit error (J/mol) seen on a dense check grid
double fit_poly_table(const rs_params& p, double T_lo, double T_hi, double Tref, PolyTable* out);

}  // namespace rs
