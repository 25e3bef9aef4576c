// rs_mixed.hpp -- the mixed-precision stepper: psi/phi and the fine-grid
// temperature in FP32; c, the solute flux divergence, the ln terms of the
// chemical potential and all time accumulation in FP64.
#pragma once

#include "rs_kernels.hpp"
#include "rs_mlp.hpp"

namespace rs {

// Reference temperature of the FP32 temperature offsets T' = T - kTref used
// by the mixed path and its SGTE polynomial table (rs_sgte_poly.inc).
constexpr double kMixedTref = 1941.0;

struct StepArgsMixed {
    const float* psi;
    const double* c;
    float* psi_n;
    double* c_n;
    double* lh;
    const float* T32;        // coarse T - kMixedTref in FP32 (NxT_all pitch)
    int Nx, Ny, band_h, nbands, lh_last_row, lh_on;
    int tfc, NxT_all, ile;
    int dxTi;                // dx*ratio as an integer row offset (6 or 8 cells)
    int has_L1;              // Redlich-Kister L1 terms present
    // frozen T: T'(i) = fta_a + fta_b * i
    double fta_a, fta_b;
    float dt, AA, q_solid, dG_den_inv;
    float eps1, epk1, lap_w_axis, lap_w_diag, inv2dx, invdx2, inv4dx2, inv2dx2;
    float ops[6];
    double K;                // dt Dl/(Gamma muk0)/dx^2 / (h0 v0)
    double L0l, L1l, RR;
    float dL0, dL1;          // L0_liq - L0_bcc, L1_liq - L1_bcc
    RK rk;                   // exact routine for cells outside the table
    // learned surrogate (surrogate kernel only): BF16 fragments, trust
    // bits, fallback counters [fallback, all], input/output normalisation
    const MlpInfer* mlp;
    const uint64_t* trust;
    unsigned long long* fb_count;
    float s_a_scale, s_a_off;  // a = T' s_a_scale + s_a_off
    float y_mean[4], y_std[4];
    int s_split;               // split (hi + lo) BF16 activations
};

// ident: the crystal orientation is the identity (the reference's only
// built orientation, alpha0 = beta0 = gamma0 = 0), which removes the z row
// of the rotated normal
// surr: thermodynamics from the learned surrogate with exact fallback
hipError_t launch_step_mixed(const StepArgsMixed& a, int wpb, bool ident, bool surr, hipStream_t s);
hipError_t launch_T32_copy(const double* T, float* T32, int64_t n, double Tref, hipStream_t s);
hipError_t launch_init_fields_f32(float* psi, double* c, const double* x0, int Nx, int Ny,
                                  double dx, int mode, double cinf, double c_liq, double c_sol,
                                  hipStream_t s);
hipError_t launch_pull_back_f32(const float* psi, const double* c, float* psi_n, double* c_n,
                                int Nx, int Ny, int off, double cinf, hipStream_t s);
hipError_t launch_tip_column_f32(const float* psi, int Nx, int Ny, float pin, int* rowbuf,
                                 int* out, hipStream_t s);

// facts of the build-time SGTE table: window [T_lo, T_hi] (K) and the
// largest FP32-evaluation error against the FP64 routine (J/mol)
void sgte_table_info(double* T_lo, double* T_hi, double* max_err);

}  // namespace rs
