// rs_mixed.hpp -- the mixed-precision stepper: psi/phi and the fine-grid
// temperature in FP32; c, the solute flux divergence, the ln terms of the
// chemical potential and all time accumulation in FP64.
#pragma once

#include "rs_antitrap.hpp"
#include "rs_kernels.hpp"
#include "rs_mlp.hpp"

namespace rs {

// Reference temperature of the FP32 temperature offsets T' = T - kTref used
// by the mixed path and its SGTE polynomial table (rs_sgte_poly.inc).
constexpr double kMixedTref = 1941.0;

constexpr int kCountSlots = 64;
constexpr int kCountStride = 16;
constexpr size_t kCountBytes = sizeof(unsigned long long) * kCountSlots * kCountStride;

struct StepArgsMixed {
    const float* psi;
    const double* c;
    float* psi_n;
    double* c_n;
    float* lh;  // latent-heat column sums (FP32: the step accumulates them in FP32)
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
    // learned surrogate (surrogate kernel only): weight fragments of the
    // activation format, trust bits, fallback counters [fallback, all],
    // input/output normalisation
    const MlpInfer* mlp;
    const uint64_t* trust;
    // surrogate counts, kCountSlots slots of kCountStride words (one
    // 128-byte line each; a workgroup adds to slot id % kCountSlots, so the
    // atomics of a launch spread over 64 lines instead of queueing on one):
    // [0] fallback cells, [1] all cells, [2] rows without the network, [3] rows
    unsigned long long* fb_count;
    float s_a_scale, s_a_off;  // a = T' s_a_scale + s_a_off
    float y_mean[4], y_std[4];
    int s_fmt;                 // activation format (kActBf16 ... kActFp16Pk)
    int s_always;              // FP16 formats: evaluate the network in every row (no skip test)
    // solute ledger (or null), as in StepArgsF64: boundary-face gains per
    // row (x faces) and per column (y faces)
    double* led;
    // dispatch order of the single-run exact-table kernel: column tiles from
    // the last down (the front is held in the upper half of the grid)
    int tiles_desc;
};

// ident: the crystal orientation is the identity (the reference's only
// built orientation, alpha0 = beta0 = gamma0 = 0), which removes the z row
// of the rotated normal
// surr: thermodynamics from the learned surrogate with exact fallback
// *variant (optional) receives the SURR template argument of the kernel it
// launched: 0 exact table, 1 BF16, 2 split BF16, 3 FP16 (rows without a
// trusted cell skip the network), 4 packed FP16, 7 FP16 in every row, 8
// packed FP16 in every row
hipError_t launch_step_mixed(const StepArgsMixed& a, int wpb, bool ident, bool surr, hipStream_t s,
                             int* variant = nullptr, LaunchEv ev = {});
hipError_t launch_T32_copy(const double* T, float* T32, int64_t n, double Tref, hipStream_t s);
hipError_t launch_init_fields_f32(float* psi, double* c, const double* x0, int Nx, int Ny,
                                  double dx, int mode, double cinf, double c_liq, double c_sol,
                                  hipStream_t s);
hipError_t launch_pull_back_f32(const float* psi, const double* c, float* psi_n, double* c_n,
                                int Nx, int Ny, int off, double cinf, hipStream_t s);
hipError_t launch_tip_column_f32(const float* psi, int Nx, int Ny, float pin, int* rowbuf,
                                 int* out, hipStream_t s);

// ---------------------------------------------------------------------------
// batched ensemble: many independent runs stepped by one persistent launch
// ---------------------------------------------------------------------------
// Per-run descriptor, fixed between pull-backs. The field pointers of the
// step arguments are chosen on the device from the run's parity.
struct BatchRun {
    StepArgsMixed a;
    float* psi[2];
    double* c[2];
    float* T32[2];
    double* T[2];
    TempArgs t;       // coarse-T arguments; T/Tn/T32n and vadv set on the device
    int ntx, nbg;     // column tiles, band groups (bands / waves per block)
    int tfc, ncT;     // coupled T; coarse cells updated per step (NxT * NyT)
    // time-dependent scalars, restated from the host loop (fun.cu:508-509,
    // 1610-1619, 1974-1981) with the same FP64 operations
    double Vp, Vp_rate, tau0, dt, S, W0, SW0;
    int ramp, ioff;
    double Tmelt, deltaT0, delta_initial, GG, l_off0, gstep;
    AtArgs at;        // antitrapping correction (extension), if at_on
    int at_on;
};
// Per-run state advanced on the device by the captured step sequence
struct BatchDyn {
    long long step;
    double l_offset_Vp;
    int cur, tcur, ile_interp, pad;
};
struct BatchLaunch {
    const BatchRun* runs;
    BatchDyn* dyn;
    const int* item0;   // nruns + 1 prefix sums of step work items
    const int* cell0;   // nruns + 1 prefix sums of coarse cells
    int nruns, items, cells;
    int step_blocks, wpb, T_blocks;
    bool ident;
    int surr;           // 0 exact, else 1 + the surrogate's activation format
    int at_any;         // some run has the antitrapping correction on
    int nsub;           // coarse-T sub-steps per step: the largest ratio_dt_dtT of a coupled run
    int led_any;        // some run keeps a solute ledger (its BatchRun's args carry led)
};
// one step of every run: fused step kernel, antitrapping pass (runs with
// it on), coarse T kernel once per sub-step, state advance
hipError_t launch_batch_step(const BatchLaunch& L, hipStream_t s);
// blocks of the step kernel resident per CU for waves-per-block wpb
int batch_step_resident(const BatchLaunch& L);

// facts of the build-time SGTE table: window [T_lo, T_hi] (K) and the
// largest FP32-evaluation error against the FP64 routine (J/mol)
void sgte_table_info(double* T_lo, double* T_hi, double* max_err);

}  // namespace rs
