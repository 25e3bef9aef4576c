// rs_host.hpp -- host-side pieces of the library shared between its
// translation units.
#pragma once

#include <string>
#include <vector>

#include "rs_kernels.hpp"
#include "rs_mixed.hpp"

namespace rs {

int common_tangent(const rs_params& p, double T, double* cl, double* cs);
int liquidus(const rs_params& p, double c_liq, double* TL, double* c_sol);

// Tip of the most advanced dendrite from a host psi array (Nx*Ny):
// integer tip by the running row scan, then a quartic x(psi) fit on five
// rows and a parabola x(y) through the five interface points
// (get_precise_tip, fun.cu:2895-2954). Returns the parabola-fit status.
struct TipFit {
    int i, j;
    double x, y, R;  // R in nm, clamped to [1, 1e5]
};
int precise_tip(const rs_params& p, const double* psi, TipFit* out);
// the same fit from per-row data (irow[Ny], vals[8 Ny], as tip_rows_kernel
// writes them): no field download
int precise_tip_rows(const rs_params& p, const int* irow, const double* vals, TipFit* out);

// restarts the context's solute ledger at the current fields (rs_capi.hip)
int solute_ledger_reset(rs_ctx* ctx);
// the message rs_last_error(NULL) returns on this thread (rs_capi.hip)
void set_thread_error(const std::string& why);

// step arguments of a context at its current state (rs_capi.hip)
StepArgsMixed mixed_step_args(rs_ctx* ctx);
TempArgs coarse_T_args(rs_ctx* ctx, int64_t step);
double tau0_value(const rs_params& p);

// learned CALPHAD surrogate (rs_surr.hip)
struct SurrState;
void surr_free(rs_ctx* ctx);
bool surr_ready(const rs_ctx* ctx);
// identity of the installed network (0 when none): equal for networks that
// evaluate alike (weights, normalisation, window, tolerance, format)
uint64_t surr_identity(const rs_ctx* ctx);
// host copy of the trust bitmap (kTrustN words), nullptr when none
const uint64_t* surr_trust_host(const rs_ctx* ctx);
void surr_step_args(rs_ctx* ctx, StepArgsMixed& a);
int surr_update_policy(rs_ctx* ctx);
// rows without the network (skip) and all rows the step kernel counted since
// the counters were last cleared (waits for the context's stream)
int surr_row_counts(rs_ctx* ctx, unsigned long long* skip, unsigned long long* rows);

}  // namespace rs

struct rs_ctx {
    rs_params p;
    rs_constants k;
    rs::Derived d;
    double orient[9];
    hipStream_t stream = nullptr;
    int device = 0;
    int64_t n = 0, nT = 0;
    double* psi[2] = {nullptr, nullptr};
    double* c[2] = {nullptr, nullptr};
    double* T[2] = {nullptr, nullptr};
    // mixed-precision path: psi in FP32 and an FP32 copy of T - Tref
    float* psi32[2] = {nullptr, nullptr};
    float* T32[2] = {nullptr, nullptr};
    bool mixed = false;
    bool surrogate = false;  // RS_PREC_SURROGATE: surrogate thermodynamics in the step
    rs::SurrState* surr = nullptr;
    double* lh = nullptr;
    double* rowd = nullptr;
    int* rowi = nullptr;
    double* outd = nullptr;
    int* outi = nullptr;
    double* x0 = nullptr;
    double* tfine = nullptr;
    double* h_pin = nullptr;  // pinned scalars for D2H
    // solute ledger: boundary-face gains per row (x faces, 2 Ny) and per
    // column (y faces, 2 Nx), and led_acc[0] the summed mass change of the
    // pull-backs, [1] the mass at the last reset, [2..3] a pull-back's mass
    // before / after
    double* led = nullptr;
    double* led_acc = nullptr;
    bool led_on = false;  // recording: from the first rs_solute_ledger call with reset
    double* tipvals = nullptr;  // per-row tip-fit data (8 Ny) and columns (Ny)
    int* tiprow = nullptr;
    int cur = 0, tcur = 0;
    int band_h = 40, wpb = 4;
    // surrogate steps: -1 choose per run whether every row evaluates the
    // network (surr_update_policy), 0 never, 1 always (RS_SURR_ALWAYS)
    int surr_always = -1;
    int64_t surr_steps = 0;  // surrogate steps taken (policy schedule)
    // step kernel of the last rs_step launch: variant (launch_step_mixed's
    // code; -1 = the FP64 kernel or none yet), identity-orientation flag,
    // and how many launches in a row used that kernel
    int last_variant = -1, last_ident = -1;
    int64_t variant_run = 0;
    std::string err;
    // HIP-event timing of the fused step kernel [0] and the coarse T kernel
    // [1] on every timing-th step (0: off)
    int timing = 0;
    int64_t timing_ctr = 0;
    // finite-value sentinel: check every finite_every-th step (0: off);
    // finite_out holds (count, first index) per checked field
    int finite_every = 0;
    int64_t finite_ctr = 0;
    unsigned long long* finite_out = nullptr;
    std::vector<hipEvent_t> ev[3];
    int nev[3] = {0, 0, 0};
};
