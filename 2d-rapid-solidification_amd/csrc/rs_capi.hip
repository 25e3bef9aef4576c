// rs_capi.hip -- the C ABI (include/rsolid.h): context lifetime, uploads,
// and the host side of the time loop (fun.cu:431-514 without its output
// calls). All device work goes to the context's own stream.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <new>

#include "rs_host.hpp"

using namespace rs;

namespace {

// why the calling thread's last rs_create failed (rs_last_error(NULL))
thread_local std::string g_create_err;

int create_fail(int code, const std::string& why) {
    g_create_err = why;
    return code;
}

// Every error return waits for the context stream first: a call that
// enqueued a copy from one of its own host buffers and then failed must not
// return (and free that buffer) while the copy is still pending.
int fail(rs_ctx* ctx, int code, const char* what, hipError_t e = hipSuccess) {
    if (ctx) {
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        char buf[512];
        if (e != hipSuccess)
            snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
        else
            snprintf(buf, sizeof buf, "%s", what);
        ctx->err = buf;
    }
    return code;
}

#define HIPCHK(ctx, call)                                              \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) return fail((ctx), RS_ERR_HIP, #call, e_); \
    } while (0)

double tau0_of(const rs_params& p) {
    const double sw = p.S * p.W0;
    return (sw * sw) / (p.Gamma * p.muk0);
}

int check_params(const rs_params& p, std::string* why) {
    if (p.Nx < 8 || p.Ny < 8) { *why = "grid must be at least 8x8"; return RS_ERR_ARG; }
    if (p.S < 1 || p.dx <= 0 || p.W0 <= 0) { *why = "S, dx and W0 must be positive"; return RS_ERR_ARG; }
    if (p.ratio_dxT_dx != 10) { *why = "ratio_dxT_dx must be 10 (10-row latent-heat blocks)"; return RS_ERR_ARG; }
    if (p.precision != RS_PREC_FP64 && p.precision != RS_PREC_MIXED && p.precision != RS_PREC_SURROGATE) {
        *why = "bad precision";
        return RS_ERR_ARG;
    }
    if (p.ifTFC && p.Ny < 22) { *why = "coupled T needs Ny >= 22"; return RS_ERR_ARG; }
    if (p.at_coef != 0.0 && !(p.k_part > 0.0 && p.V_D >= 0.0)) {
        *why = "antitrapping needs 0 < k_part and V_D >= 0";
        return RS_ERR_ARG;
    }
    if (p.i_tip_target < 1 || p.i_tip_target > p.Nx - 2) { *why = "i_tip_target outside the grid"; return RS_ERR_ARG; }
    if ((int64_t)p.Nx * p.Ny >= (int64_t)1 << 31) { *why = "grid exceeds 2^31 cells (32-bit offsets)"; return RS_ERR_ARG; }
    if (p.precision != RS_PREC_FP64) {
        // the mixed stepper indexes the coarse-T rows with integer arithmetic
        const double dxT = p.dx * p.ratio_dxT_dx;
        if (fabs(dxT - (double)lround(dxT)) > 1e-12) { *why = "mixed path needs dx*ratio_dxT_dx integral"; return RS_ERR_ARG; }
        // its raw buffer views address an FP64 field with 32-bit byte offsets
        if ((int64_t)p.Nx * p.Ny * 8 >= (int64_t)1 << 32) { *why = "mixed path needs Nx*Ny < 2^29 cells"; return RS_ERR_ARG; }
    }
    return RS_OK;
}

// T window check: the fine window must interpolate from interior coarse
// columns (the reference reads past the coarse grid otherwise,
// fun.cu:1588-1599)
int check_window(rs_ctx* ctx, int ile) {
    if (!ctx->p.ifTFC) return RS_OK;
    const int r = ctx->p.ratio_dxT_dx;
    if (ile - 1 < 0 || (ctx->p.Nx - 2 + ile) / r + 1 > ctx->k.NxT - 1) {
        char b[256];
        snprintf(b, sizeof b, "fine window left the temperature sample (i_left_end=%d, NxT=%d)", ile,
                 ctx->k.NxT);
        return fail(ctx, RS_ERR_DOMAIN, b);
    }
    return RS_OK;
}

void free_all(rs_ctx* ctx) {
    double* ds[] = {ctx->psi[0], ctx->psi[1], ctx->c[0], ctx->c[1], ctx->T[0], ctx->T[1],
                    ctx->lh, ctx->rowd, ctx->outd, ctx->x0, ctx->tfine};
    for (double* d : ds)
        if (d) (void)hipFree(d);
    float* fs[] = {ctx->psi32[0], ctx->psi32[1], ctx->T32[0], ctx->T32[1]};
    for (float* f : fs)
        if (f) (void)hipFree(f);
    surr_free(ctx);
    if (ctx->rowi) (void)hipFree(ctx->rowi);
    if (ctx->tipvals) (void)hipFree(ctx->tipvals);
    if (ctx->tiprow) (void)hipFree(ctx->tiprow);
    if (ctx->outi) (void)hipFree(ctx->outi);
    if (ctx->finite_out) (void)hipFree(ctx->finite_out);
    if (ctx->led) (void)hipFree(ctx->led);
    if (ctx->led_acc) (void)hipFree(ctx->led_acc);
    if (ctx->h_pin) (void)hipHostFree(ctx->h_pin);
    for (int q = 0; q < 3; ++q)
        for (hipEvent_t e : ctx->ev[q]) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
}

// the next (start, stop) pair of timing series `which`, written by the
// dispatches of the launches it is handed to; left null once the pool is full
int ev_pair(rs_ctx* ctx, int which, LaunchEv* out) {
    std::vector<hipEvent_t>& v = ctx->ev[which];
    const int need = ctx->nev[which] + 2;
    if ((int)v.size() < need) {
        if (v.size() >= 8192) return RS_OK;  // pool full: stop sampling
        while ((int)v.size() < need) {
            hipEvent_t e;
            HIPCHK(ctx, hipEventCreate(&e));
            v.push_back(e);
        }
    }
    out->start = v[ctx->nev[which]];
    out->stop = v[ctx->nev[which] + 1];
    ctx->nev[which] += 2;
    return RS_OK;
}

StepArgsF64 step_args(rs_ctx* ctx) {
    const rs_params& p = ctx->p;
    StepArgsF64 a;
    a.psi = ctx->psi[ctx->cur];
    a.c = ctx->c[ctx->cur];
    a.psi_n = ctx->psi[ctx->cur ^ 1];
    a.c_n = ctx->c[ctx->cur ^ 1];
    a.lh = ctx->lh;
    a.Nx = p.Nx;
    a.Ny = p.Ny;
    a.band_h = ctx->band_h;
    a.nbands = (p.Ny - 2 + a.band_h - 1) / a.band_h;
    a.lh_last_row = p.ratio_dxT_dx * (ctx->k.NyT - 2);
    a.lh_on = p.ifTFC;
    a.ts = make_tempsrc(p, ctx->k, ctx->T[ctx->tcur], ctx->k.i_left_end_interp);
    a.dt = ctx->k.dt;
    a.AA = p.AA;
    a.q_solid = p.q_solid;
    a.dG_den = ctx->d.dG_den;
    a.mu_den = ctx->d.mu_den;
    a.kdc = ctx->d.kdc;
    a.dx2 = p.dx * p.dx;
    a.rk = rk_of(p);
    a.an.ops[0] = ctx->orient[0];
    a.an.ops[1] = ctx->orient[1];
    a.an.ops[2] = ctx->orient[3];
    a.an.ops[3] = ctx->orient[4];
    a.an.ops[4] = ctx->orient[6];
    a.an.ops[5] = ctx->orient[7];
    a.an.ops[6] = a.an.ops[7] = 0.0;
    a.an.eps1 = p.eps1;
    a.an.epk1 = p.epk1;
    a.an.lap_w_axis = p.lap_w_axis;
    a.an.lap_w_diag = p.lap_w_diag;
    a.an.inv2dx = 1.0 / (2.0 * p.dx);
    a.an.dx = p.dx;
    a.led = ctx->led_on ? ctx->led : nullptr;
    return a;
}

// Solute ledger reset: clears the edge records and the pull-back sum, and
// takes the current interior mass as the reference (nothing while the
// ledger is off).
int ledger_reset(rs_ctx* ctx) {
    if (!ctx->led_on) return RS_OK;
    HIPCHK(ctx, hipMemsetAsync(ctx->led, 0, sizeof(double) * 2 * (size_t)(ctx->p.Ny + ctx->p.Nx), ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(ctx->led_acc, 0, sizeof(double), ctx->stream));
    HIPCHK(ctx, launch_mass(ctx->c[ctx->cur], ctx->p.Nx, ctx->p.Ny, ctx->rowd, ctx->led_acc + 1,
                            ctx->stream));
    return RS_OK;
}

// A shift that would move the fine window off the coarse T sample is
// rejected before anything changes (the context keeps its state).
int do_pull_back(rs_ctx* ctx, int off) {
    rs_constants& k = ctx->k;
    if (ctx->p.ifTFC) {
        const int ile = k.i_left_end_initial + (k.i_offset_history + off) -
                        (int)(k.l_offset_Vp / ctx->p.dx);
        const int rc = check_window(ctx, ile);
        if (rc) return rc;
        k.i_left_end = ile;
    }
    k.i_offset = off;
    k.i_offset_history += off;
    const int c0 = ctx->cur;
    // solute ledger: the interior mass before and after the shift
    if (ctx->led_on)
        HIPCHK(ctx, launch_mass(ctx->c[c0], ctx->p.Nx, ctx->p.Ny, ctx->rowd, ctx->led_acc + 2,
                                ctx->stream));
    if (ctx->mixed)
        HIPCHK(ctx, launch_pull_back_f32(ctx->psi32[c0], ctx->c[c0], ctx->psi32[c0 ^ 1],
                                         ctx->c[c0 ^ 1], ctx->p.Nx, ctx->p.Ny, off, ctx->p.cinf,
                                         ctx->stream));
    else
        HIPCHK(ctx, launch_pull_back(ctx->psi[c0], ctx->c[c0], ctx->psi[c0 ^ 1], ctx->c[c0 ^ 1],
                                     ctx->p.Nx, ctx->p.Ny, off, ctx->p.cinf, ctx->stream));
    if (ctx->led_on) {
        HIPCHK(ctx, launch_mass(ctx->c[c0 ^ 1], ctx->p.Nx, ctx->p.Ny, ctx->rowd, ctx->led_acc + 3,
                                ctx->stream));
        HIPCHK(ctx, launch_ledger_pull(ctx->led_acc + 2, ctx->led_acc, ctx->stream));
    }
    ctx->cur ^= 1;
    return RS_OK;
}

int do_tip_column(rs_ctx* ctx, int* it) {
    if (ctx->mixed)
        HIPCHK(ctx, launch_tip_column_f32(ctx->psi32[ctx->cur], ctx->p.Nx, ctx->p.Ny,
                                          (float)ctx->p.phi_interface, ctx->rowi, ctx->outi,
                                          ctx->stream));
    else
        HIPCHK(ctx, launch_tip_column(ctx->psi[ctx->cur], ctx->p.Nx, ctx->p.Ny, ctx->p.phi_interface,
                                      ctx->rowi, ctx->outi, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->h_pin, ctx->outi, sizeof(int), hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    memcpy(it, ctx->h_pin, sizeof(int));
    return RS_OK;
}

// FP32 copies of both coarse T buffers (mixed path), spare row included
int refresh_T32(rs_ctx* ctx) {
    if (!ctx->mixed) return RS_OK;
    const int64_t n = ctx->nT + ctx->k.NxT_all;
    for (int q = 0; q < 2; ++q)
        HIPCHK(ctx, launch_T32_copy(ctx->T[q], ctx->T32[q], n, kMixedTref, ctx->stream));
    return RS_OK;
}

StepArgsMixed step_args_mixed(rs_ctx* ctx) {
    const rs_params& p = ctx->p;
    const rs_constants& k = ctx->k;
    StepArgsMixed a{};
    a.psi = ctx->psi32[ctx->cur];
    a.c = ctx->c[ctx->cur];
    a.psi_n = ctx->psi32[ctx->cur ^ 1];
    a.c_n = ctx->c[ctx->cur ^ 1];
    a.lh = reinterpret_cast<float*>(ctx->lh);  // FP32 sums in the FP64-sized buffer
    a.T32 = ctx->T32[ctx->tcur];
    a.Nx = p.Nx;
    a.Ny = p.Ny;
    a.band_h = ctx->band_h;
    a.nbands = (p.Ny - 2 + a.band_h - 1) / a.band_h;
    a.lh_last_row = p.ratio_dxT_dx * (k.NyT - 2);
    a.lh_on = p.ifTFC;
    a.tfc = p.ifTFC;
    a.NxT_all = k.NxT_all;
    a.ile = k.i_left_end_interp;
    a.dxTi = (int)lround(p.dx * p.ratio_dxT_dx);
    a.has_L1 = (p.L1_liq != 0.0 || p.L1_bcc != 0.0) ? 1 : 0;
    // frozen T: T(i) = Tmelt + dT0 (d0 - dl) + GG dx S W0 (i + ioff), as T - Tref
    const double gstep = p.GG * p.dx * p.S * p.W0;
    const int ioff = -p.i_tip_target + k.i_offset_history - k.i_offset_history_initial;
    const double dl = p.GG * (k.l_offset_Vp - k.l_offset_Vp_initial) * p.S * p.W0 / p.deltaT0;
    a.fta_a = p.Tmelt + p.deltaT0 * (k.delta_initial - dl) + gstep * ioff - kMixedTref;
    a.fta_b = gstep;
    a.dt = (float)k.dt;
    a.AA = (float)p.AA;
    a.q_solid = (float)p.q_solid;
    a.dG_den_inv = (float)(1.0 / ctx->d.dG_den);
    a.eps1 = (float)p.eps1;
    a.epk1 = (float)p.epk1;
    a.lap_w_axis = (float)p.lap_w_axis;
    a.lap_w_diag = (float)p.lap_w_diag;
    a.inv2dx = (float)(1.0 / (2.0 * p.dx));
    a.invdx2 = (float)(1.0 / (p.dx * p.dx));
    a.inv4dx2 = (float)(1.0 / (4.0 * p.dx * p.dx));
    a.inv2dx2 = (float)(1.0 / (2.0 * p.dx * p.dx));
    const int oi[6] = {0, 1, 3, 4, 6, 7};
    for (int q = 0; q < 6; ++q) a.ops[q] = (float)ctx->orient[oi[q]];
    a.K = ctx->d.kdc / (p.dx * p.dx) * ctx->d.mu_den;
    a.L0l = p.L0_liq;
    a.L1l = p.L1_liq;
    a.RR = p.RR;
    a.dL0 = (float)(p.L0_liq - p.L0_bcc);
    a.dL1 = (float)(p.L1_liq - p.L1_bcc);
    a.rk = rk_of(p);
    a.led = ctx->led_on ? ctx->led : nullptr;
    a.tiles_desc = 2 * p.i_tip_target > p.Nx ? 1 : 0;
    return a;
}

int init_common(rs_ctx* ctx, const std::vector<double>& x0, int mode) {
    const rs_params& p = ctx->p;
    ctx->k = rs_constants{};
    rs_derive(&p, &ctx->k);
    ctx->cur = 0;
    ctx->tcur = 0;
    orientation(ctx->orient, ctx->p.alpha0, ctx->p.beta0, ctx->p.gamma0);
    HIPCHK(ctx, hipMemcpyAsync(ctx->x0, x0.data(), sizeof(double) * p.Ny, hipMemcpyHostToDevice,
                               ctx->stream));
    if (ctx->mixed)
        HIPCHK(ctx, launch_init_fields_f32(ctx->psi32[0], ctx->c[0], ctx->x0, p.Nx, p.Ny, p.dx,
                                           mode, p.cinf, p.c_liq0, p.c_sol0, ctx->stream));
    else
        HIPCHK(ctx, launch_init_fields(ctx->psi[0], ctx->c[0], ctx->x0, p.Nx, p.Ny, p.dx, mode,
                                       p.cinf, p.c_liq0, p.c_sol0, ctx->stream));
    if (p.ifTFC) {
        // NyT_all rows plus the spare row the bilinear read may touch
        HIPCHK(ctx, launch_init_T(ctx->T[0], ctx->T[1], ctx->k.NxT_all, ctx->k.NyT_all + 1,
                                  ctx->k.Tcold, p.S, p.W0, p.dx, p.ratio_dxT_dx, p.GG, ctx->stream));
        int rc = refresh_T32(ctx);
        if (rc) return rc;
    }
    int rc = ledger_reset(ctx);
    if (rc) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->k.i_left_end_interp = ctx->k.i_left_end;
    return check_window(ctx, ctx->k.i_left_end);
}

// coarse-T kernel arguments for the step `step` (the T/Tn/T32n pointers
// are set by the caller per sub-step)
TempArgs temp_args(rs_ctx* ctx, int64_t step) {
    const rs_params& p = ctx->p;
    const rs_constants& k = ctx->k;
    const double tau0 = tau0_of(p);
    TempArgs t;
    // the FP32 paths keep their latent-heat column sums in FP32
    t.lh = ctx->mixed ? nullptr : ctx->lh;
    t.lh32 = ctx->mixed ? reinterpret_cast<const float*>(ctx->lh) : nullptr;
    t.Nx = p.Nx;
    t.NxT = k.NxT;
    t.NyT = k.NyT;
    t.NxT_all = k.NxT_all;
    t.ratio = p.ratio_dxT_dx;
    t.ile = k.i_left_end;
    t.ratio_dt_dtT = k.ratio_dt_dtT;
    t.vadv = vp_ramp(p, step * tau0 * k.dt) * tau0 / (p.S * p.W0);
    t.dxT = p.dx * p.ratio_dxT_dx;
    t.dtT = k.dtT;
    t.DT = p.DT;
    t.Gmu = p.Gamma * p.muk0;
    t.Lcp = ctx->d.Lcp;
    t.Tcold = k.Tcold;
    t.Thot = k.Thot;
    t.Tref = kMixedTref;
    t.T = nullptr;
    t.Tn = nullptr;
    t.T32n = nullptr;
    return t;
}

void note_variant(rs_ctx* ctx, int v, int ident) {
    if (v == ctx->last_variant && ident == ctx->last_ident) {
        ++ctx->variant_run;
    } else {
        ctx->last_variant = v;
        ctx->last_ident = ident;
        ctx->variant_run = 1;
    }
}

uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Finite-value sentinel over the current psi, c and (coupled T) coarse T:
// RS_ERR_NONFINITE naming the first field with a NaN or Inf, its count and
// the first such cell in row-major order.
int finite_check(rs_ctx* ctx) {
    if (!ctx->finite_out)
        HIPCHK(ctx, hipMalloc(&ctx->finite_out, 6 * sizeof(unsigned long long)));
    const unsigned long long init[6] = {0ull, ~0ull, 0ull, ~0ull, 0ull, ~0ull};
    unsigned long long h[6];
    HIPCHK(ctx, hipMemcpyAsync(ctx->finite_out, init, sizeof init, hipMemcpyHostToDevice,
                               ctx->stream));
    const int c0 = ctx->cur;
    if (ctx->mixed)
        HIPCHK(ctx, launch_nonfinite_f32(ctx->psi32[c0], ctx->n, ctx->finite_out, ctx->stream));
    else
        HIPCHK(ctx, launch_nonfinite(ctx->psi[c0], ctx->n, ctx->finite_out, ctx->stream));
    HIPCHK(ctx, launch_nonfinite(ctx->c[c0], ctx->n, ctx->finite_out + 2, ctx->stream));
    if (ctx->p.ifTFC)
        HIPCHK(ctx, launch_nonfinite(ctx->T[ctx->tcur], ctx->nT, ctx->finite_out + 4, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(h, ctx->finite_out, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const char* names[3] = {"psi", "c", "coarse T"};
    for (int f = 0; f < 3; ++f) {
        if (h[2 * f] == 0) continue;
        const int64_t w = f == 2 ? ctx->k.NxT_all : ctx->p.Nx;
        const int64_t q = (int64_t)h[2 * f + 1];
        char b[256];
        snprintf(b, sizeof b,
                 "non-finite value after step %lld: %s has %llu such cells, the first at "
                 "(i=%lld, j=%lld)",
                 (long long)ctx->k.step, names[f], h[2 * f], (long long)(q % w), (long long)(q / w));
        return fail(ctx, RS_ERR_NONFINITE, b);
    }
    return RS_OK;
}

}  // namespace

namespace rs {
void set_thread_error(const std::string& why) { g_create_err = why; }
int solute_ledger_reset(rs_ctx* ctx) { return ledger_reset(ctx); }
StepArgsMixed mixed_step_args(rs_ctx* ctx) { return step_args_mixed(ctx); }
TempArgs coarse_T_args(rs_ctx* ctx, int64_t step) { return temp_args(ctx, step); }
double tau0_value(const rs_params& p) { return tau0_of(p); }
}  // namespace rs

extern "C" {

int rs_abi_info(int32_t* s) {
    if (!s) return RS_ERR_ARG;
    s[0] = (int32_t)sizeof(rs_params);
    s[1] = (int32_t)sizeof(rs_constants);
    s[2] = (int32_t)sizeof(rs_tip_row);
    s[3] = RS_ABI_VERSION;
    return RS_OK;
}

int rs_create(const rs_params* p, rs_ctx** out) {
    if (!p || !out) return create_fail(RS_ERR_ARG, "null argument");
    *out = nullptr;
    std::string why;
    int rc = check_params(*p, &why);
    if (rc) return create_fail(rc, why);
    rs_ctx* ctx = new (std::nothrow) rs_ctx;
    if (!ctx) return create_fail(RS_ERR_NOMEM, "host allocation failed");
    ctx->p = *p;
    rs_derive(p, &ctx->k);
    ctx->d = derive_constants(*p, ctx->k.dt);
    orientation(ctx->orient, ctx->p.alpha0, ctx->p.beta0, ctx->p.gamma0);
    ctx->n = (int64_t)p->Nx * p->Ny;
    ctx->nT = (int64_t)ctx->k.NxT_all * ctx->k.NyT_all;
    // the FP32 exact-table kernel takes 20-row bands in workgroups of two
    // wavefronts: a 4096^2 step on a developed front takes 127 us so, 134 us
    // with four wavefronts, 130 us with one, 154 us with 40-row bands and 137
    // us with 10 (its launch ends with its slowest workgroups, and shorter
    // bands finish sooner; below 20 rows the two halo rows of each band cost
    // more, and the XCD strips are whole band groups: finer groups even them)
    // the FP64 kernel (three wavefronts per SIMD) likewise: 537 us with
    // 20-row bands in two-wavefront workgroups, 545 us with 40-row bands
    // in four, 596 us with 60
    if (p->precision == RS_PREC_MIXED || p->precision == RS_PREC_FP64) {
        ctx->band_h = 20;
        ctx->wpb = 2;
    }
    if (const char* e = getenv("RS_BAND_H")) {
        const int v = atoi(e);
        if (v >= 10 && v <= 2000 && v % 10 == 0) ctx->band_h = v;
    }
    if (const char* e = getenv("RS_WPB")) {
        const int v = atoi(e);
        if (v >= 1 && v <= 4) ctx->wpb = v;
    }
    if (const char* e = getenv("RS_SURR_ALWAYS")) {
        const int v = atoi(e);
        if (v == 0 || v == 1) ctx->surr_always = v;
    }
    hipError_t e;
    if ((e = hipGetDevice(&ctx->device)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess) {
        free_all(ctx);
        delete ctx;
        return create_fail(RS_ERR_HIP, std::string("no usable HIP device: ") + hipGetErrorString(e));
    }
    const size_t fb = sizeof(double) * (size_t)ctx->n;
    // coarse T: NxT_all x NyT_all plus one spare row (the bilinear read of
    // row jT+1 reaches row NyT_all, with weight 0, when NyT % 4 == 0)
    const size_t tb = sizeof(double) * (size_t)(ctx->nT + ctx->k.NxT_all);
    const size_t lhb = sizeof(double) * (size_t)p->Nx * (size_t)(ctx->k.NyT > 2 ? ctx->k.NyT - 2 : 1);
    ctx->mixed = p->precision != RS_PREC_FP64;
    ctx->surrogate = p->precision == RS_PREC_SURROGATE;
    const size_t fb32 = sizeof(float) * (size_t)ctx->n;
    const size_t tb32 = sizeof(float) * (size_t)(ctx->nT + ctx->k.NxT_all);
    bool ok = true;
    if (ctx->mixed) {
        ok = hipMalloc(&ctx->psi32[0], fb32) == hipSuccess &&
             hipMalloc(&ctx->psi32[1], fb32) == hipSuccess &&
             hipMalloc(&ctx->T32[0], tb32) == hipSuccess && hipMalloc(&ctx->T32[1], tb32) == hipSuccess &&
             hipMemsetAsync(ctx->psi32[0], 0, fb32, ctx->stream) == hipSuccess &&
             hipMemsetAsync(ctx->psi32[1], 0, fb32, ctx->stream) == hipSuccess &&
             hipMemsetAsync(ctx->T32[0], 0, tb32, ctx->stream) == hipSuccess &&
             hipMemsetAsync(ctx->T32[1], 0, tb32, ctx->stream) == hipSuccess;
    } else {
        ok = hipMalloc(&ctx->psi[0], fb) == hipSuccess && hipMalloc(&ctx->psi[1], fb) == hipSuccess &&
             hipMemsetAsync(ctx->psi[0], 0, fb, ctx->stream) == hipSuccess &&
             hipMemsetAsync(ctx->psi[1], 0, fb, ctx->stream) == hipSuccess;
    }
    ok = ok && hipMalloc(&ctx->c[0], fb) == hipSuccess && hipMalloc(&ctx->c[1], fb) == hipSuccess &&
              hipMalloc(&ctx->T[0], tb) == hipSuccess && hipMalloc(&ctx->T[1], tb) == hipSuccess &&
              hipMalloc(&ctx->lh, lhb) == hipSuccess &&
              hipMalloc(&ctx->rowd, sizeof(double) * p->Ny) == hipSuccess &&
              hipMalloc(&ctx->rowi, sizeof(int) * p->Ny) == hipSuccess &&
              hipMalloc(&ctx->outd, sizeof(double) * 8) == hipSuccess &&
              hipMalloc(&ctx->outi, sizeof(int) * 8) == hipSuccess &&
              hipMalloc(&ctx->x0, sizeof(double) * p->Ny) == hipSuccess &&
              hipMalloc(&ctx->led, sizeof(double) * 2 * (size_t)(p->Ny + p->Nx)) == hipSuccess &&
              hipMalloc(&ctx->led_acc, sizeof(double) * 4) == hipSuccess &&
              hipHostMalloc(&ctx->h_pin, 64) == hipSuccess;
    if (!ok) {
        free_all(ctx);
        delete ctx;
        return create_fail(RS_ERR_NOMEM, "device allocation of the fields failed");
    }
    // defined contents everywhere (ghosts, padding, unused latent-heat rows)
    ok = hipMemsetAsync(ctx->c[0], 0, fb, ctx->stream) == hipSuccess &&
         hipMemsetAsync(ctx->c[1], 0, fb, ctx->stream) == hipSuccess &&
         hipMemsetAsync(ctx->T[0], 0, tb, ctx->stream) == hipSuccess &&
         hipMemsetAsync(ctx->T[1], 0, tb, ctx->stream) == hipSuccess &&
         hipMemsetAsync(ctx->lh, 0, lhb, ctx->stream) == hipSuccess &&
         hipMemsetAsync(ctx->led, 0, sizeof(double) * 2 * (size_t)(p->Ny + p->Nx), ctx->stream) == hipSuccess &&
         hipMemsetAsync(ctx->led_acc, 0, sizeof(double) * 4, ctx->stream) == hipSuccess &&
         hipStreamSynchronize(ctx->stream) == hipSuccess;
    if (!ok) {
        free_all(ctx);
        delete ctx;
        return create_fail(RS_ERR_HIP, "clearing the new fields failed");
    }
    g_create_err.clear();
    *out = ctx;
    return RS_OK;
}

int rs_context_info(rs_ctx* ctx, double* out, int32_t n) {
    if (!ctx || !out || n < 1) return RS_ERR_ARG;
    double lo = 0.0, hi = 0.0, err = 0.0;
    if (ctx->mixed) sgte_table_info(&lo, &hi, &err);
    const double v[9] = {err, ctx->mixed ? kMixedTref : 0.0, lo, hi, (double)ctx->band_h,
                         (double)ctx->wpb, (double)ctx->last_variant, (double)ctx->last_ident,
                         (double)ctx->variant_run};
    for (int q = 0; q < n && q < 9; ++q) out[q] = v[q];
    return RS_OK;
}

int rs_destroy(rs_ctx* ctx) {
    if (!ctx) return RS_ERR_ARG;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
    return RS_OK;
}

const char* rs_last_error(const rs_ctx* ctx) {
    if (ctx) return ctx->err.c_str();
    return g_create_err.empty() ? "null context" : g_create_err.c_str();
}

int rs_device_name(char* buf, int32_t len) {
    if (!buf || len < 1) return RS_ERR_ARG;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return RS_ERR_HIP;
    snprintf(buf, (size_t)len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName,
             prop.multiProcessorCount);
    return RS_OK;
}

int rs_init_reference(rs_ctx* ctx) {
    if (!ctx) return RS_ERR_ARG;
    const rs_params& p = ctx->p;
    std::vector<double> x0(p.Ny);
    const double PI_REF = 3.1415927;  // fun.cu:59
    for (int j = 0; j < p.Ny; ++j) x0[j] = p.i_tip_target + 0.5 * sin(2.0 * PI_REF * j / p.Ny);
    return init_common(ctx, x0, 0);
}

int rs_init_config(rs_ctx* ctx, uint64_t seed) {
    if (!ctx) return RS_ERR_ARG;
    const rs_params& p = ctx->p;
    std::vector<double> x0(p.Ny);
    uint64_t st = seed;
    for (int j = 1; j <= p.Ny - 2; ++j) {
        const double u = (double)(splitmix64(&st) >> 11) * 0x1.0p-53;
        x0[j] = p.i_tip_target + (2.0 * u - 1.0);
    }
    x0[0] = x0[p.Ny - 2];
    x0[p.Ny - 1] = x0[1];
    return init_common(ctx, x0, 1);
}

int rs_upload(rs_ctx* ctx, int32_t field, const double* host, int64_t count) {
    if (!ctx || !host) return RS_ERR_ARG;
    std::vector<float> f32;
    switch (field) {
    case RS_FIELD_C:
    case RS_FIELD_PSI:
        if (count != ctx->n) return fail(ctx, RS_ERR_ARG, "field size mismatch");
        if (field == RS_FIELD_PSI && ctx->mixed) {
            f32.resize((size_t)count);
            for (int64_t q = 0; q < count; ++q) f32[q] = (float)host[q];
            HIPCHK(ctx, hipMemcpyAsync(ctx->psi32[ctx->cur], f32.data(), sizeof(float) * count,
                                       hipMemcpyHostToDevice, ctx->stream));
        } else {
            HIPCHK(ctx, hipMemcpyAsync(field == RS_FIELD_C ? ctx->c[ctx->cur] : ctx->psi[ctx->cur],
                                       host, sizeof(double) * count, hipMemcpyHostToDevice,
                                       ctx->stream));
        }
        break;
    case RS_FIELD_T: {
        if (count != ctx->nT) return fail(ctx, RS_ERR_ARG, "T size mismatch");
        HIPCHK(ctx, hipMemcpyAsync(ctx->T[0], host, sizeof(double) * count, hipMemcpyHostToDevice,
                                   ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(ctx->T[1], host, sizeof(double) * count, hipMemcpyHostToDevice,
                                   ctx->stream));
        ctx->tcur = 0;
        int rc = refresh_T32(ctx);
        if (rc) return rc;
        break;
    }
    default:
        return fail(ctx, RS_ERR_ARG, "field cannot be uploaded");
    }
    if (field == RS_FIELD_C) {  // a new composition field: the ledger starts over
        const int rc = ledger_reset(ctx);
        if (rc) return rc;
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return RS_OK;
}

int rs_download(rs_ctx* ctx, int32_t field, double* host, int64_t count) {
    if (!ctx || !host) return RS_ERR_ARG;
    const int64_t need = (field == RS_FIELD_T) ? ctx->nT : ctx->n;
    if (count != need) return fail(ctx, RS_ERR_ARG, "field size mismatch");
    const double* src = nullptr;
    if ((field == RS_FIELD_PSI || field == RS_FIELD_PHI) && ctx->mixed) {
        std::vector<float> f32((size_t)count);
        HIPCHK(ctx, hipMemcpyAsync(f32.data(), ctx->psi32[ctx->cur], sizeof(float) * count,
                                   hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        for (int64_t q = 0; q < count; ++q)
            host[q] = field == RS_FIELD_PHI ? fmax((double)f32[q], -1.0) : (double)f32[q];
        return RS_OK;
    }
    switch (field) {
    case RS_FIELD_C: src = ctx->c[ctx->cur]; break;
    case RS_FIELD_PSI:
    case RS_FIELD_PHI: src = ctx->psi[ctx->cur]; break;
    case RS_FIELD_T: src = ctx->T[ctx->tcur]; break;
    case RS_FIELD_TFINE: {
        if (!ctx->tfine) HIPCHK(ctx, hipMalloc(&ctx->tfine, sizeof(double) * (size_t)ctx->n));
        TempSrc ts = make_tempsrc(ctx->p, ctx->k, ctx->T[ctx->tcur], ctx->k.i_left_end_interp);
        HIPCHK(ctx, launch_interp(ts, ctx->tfine, ctx->p.Nx, ctx->p.Ny, ctx->stream));
        src = ctx->tfine;
        break;
    }
    default: return fail(ctx, RS_ERR_ARG, "bad field");
    }
    HIPCHK(ctx, hipMemcpyAsync(host, src, sizeof(double) * count, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (field == RS_FIELD_PHI)
        for (int64_t q = 0; q < count; ++q) host[q] = fmax(host[q], -1.0);
    return RS_OK;
}

int rs_get_constants(rs_ctx* ctx, rs_constants* k) {
    if (!ctx || !k) return RS_ERR_ARG;
    *k = ctx->k;
    return RS_OK;
}

int rs_set_constants(rs_ctx* ctx, const rs_constants* k) {
    if (!ctx || !k) return RS_ERR_ARG;
    // every rule is checked before anything changes: the buffers were sized
    // for the context's coarse grid, the loop takes step % cadence, and the
    // fine window must interpolate from interior coarse columns
    if (k->NxT_all != ctx->k.NxT_all || k->NyT_all != ctx->k.NyT_all || k->NxT != ctx->k.NxT ||
        k->NyT != ctx->k.NyT)
        return fail(ctx, RS_ERR_ARG, "coarse grid geometry cannot change");
    if (k->step_fields <= 0 || k->step_tip <= 0 || k->step_check_point <= 0 || k->step_pull_back <= 0)
        return fail(ctx, RS_ERR_ARG, "output and pull-back cadences must be positive");
    if (k->ratio_dt_dtT < 1) return fail(ctx, RS_ERR_ARG, "ratio_dt_dtT must be at least 1");
    if (!(k->dt > 0.0) || !std::isfinite(k->dt) || !(k->dtT > 0.0) || !std::isfinite(k->dtT))
        return fail(ctx, RS_ERR_ARG, "dt and dtT must be positive and finite");
    if (k->step < 0) return fail(ctx, RS_ERR_ARG, "step must be non-negative");
    if (ctx->p.ifTFC) {
        const int r = ctx->p.ratio_dxT_dx;
        for (int ile : {k->i_left_end, k->i_left_end_interp})
            if (ile - 1 < 0 || (ctx->p.Nx - 2 + ile) / r + 1 > ctx->k.NxT - 1) {
                char b[256];
                snprintf(b, sizeof b,
                         "i_left_end / i_left_end_interp = %d puts the fine window off the "
                         "temperature sample (NxT=%d)",
                         ile, ctx->k.NxT);
                return fail(ctx, RS_ERR_ARG, b);
            }
    }
    ctx->k = *k;
    ctx->d = derive_constants(ctx->p, ctx->k.dt);
    return RS_OK;
}

int rs_tip_column(rs_ctx* ctx, int32_t* i_tip) {
    if (!ctx || !i_tip) return RS_ERR_ARG;
    int it = 0;
    int rc = do_tip_column(ctx, &it);
    if (rc) return rc;
    *i_tip = it;
    return RS_OK;
}

int rs_pull_back(rs_ctx* ctx, int32_t i_offset) {
    if (!ctx) return RS_ERR_ARG;
    return do_pull_back(ctx, i_offset);
}

int rs_step(rs_ctx* ctx, int64_t nsteps) {
    if (!ctx || nsteps < 0) return RS_ERR_ARG;
    const rs_params& p = ctx->p;
    rs_constants& k = ctx->k;
    const double tau0 = tau0_of(p);
    for (int64_t q = 0; q < nsteps; ++q) {
        const int64_t step = k.step;
        if (step % k.step_pull_back == 0) {  // fun.cu:448 (includes step 0)
            int it = 0;
            int rc = do_tip_column(ctx, &it);
            if (rc) return rc;
            rc = do_pull_back(ctx, it - p.i_tip_target);
            if (rc) return rc;
        }
        // every timing-th step: the dispatches of its step kernel (with the
        // antitrapping pass) and of its coarse-T kernels record their times
        const bool timed = ctx->timing > 0 && ctx->timing_ctr++ % ctx->timing == 0;
        LaunchEv ev_step, ev_T;
        if (timed) {
            int rc = ev_pair(ctx, 0, &ev_step);
            if (!rc && p.ifTFC) rc = ev_pair(ctx, 1, &ev_T);
            if (rc) return rc;
        }
        const bool at_on = p.at_coef != 0.0;
        const LaunchEv ev_k{ev_step.start, at_on ? nullptr : ev_step.stop};
        const LaunchEv ev_at{nullptr, ev_step.stop};
        if (ctx->mixed) {
            StepArgsMixed a = step_args_mixed(ctx);
            if (ctx->surrogate) {
                if (!surr_ready(ctx))
                    return fail(ctx, RS_ERR_STATE, "surrogate path needs trained weights");
                // after the first 64 steps and then every 1024
                if (ctx->surr_always < 0 && ctx->surr_steps++ % 1024 == 64) {
                    const int rc = surr_update_policy(ctx);
                    if (rc) return rc;
                }
                surr_step_args(ctx, a);
            }
            const double* o = ctx->orient;
            const bool ident = o[0] == 1.0 && o[1] == 0.0 && o[3] == 0.0 && o[4] == 1.0 &&
                               o[6] == 0.0 && o[7] == 0.0;
            int v = 0;
            HIPCHK(ctx, launch_step_mixed(a, ctx->wpb, ident, ctx->surrogate, ctx->stream, &v, ev_k));
            note_variant(ctx, v, ident ? 1 : 0);
        } else {
            const StepArgsF64 a = step_args(ctx);
            HIPCHK(ctx, launch_step_fp64(a, ctx->wpb, ctx->stream, ev_k));
            note_variant(ctx, -1, 1);
        }
        if (at_on) {
            const double vdn = p.V_D > 0.0 ? p.V_D * tau0 / (p.S * p.W0) : 0.0;
            const int c0 = ctx->cur;
            if (ctx->mixed)
                HIPCHK(ctx, launch_antitrap_f32(ctx->psi32[c0], ctx->psi32[c0 ^ 1], ctx->c[c0],
                                                ctx->c[c0 ^ 1], p.Nx, p.Ny, p.dx, k.dt, p.at_coef,
                                                p.k_part, vdn, ctx->stream, ev_at));
            else
                HIPCHK(ctx, launch_antitrap(ctx->psi[c0], ctx->psi[c0 ^ 1], ctx->c[c0], ctx->c[c0 ^ 1],
                                            p.Nx, p.Ny, p.dx, k.dt, p.at_coef, p.k_part, vdn,
                                            ctx->stream, ev_at));
        }
        if (p.ifTFC) {
            TempArgs t = temp_args(ctx, step);
            for (int sub = 0; sub < k.ratio_dt_dtT; ++sub) {
                t.T = ctx->T[ctx->tcur];
                t.Tn = ctx->T[ctx->tcur ^ 1];
                t.T32n = ctx->mixed ? ctx->T32[ctx->tcur ^ 1] : nullptr;
                const LaunchEv ev{sub == 0 ? ev_T.start : nullptr,
                                  sub == k.ratio_dt_dtT - 1 ? ev_T.stop : nullptr};
                HIPCHK(ctx, launch_temperature(t, ctx->stream, ev));
                ctx->tcur ^= 1;
            }
            k.i_left_end_interp = k.i_left_end;
        }
        ctx->cur ^= 1;
        k.l_offset_Vp += vp_ramp(p, step * tau0 * k.dt) * k.dt * tau0 / (p.S * p.W0);
        k.step = step + 1;
        if (ctx->finite_every > 0 && ++ctx->finite_ctr % ctx->finite_every == 0) {
            const int rc = finite_check(ctx);
            if (rc) return rc;
        }
    }
    return RS_OK;
}

int rs_solute_mass(rs_ctx* ctx, double* mass) {
    if (!ctx || !mass) return RS_ERR_ARG;
    HIPCHK(ctx, launch_mass(ctx->c[ctx->cur], ctx->p.Nx, ctx->p.Ny, ctx->rowd, ctx->outd, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->h_pin, ctx->outd, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *mass = ctx->h_pin[0];
    return RS_OK;
}

int rs_solute_ledger(rs_ctx* ctx, double* out, int32_t n, int32_t reset) {
    if (!ctx || !out || n < 1) return RS_ERR_ARG;
    if (!ctx->led_on) {
        if (!reset) return fail(ctx, RS_ERR_STATE, "the solute ledger is off (start it with reset != 0)");
        // start: the record begins at the current fields
        ctx->led_on = true;
        const int rc = ledger_reset(ctx);
        if (rc) return rc;
    }
    const int Ny = ctx->p.Ny, Nx = ctx->p.Nx;
    // outd: [0] mass now, [1..2] x faces, [3..4] y faces
    HIPCHK(ctx, launch_mass(ctx->c[ctx->cur], Nx, Ny, ctx->rowd, ctx->outd, ctx->stream));
    HIPCHK(ctx, launch_sum(ctx->led, 1, Ny - 2, ctx->outd + 1, ctx->stream));
    HIPCHK(ctx, launch_sum(ctx->led + Ny, 1, Ny - 2, ctx->outd + 2, ctx->stream));
    HIPCHK(ctx, launch_sum(ctx->led + 2 * Ny, 1, Nx - 2, ctx->outd + 3, ctx->stream));
    HIPCHK(ctx, launch_sum(ctx->led + 2 * Ny + Nx, 1, Nx - 2, ctx->outd + 4, ctx->stream));
    double h[5], acc[2];
    HIPCHK(ctx, hipMemcpyAsync(h, ctx->outd, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(acc, ctx->led_acc, sizeof acc, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const double ex = h[1] + h[2], ey = h[3] + h[4];
    const double v[6] = {h[0], acc[1], acc[0], ex, ey, ((h[0] - acc[1]) - acc[0]) - (ex + ey)};
    for (int q = 0; q < n && q < 6; ++q) out[q] = v[q];
    if (reset) {
        const int rc = ledger_reset(ctx);
        if (rc) return rc;
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    return RS_OK;
}

// set_tip_output (fun.cu:3024-3072) producing a tip.txt row
// (src/main.cu:2671-2709)
int rs_tip(rs_ctx* ctx, rs_tip_row* t) {
    if (!ctx || !t) return RS_ERR_ARG;
    const rs_params& p = ctx->p;
    rs_constants& k = ctx->k;
    // per-row fit data from the GPU (36 B per row), not the field
    if (!ctx->tipvals) {
        HIPCHK(ctx, hipMalloc(&ctx->tipvals, sizeof(double) * 8 * (size_t)p.Ny));
        HIPCHK(ctx, hipMalloc(&ctx->tiprow, sizeof(int) * (size_t)p.Ny));
    }
    if (ctx->mixed)
        HIPCHK(ctx, launch_tip_detail_f32(ctx->psi32[ctx->cur], p.Nx, p.Ny, p.phi_interface,
                                          ctx->tiprow, ctx->tipvals, ctx->stream));
    else
        HIPCHK(ctx, launch_tip_detail(ctx->psi[ctx->cur], p.Nx, p.Ny, p.phi_interface, ctx->tiprow,
                                      ctx->tipvals, ctx->stream));
    std::vector<int> irow((size_t)p.Ny);
    std::vector<double> vals((size_t)p.Ny * 8);
    HIPCHK(ctx, hipMemcpyAsync(irow.data(), ctx->tiprow, sizeof(int) * p.Ny, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(vals.data(), ctx->tipvals, sizeof(double) * 8 * p.Ny,
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    int rc = RS_OK;
    TipFit f;
    const int st = precise_tip_rows(p, irow.data(), vals.data(), &f);
    const double tau0 = tau0_of(p);
    const double dt = k.dt;
    const double xnow = (k.i_offset_history + f.x) * p.dx;
    double V = (xnow - k.x_tip_history) * p.S * p.W0 / (k.step_tip * tau0 * dt);
    k.x_tip_history = xnow;
    if (V < 0.0) V = 0.01;
    if (V > 10.0) V = 10.0;
    double Ttip, delta;
    if (p.ifTFC) {
        // the reference re-interpolates T_fine with the current i_left_end and
        // samples row 0 (fun.cu:3052-3056)
        k.i_left_end_interp = k.i_left_end;
        const int r = p.ratio_dxT_dx;
        const int i = f.i, j = 0, ile = k.i_left_end;
        const double dxT = p.dx * r;
        const int iT = (i - 1 + ile) / r;
        const int jT = (int)((j - 1 + dxT) / r);
        const double x = (i - 1.0 + ile) / r - iT;
        const double y = (j - 1.0 + dxT) / r - jT;
        // the four coarse values of the bilinear read
        double q4[4];
        const double* Tbuf = ctx->T[ctx->tcur];
        const int64_t o0 = (int64_t)jT * k.NxT_all + iT, o1 = o0 + k.NxT_all;
        if (iT < 0 || iT + 1 >= k.NxT_all || jT < 0 || jT + 1 > k.NyT_all)
            return fail(ctx, RS_ERR_DOMAIN, "tip outside the coarse temperature grid");
        HIPCHK(ctx, hipMemcpyAsync(q4, Tbuf + o0, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(q4 + 2, Tbuf + o1, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                   ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        const double f00 = q4[0], f10 = q4[1], f01 = q4[2], f11 = q4[3];
        const double a10 = f10 - f00, a01 = f01 - f00, a11 = f11 - f10 - f01 + f00;
        Ttip = f00 + a10 * x + a01 * y + a11 * x * y;
        delta = (Ttip - p.Tmelt) / p.deltaT0;
    } else {
        delta = k.delta_initial +
                p.GG * (f.i - p.i_tip_target + k.i_offset_history - k.i_offset_history_initial) * p.dx *
                    p.S * p.W0 / p.deltaT0 -
                p.GG * (k.l_offset_Vp - k.l_offset_Vp_initial) * p.S * p.W0 / p.deltaT0;
        Ttip = delta * p.deltaT0 + p.Tmelt;
    }
    k.delta = delta;
    if (k.if_begin_tip == 1) {
        k.if_begin_tip = 0;
        V = NAN;
    }
    t->step = k.step;
    t->t_us = k.step * dt * tau0 * 1e-3;
    t->delta = delta;
    t->T_tip = Ttip;
    t->V_tip = V;
    t->Vp = vp_ramp(p, k.step * tau0 * dt);
    t->R = f.R;
    t->i = f.i;
    t->j = f.j;
    t->x = f.x;
    t->y = f.y;
    return st == 0 ? RS_OK : fail(ctx, RS_ERR_DOMAIN, "tip parabola fit failed");
}

int rs_sync(rs_ctx* ctx) {
    if (!ctx) return RS_ERR_ARG;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return RS_OK;
}

int rs_timing_enable(rs_ctx* ctx, int32_t every) {
    if (!ctx || every < 0) return RS_ERR_ARG;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->timing = every;
    ctx->timing_ctr = 0;
    for (int q = 0; q < 3; ++q) ctx->nev[q] = 0;
    return RS_OK;
}

int rs_set_finite_check(rs_ctx* ctx, int32_t every) {
    if (!ctx || every < 0) return RS_ERR_ARG;
    ctx->finite_every = every;
    ctx->finite_ctr = 0;
    return RS_OK;
}

int rs_calibrate_copy(rs_ctx* ctx, int64_t n, int32_t reps, double* avg_ms) {
    if (!ctx || !avg_ms || n < 1 || n > ctx->n || reps < 1) return RS_ERR_ARG;
    hipEvent_t e0, e1;
    HIPCHK(ctx, hipEventCreate(&e0));
    HIPCHK(ctx, hipEventCreate(&e1));
    double tot = 0.0;
    for (int q = 0; q < reps; ++q) {
        HIPCHK(ctx, hipEventRecord(e0, ctx->stream));
        // current c into the other buffer, which the next step overwrites
        HIPCHK(ctx, launch_copy(ctx->c[ctx->cur], ctx->c[ctx->cur ^ 1], n, ctx->stream));
        HIPCHK(ctx, hipEventRecord(e1, ctx->stream));
        HIPCHK(ctx, hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, e0, e1));
        tot += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = tot / reps;
    return RS_OK;
}

int rs_timing(rs_ctx* ctx, int32_t which, double* avg_ms, int64_t* count) {
    if (!ctx || !avg_ms || !count || which < 0 || which > 1) return RS_ERR_ARG;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    double tot = 0.0;
    int64_t nrec = 0;
    for (int q = 0; q + 1 < ctx->nev[which]; q += 2) {
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[which][q], ctx->ev[which][q + 1]));
        tot += ms;
        ++nrec;
    }
    *avg_ms = nrec ? tot / nrec : 0.0;
    *count = nrec;
    return RS_OK;
}

}  // extern "C"
