// rs_surr.hip -- host side of the learned CALPHAD surrogate: sample and
// curve set-up, the training loop (kernels in rs_mlp.hip), the trust table
// and the surrogate part of the C ABI.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "rs_host.hpp"
#include "rs_mlp.hpp"

using namespace rs;

namespace rs {

struct SurrState {
    float* P = nullptr;  // FP32 master weights
    float* m = nullptr;  // Adam moments
    float* v = nullptr;
    MlpFrags* F = nullptr;
    uint64_t* trust = nullptr;               // device copy, kTrustN words
    // kCountSlots slots of [0] fallback cells, [1] all cells, [2] rows
    // without the network, [3] all rows (one row = one wavefront's cells)
    unsigned long long* count = nullptr;
    unsigned long long seen_rows = 0, seen_skip = 0;  // [3], [2] at the last policy update
    bool always = false;  // evaluate the network in every row
    MlpNorm norm{};
    double T_lo = 0.0, T_hi = 0.0, tol = 1e-4;
    int fmt = kActFp16;  // inference activation format (training: BF16, split unless kActBf16)
    uint64_t h_trust[kTrustN] = {};
    // FNV-1a hash of the FP32 weights, the normalisation, the window, the
    // tolerance and the format: equal for networks that evaluate alike
    uint64_t ident = 0;
    bool ready = false;
};

}  // namespace rs

namespace {

// waits for the context stream before returning an error (see rs_capi.hip:
// no pending copy may outlive a host buffer of the failed call)
int fail(rs_ctx* ctx, int code, const char* what, hipError_t e = hipSuccess) {
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    char buf[512];
    if (e != hipSuccess)
        snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    else
        snprintf(buf, sizeof buf, "%s", what);
    ctx->err = buf;
    return code;
}

#define HIPCHK(ctx, call)                                                \
    do {                                                                 \
        hipError_t e_ = (call);                                          \
        if (e_ != hipSuccess) return fail((ctx), RS_ERR_HIP, #call, e_); \
    } while (0)

// device buffers that live for one call
struct Scratch {
    std::vector<void*> p;
    ~Scratch() {
        for (void* q : p) (void)hipFree(q);
    }
    template <typename T>
    hipError_t get(T** out, size_t count) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, sizeof(T) * std::max<size_t>(count, 1));
        if (e == hipSuccess) p.push_back(q);
        *out = static_cast<T*>(q);
        return e;
    }
};

int ensure_state(rs_ctx* ctx) {
    if (ctx->surr) return RS_OK;
    SurrState* s = new (std::nothrow) SurrState;
    if (!s) return fail(ctx, RS_ERR_NOMEM, "surrogate state");
    ctx->surr = s;
    HIPCHK(ctx, hipMalloc(&s->P, sizeof(float) * kNParam));
    HIPCHK(ctx, hipMalloc(&s->m, sizeof(float) * kNParam));
    HIPCHK(ctx, hipMalloc(&s->v, sizeof(float) * kNParam));
    HIPCHK(ctx, hipMalloc(&s->F, sizeof(MlpFrags)));
    HIPCHK(ctx, hipMalloc(&s->trust, sizeof(uint64_t) * kTrustN));
    HIPCHK(ctx, hipMalloc(&s->count, kCountBytes));
    HIPCHK(ctx, hipMemsetAsync(s->count, 0, kCountBytes, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(s->trust, 0, sizeof(uint64_t) * kTrustN, ctx->stream));
    return RS_OK;
}

uint64_t splitmix(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double unif(uint64_t* s) { return (double)(splitmix(s) >> 11) * 0x1.0p-53; }

// temperature range of the run's fine window (K): coarse field range in
// the coupled mode, the frozen profile over the grid otherwise
void run_T_range(rs_ctx* ctx, double* lo, double* hi) {
    const rs_params& p = ctx->p;
    const rs_constants& k = ctx->k;
    if (p.ifTFC) {
        *lo = k.Tcold;
        *hi = k.Thot;
    } else {
        const double gstep = p.GG * p.dx * p.S * p.W0;
        const int ioff = -p.i_tip_target + k.i_offset_history - k.i_offset_history_initial;
        const double dl = p.GG * (k.l_offset_Vp - k.l_offset_Vp_initial) * p.S * p.W0 / p.deltaT0;
        const double t0 = p.Tmelt + p.deltaT0 * (k.delta_initial - dl) + gstep * ioff;
        *lo = t0;
        *hi = t0 + gstep * (p.Nx - 1);
    }
}

// largest nondimensional residual of the quantities the stepper consumes
double residual(const Derived& d, const double* net, const double* ex) {
    const double e0 = fabs(net[0] - ex[0]) / d.dG_den;
    const double e2 = fabs(net[2] - ex[2]) * d.mu_den;
    const double e3 = fabs((net[2] + net[3]) - (ex[2] + ex[3])) * d.mu_den;
    return std::max(e0, std::max(e2, e3));
}

// The packed-FP16 format rounds the activation's input, clamp-FMA and
// product to FP16 each: its output noise varies between the validation
// nodes by up to 24% of the residual bound (measured), so its bins are
// judged against 0.8 of the bound, which keeps the points between the
// nodes within it.
constexpr double kPackedTolScale = 0.8;

// network and exact targets on the (8 kTrustN + 1)^2 validation grid; bins
// whose every grid point (edges included) meets the tolerance are trusted
int build_trust(rs_ctx* ctx, double* max_err4, double* max_res_trusted, double* frac) {
    SurrState* s = ctx->surr;
    const double tol = s->fmt == kActFp16Pk ? kPackedTolScale * s->tol : s->tol;
    const int G = 8 * kTrustN + 1;
    const int64_t n = (int64_t)G * G;
    std::vector<double> tc((size_t)(2 * n));
    for (int pb = 0; pb < G; ++pb)
        for (int pa = 0; pa < G; ++pa) {
            const double a = -1.0 + 2.0 * pa / (G - 1), b = -1.0 + 2.0 * pb / (G - 1);
            const int64_t q = (int64_t)pb * G + pa;
            tc[2 * q] = s->norm.t_mid + a * s->norm.t_half;
            tc[2 * q + 1] = 1.0 - (0.2 + 0.2 * b);
        }
    Scratch sc;
    double *dTc, *dY;
    float *dX, *dO;
    HIPCHK(ctx, sc.get(&dTc, 2 * n));
    HIPCHK(ctx, sc.get(&dY, 4 * n));
    HIPCHK(ctx, sc.get(&dX, 2 * n));
    HIPCHK(ctx, sc.get(&dO, 4 * n));
    HIPCHK(ctx, hipMemcpyAsync(dTc, tc.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, launch_mlp_inputs(dTc, n, s->norm.t_mid, s->norm.t_half, dX, ctx->stream));
    HIPCHK(ctx, launch_mlp_eval(dX, n, s->F, dO, s->fmt, ctx->stream));
    HIPCHK(ctx, launch_mlp_targets(dTc, n, ctx->p.L0_bcc, ctx->p.L0_liq, ctx->p.L1_bcc, ctx->p.L1_liq,
                                   dY, ctx->stream));
    std::vector<double> Y((size_t)(4 * n));
    std::vector<float> O((size_t)(4 * n));
    HIPCHK(ctx, hipMemcpyAsync(Y.data(), dY, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(O.data(), dO, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<double> res((size_t)n);
    for (int k = 0; k < 4; ++k) max_err4[k] = 0.0;
    for (int64_t q = 0; q < n; ++q) {
        double net[4];
        for (int k = 0; k < 4; ++k) {
            net[k] = (double)O[4 * q + k] * s->norm.y_std[k] + s->norm.y_mean[k];
            max_err4[k] = std::max(max_err4[k], fabs(net[k] - Y[4 * q + k]));
        }
        res[q] = residual(ctx->d, net, &Y[4 * q]);
    }
    int trusted = 0;
    *max_res_trusted = 0.0;
    for (int ib = 0; ib < kTrustN; ++ib) {
        uint64_t w = 0;
        for (int ia = 0; ia < kTrustN; ++ia) {
            double r = 0.0;
            for (int pb = 8 * ib; pb <= 8 * ib + 8; ++pb)
                for (int pa = 8 * ia; pa <= 8 * ia + 8; ++pa) r = std::max(r, res[(size_t)pb * G + pa]);
            if (r <= tol && std::isfinite(r)) {
                w |= 1ULL << ia;
                ++trusted;
                *max_res_trusted = std::max(*max_res_trusted, r);
            }
        }
        s->h_trust[ib] = w;
    }
    *frac = (double)trusted / (kTrustN * kTrustN);
    HIPCHK(ctx, hipMemcpyAsync(s->trust, s->h_trust, sizeof(uint64_t) * kTrustN, hipMemcpyHostToDevice,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return RS_OK;
}

// mean squared error (normalised outputs) of the network over a sample set
int mse_of(rs_ctx* ctx, const float* dX, const float* dYn, int64_t n, float* dO, double* mse) {
    HIPCHK(ctx, launch_mlp_eval(dX, n, ctx->surr->F, dO, ctx->surr->fmt, ctx->stream));
    std::vector<float> O((size_t)(4 * n)), Yn((size_t)(4 * n));
    HIPCHK(ctx, hipMemcpyAsync(O.data(), dO, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(Yn.data(), dYn, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    double acc = 0.0;
    for (int64_t q = 0; q < 4 * n; ++q) {
        const double e = (double)O[q] - (double)Yn[q];
        acc += e * e;
    }
    *mse = acc / (double)(4 * n);
    return RS_OK;
}

// the four counts summed over the slots
void sum_counts(const unsigned long long* slots, unsigned long long c[4]) {
    for (int k = 0; k < 4; ++k) c[k] = 0;
    for (int q = 0; q < kCountSlots; ++q)
        for (int k = 0; k < 4; ++k) c[k] += slots[(size_t)q * kCountStride + k];
}

// the row counts skip (rows without the network) and rows
// since the counters were last cleared; see surr_update_policy
void surr_policy_apply(SurrState* s, unsigned long long skip_all, unsigned long long rows_all) {
    if (rows_all < s->seen_rows || skip_all < s->seen_skip) s->seen_rows = s->seen_skip = 0;
    const unsigned long long rows = rows_all - s->seen_rows, skip = skip_all - s->seen_skip;
    if (rows < 4096) return;
    s->always = skip * 50 < rows;
    s->seen_rows = rows_all;
    s->seen_skip = skip_all;
}

}  // namespace

namespace rs {

void surr_free(rs_ctx* ctx) {
    SurrState* s = ctx->surr;
    if (!s) return;
    void* ps[] = {s->P, s->m, s->v, s->F, s->trust, s->count};
    for (void* q : ps)
        if (q) (void)hipFree(q);
    delete s;
    ctx->surr = nullptr;
}

bool surr_ready(const rs_ctx* ctx) { return ctx->surr && ctx->surr->ready; }

uint64_t surr_identity(const rs_ctx* ctx) { return surr_ready(ctx) ? ctx->surr->ident : 0; }

const uint64_t* surr_trust_host(const rs_ctx* ctx) {
    return surr_ready(ctx) ? ctx->surr->h_trust : nullptr;
}

void surr_step_args(rs_ctx* ctx, StepArgsMixed& a) {
    const SurrState* s = ctx->surr;
    a.mlp = fp16_fmt(s->fmt) ? &s->F->h : &s->F->f;
    a.trust = s->trust;
    a.fb_count = s->count;
    // a = (T - t_mid)/t_half with T = Tref + T'
    a.s_a_scale = (float)(1.0 / s->norm.t_half);
    a.s_a_off = (float)((kMixedTref - s->norm.t_mid) / s->norm.t_half);
    for (int k = 0; k < 4; ++k) {
        a.y_mean[k] = s->norm.y_mean[k];
        a.y_std[k] = s->norm.y_std[k];
    }
    a.s_fmt = s->fmt;
    a.s_always = ctx->surr_always >= 0 ? ctx->surr_always : s->always;
}

// Rows with no trusted cell skip the network behind a branch that costs ~4%
// of a step whenever such rows are rare (a front grown inside the sampled
// (T, c) box); once the rows counted since the last update number at least
// 4096 and fewer than 2% of them skipped the network, the steps evaluate it
// in every row (same results: such rows discard it), and go back to the
// branch when the share rises again. Waits for the stream (it reads the
// counts), so rs_step calls it once per 1024 surrogate steps.
int surr_row_counts(rs_ctx* ctx, unsigned long long* skip, unsigned long long* rows) {
    SurrState* s = ctx->surr;
    *skip = *rows = 0;
    if (!s || !s->ready) return RS_OK;
    std::vector<unsigned long long> slots(kCountBytes / sizeof(unsigned long long));
    HIPCHK(ctx, hipMemcpyAsync(slots.data(), s->count, kCountBytes, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    unsigned long long c[4];
    sum_counts(slots.data(), c);
    *skip = c[2];
    *rows = c[3];
    return RS_OK;
}

int surr_update_policy(rs_ctx* ctx) {
    SurrState* s = ctx->surr;
    if (!s || !s->ready) return RS_OK;
    std::vector<unsigned long long> slots(kCountBytes / sizeof(unsigned long long));
    HIPCHK(ctx, hipMemcpyAsync(slots.data(), s->count, kCountBytes, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    unsigned long long c[4];
    sum_counts(slots.data(), c);
    surr_policy_apply(s, c[2], c[3]);
    return RS_OK;
}

}  // namespace rs

namespace {

uint64_t fnv1a(uint64_t h, const void* data, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(data);
    for (size_t q = 0; q < n; ++q) {
        h ^= b[q];
        h *= 0x100000001B3ULL;
    }
    return h;
}

uint64_t network_identity(const SurrState* s, const float* P) {
    uint64_t h = fnv1a(0xCBF29CE484222325ULL, P, sizeof(float) * kNParam);
    h = fnv1a(h, &s->norm, sizeof s->norm);
    const double w[3] = {s->T_lo, s->T_hi, s->tol};
    h = fnv1a(h, w, sizeof w);
    return fnv1a(h, &s->fmt, sizeof s->fmt);
}

}  // namespace

extern "C" {

int rs_surrogate_defaults(rs_surrogate_opts* o) {
    if (!o) return RS_ERR_ARG;
    memset(o, 0, sizeof *o);
    o->n_uniform = 1 << 20;
    o->n_near = 1 << 18;
    o->seed = 20240607ULL;
    o->epochs = 40;
    o->batch = 4096;
    o->T_lo = 0.0;
    o->T_hi = 0.0;
    o->lr = 3e-3;
    o->lr_final = 3e-5;
    o->residual_tol = 1e-4;
    o->n_curve = 256;
    o->act_format = RS_ACT_FP16_PACKED;
    return RS_OK;
}

int rs_surrogate_train(rs_ctx* ctx, const rs_surrogate_opts* o, double* report, int32_t nrep) {
    if (!ctx || !o) return RS_ERR_ARG;
    const rs_params& p = ctx->p;
    const int64_t n = o->n_uniform + o->n_near;
    if (o->n_uniform < 0 || o->n_near < 0 || n < 64 || n % 64 != 0 || n > (1LL << 26))
        return fail(ctx, RS_ERR_ARG, "sample count must be a positive multiple of 64 (<= 2^26)");
    if (o->batch < 64 || o->batch % 64 != 0 || o->batch > n)
        return fail(ctx, RS_ERR_ARG, "batch must be a multiple of 64 and at most the sample count");
    if (o->epochs < 0 || o->n_curve < 2 || !(o->residual_tol > 0.0) || !(o->lr > 0.0))
        return fail(ctx, RS_ERR_ARG, "bad surrogate options");
    if (o->act_format < RS_ACT_BF16 || o->act_format > RS_ACT_FP16_PACKED)
        return fail(ctx, RS_ERR_ARG, "act_format must be one of RS_ACT_*");
    int rc = ensure_state(ctx);
    if (rc) return rc;
    SurrState* s = ctx->surr;
    s->ready = false;

    // training window: the run's temperature range widened by 20 K below
    // and 25 K above (latent heat raises the coupled field by ~10 K over
    // 10^4 steps of the 4096^2 preset), across the 1941 K breakpoint of the
    // G_Ti descriptions (fun.cu:1628-1646) where the range spans it: the
    // pieces meet there within 0.1 J/mol (liquid; its heat capacity steps
    // from 40.7 to 46.5 J/mol/K), far below the residual bound. Measured at
    // 4096^2 over steps 2000-6000: the spanning window [1907.6, 1969.0] K
    // trusts 99.9% of its bins, 0.19% of the cells fall back, tip and front
    // velocity within 5.5e-4 and 2.3e-4 of the FP64 path; a window stopped
    // at the breakpoint sent the 6% of cells above it to the exact routine
    // for tip and front errors of 2.5e-4 and 1.7e-5; two networks, one per
    // side, matched the spanning one's accuracy but cost the step kernel
    // 15-25% (a second evaluation for the rows across the switch isotherm,
    // which a cell's choice must not depend on the wavefront to keep the
    // face fluxes antisymmetric).
    double lo = o->T_lo, hi = o->T_hi;
    if (!(lo < hi)) {
        double rlo, rhi;
        run_T_range(ctx, &rlo, &rhi);
        lo = rlo - 20.0;
        hi = rhi + 25.0;
    }
    if (!(lo > 300.0 && hi < 2740.0 && hi - lo > 1.0))
        return fail(ctx, RS_ERR_DOMAIN, "surrogate temperature window outside the SGTE ranges");
    s->T_lo = lo;
    s->T_hi = hi;
    s->tol = o->residual_tol;
    s->fmt = o->act_format;
    s->norm.t_mid = (float)(0.5 * (lo + hi));
    s->norm.t_half = (float)(0.5 * (hi - lo));

    // liquidus / solidus partner compositions on the window (common
    // tangent); where no tangent exists the nearest solved point is used,
    // and the run's nominal pair if none is
    std::vector<double> cl((size_t)o->n_curve), cs((size_t)o->n_curve);
    std::vector<char> okc((size_t)o->n_curve, 0);
    double gl = p.c_liq0, gs = p.c_sol0;
    for (int q = 0; q < o->n_curve; ++q) {
        const double T = lo + (hi - lo) * q / (o->n_curve - 1);
        double a = gl, b = gs;
        if (common_tangent(p, T, &a, &b) == 0 && a > 0.0 && a < 1.0 && b > 0.0 && b < 1.0 &&
            fabs(a - b) > 1e-6) {
            cl[q] = a;
            cs[q] = b;
            okc[q] = 1;
            gl = a;
            gs = b;
        }
    }
    int any = -1;
    for (int q = 0; q < o->n_curve; ++q)
        if (okc[q]) { any = q; break; }
    for (int q = 0; q < o->n_curve; ++q) {
        if (okc[q]) continue;
        if (any < 0) {
            cl[q] = p.c_liq0;
            cs[q] = p.c_sol0;
            continue;
        }
        int best = any;
        for (int r = 0; r < o->n_curve; ++r)
            if (okc[r] && abs(r - q) < abs(best - q)) best = r;
        cl[q] = cl[best];
        cs[q] = cs[best];
    }

    Scratch sc;
    float *dX, *dYn, *dO, *partial, *lossp;
    double *dY, *dTc, *dcl, *dcs;
    HIPCHK(ctx, sc.get(&dX, 2 * n));
    HIPCHK(ctx, sc.get(&dYn, 4 * n));
    HIPCHK(ctx, sc.get(&dO, 4 * n));
    HIPCHK(ctx, sc.get(&dY, 4 * n));
    HIPCHK(ctx, sc.get(&dTc, 2 * n));
    HIPCHK(ctx, sc.get(&dcl, o->n_curve));
    HIPCHK(ctx, sc.get(&dcs, o->n_curve));
    HIPCHK(ctx, sc.get(&partial, (size_t)(o->batch / 64) * kNQ));
    HIPCHK(ctx, sc.get(&lossp, o->batch / 64));
    HIPCHK(ctx, hipMemcpyAsync(dcl, cl.data(), sizeof(double) * o->n_curve, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(dcs, cs.data(), sizeof(double) * o->n_curve, hipMemcpyHostToDevice, ctx->stream));

    SampleArgs sa;
    sa.X = dX;
    sa.Y = dY;
    sa.Tc = dTc;
    sa.n_uni = o->n_uniform;
    sa.n_near = o->n_near;
    sa.seed = o->seed;
    sa.T_lo = lo;
    sa.T_hi = hi;
    sa.t_mid = s->norm.t_mid;
    sa.t_half = s->norm.t_half;
    sa.cl = dcl;
    sa.cs = dcs;
    sa.n_curve = o->n_curve;
    sa.L0_bcc = p.L0_bcc;
    sa.L0_liq = p.L0_liq;
    sa.L1_bcc = p.L1_bcc;
    sa.L1_liq = p.L1_liq;
    sa.RR = p.RR;
    HIPCHK(ctx, launch_mlp_samples(sa, ctx->stream));

    // output normalisation from the sample set
    {
        std::vector<double> Y((size_t)(4 * n));
        HIPCHK(ctx, hipMemcpyAsync(Y.data(), dY, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        for (int k = 0; k < 4; ++k) {
            double m1 = 0.0, m2 = 0.0;
            for (int64_t q = 0; q < n; ++q) m1 += Y[4 * q + k];
            m1 /= (double)n;
            for (int64_t q = 0; q < n; ++q) {
                const double e = Y[4 * q + k] - m1;
                m2 += e * e;
            }
            const double sd = sqrt(m2 / (double)n);
            if (!std::isfinite(m1) || !std::isfinite(sd))
                return fail(ctx, RS_ERR_DOMAIN, "non-finite CALPHAD targets in the training window");
            s->norm.y_mean[k] = (float)m1;
            s->norm.y_std[k] = (float)(sd > 1e-6 ? sd : 1.0);
        }
    }
    HIPCHK(ctx, launch_mlp_normalise(dY, dYn, n, s->norm, ctx->stream));

    // initial weights: uniform Glorot from the seed, zero biases
    {
        std::vector<float> P0((size_t)kNParam, 0.0f);
        uint64_t st = o->seed ^ 0x5DEECE66DULL;
        auto fill = [&](int off, int fan_in, int fan_out, int count) {
            const double r = sqrt(6.0 / (fan_in + fan_out));
            for (int q = 0; q < count; ++q) P0[off + q] = (float)((2.0 * unif(&st) - 1.0) * r);
        };
        fill(kOffW1, kIn, kH, kH * kIn);
        fill(kOffW2, kH, kH, kH * kH);
        fill(kOffW3, kH, kOut, kOut * kH);
        HIPCHK(ctx, hipMemcpyAsync(s->P, P0.data(), sizeof(float) * kNParam, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    HIPCHK(ctx, hipMemsetAsync(s->m, 0, sizeof(float) * kNParam, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(s->v, 0, sizeof(float) * kNParam, ctx->stream));
    HIPCHK(ctx, launch_mlp_pack(s->P, s->F, ctx->stream));
    double mse0 = 0.0;
    rc = mse_of(ctx, dX, dYn, n, dO, &mse0);
    if (rc) return rc;

    hipEvent_t e0, e1;
    HIPCHK(ctx, hipEventCreate(&e0));
    HIPCHK(ctx, hipEventCreate(&e1));
    HIPCHK(ctx, hipEventRecord(e0, ctx->stream));
    const int64_t spe = n / o->batch;
    const int64_t total = spe * o->epochs;
    int64_t t = 0;
    uint64_t st = o->seed * 0x2545F4914F6CDD1DULL + 1;
    for (int ep = 0; ep < o->epochs; ++ep) {
        // a fresh bijective visiting order per epoch: r -> (r mult + off) mod n
        int64_t mult;
        do {
            mult = (int64_t)(splitmix(&st) % (uint64_t)n) | 1;
        } while (std::gcd(mult, n) != 1);
        const int64_t off = (int64_t)(splitmix(&st) % (uint64_t)n);
        for (int64_t b = 0; b < spe; ++b) {
            ++t;
            const double fr = (double)(t - 1) / (double)std::max<int64_t>(total - 1, 1);
            const double lr = o->lr_final + 0.5 * (o->lr - o->lr_final) * (1.0 + cos(M_PI * fr));
            hipError_t e = launch_mlp_train_step(dX, dYn, n, b * o->batch, o->batch, mult, off, s->F,
                                                 partial, lossp, s->P, s->m, s->v, (float)lr, 0.9f,
                                                 0.999f, 1e-8f, (int)t, s->fmt != kActBf16, ctx->stream);
            if (e != hipSuccess) {
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                return fail(ctx, RS_ERR_HIP, "surrogate train step", e);
            }
        }
    }
    HIPCHK(ctx, hipEventRecord(e1, ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);

    double mse1 = 0.0;
    rc = mse_of(ctx, dX, dYn, n, dO, &mse1);
    if (rc) return rc;
    double err4[4], rmax, frac;
    rc = build_trust(ctx, err4, &rmax, &frac);
    if (rc) return rc;
    {
        std::vector<float> Ph(kNParam);
        HIPCHK(ctx, hipMemcpyAsync(Ph.data(), s->P, sizeof(float) * kNParam, hipMemcpyDeviceToHost,
                                   ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        s->ident = network_identity(s, Ph.data());
    }
    s->ready = true;
    const double rep[RS_SURR_NREPORT] = {mse1,    frac, err4[0],      err4[1], err4[2], err4[3],
                                         rmax,    ms * 1e-3, (double)t, lo,     hi,      mse0};
    if (report)
        for (int q = 0; q < nrep && q < RS_SURR_NREPORT; ++q) report[q] = rep[q];
    return RS_OK;
}

int rs_surrogate_get(rs_ctx* ctx, float* params, int32_t np, double* meta, int32_t nm) {
    if (!ctx) return RS_ERR_ARG;
    if (!surr_ready(ctx)) return fail(ctx, RS_ERR_STATE, "no surrogate weights");
    SurrState* s = ctx->surr;
    if (params) {
        if (np != kNParam) return fail(ctx, RS_ERR_ARG, "parameter count");
        HIPCHK(ctx, hipMemcpyAsync(params, s->P, sizeof(float) * kNParam, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (meta) {
        const double v[RS_SURR_NMETA] = {s->norm.t_mid,    s->norm.t_half,   s->norm.y_mean[0],
                                         s->norm.y_mean[1], s->norm.y_mean[2], s->norm.y_mean[3],
                                         s->norm.y_std[0],  s->norm.y_std[1],  s->norm.y_std[2],
                                         s->norm.y_std[3],  s->T_lo,          s->T_hi,
                                         s->tol,            (double)s->fmt};
        for (int q = 0; q < nm && q < RS_SURR_NMETA; ++q) meta[q] = v[q];
    }
    return RS_OK;
}

int rs_surrogate_set(rs_ctx* ctx, const float* params, int32_t np, const double* meta, int32_t nm) {
    if (!ctx || !params || !meta) return RS_ERR_ARG;
    if (np != kNParam || nm < RS_SURR_NMETA) return fail(ctx, RS_ERR_ARG, "parameter or meta count");
    if (!(meta[1] > 0.0) || !(meta[12] > 0.0)) return fail(ctx, RS_ERR_ARG, "bad surrogate meta");
    if (!(meta[13] == RS_ACT_BF16 || meta[13] == RS_ACT_BF16_SPLIT || meta[13] == RS_ACT_FP16 ||
          meta[13] == RS_ACT_FP16_PACKED))
        return fail(ctx, RS_ERR_ARG, "bad surrogate meta: act_format");
    for (int k = 0; k < 4; ++k)
        if (!(meta[6 + k] > 0.0)) return fail(ctx, RS_ERR_ARG, "bad surrogate meta");
    int rc = ensure_state(ctx);
    if (rc) return rc;
    SurrState* s = ctx->surr;
    s->ready = false;
    s->norm.t_mid = (float)meta[0];
    s->norm.t_half = (float)meta[1];
    for (int k = 0; k < 4; ++k) {
        s->norm.y_mean[k] = (float)meta[2 + k];
        s->norm.y_std[k] = (float)meta[6 + k];
    }
    s->T_lo = meta[10];
    s->T_hi = meta[11];
    s->tol = meta[12];
    s->fmt = (int)meta[13];
    HIPCHK(ctx, hipMemcpyAsync(s->P, params, sizeof(float) * kNParam, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(s->m, 0, sizeof(float) * kNParam, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(s->v, 0, sizeof(float) * kNParam, ctx->stream));
    HIPCHK(ctx, launch_mlp_pack(s->P, s->F, ctx->stream));
    double err4[4], rmax, frac;
    rc = build_trust(ctx, err4, &rmax, &frac);
    if (rc) return rc;
    s->ident = network_identity(s, params);
    s->ready = true;
    return RS_OK;
}

int rs_surrogate_grad(rs_ctx* ctx, const float* X, const float* Yn, int32_t n, float* grad,
                      double* loss) {
    if (!ctx || !X || !Yn || !grad) return RS_ERR_ARG;
    if (n < 64 || n % 64 != 0 || n > 65536) return fail(ctx, RS_ERR_ARG, "n must be a multiple of 64, <= 65536");
    if (!surr_ready(ctx)) return fail(ctx, RS_ERR_STATE, "no surrogate weights");
    SurrState* s = ctx->surr;
    Scratch sc;
    float *dX, *dYn, *partial, *lossp, *dg;
    HIPCHK(ctx, sc.get(&dX, 2 * (size_t)n));
    HIPCHK(ctx, sc.get(&dYn, 4 * (size_t)n));
    HIPCHK(ctx, sc.get(&partial, (size_t)(n / 64) * kNQ));
    HIPCHK(ctx, sc.get(&lossp, n / 64));
    HIPCHK(ctx, sc.get(&dg, kNParam));
    HIPCHK(ctx, hipMemcpyAsync(dX, X, sizeof(float) * 2 * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(dYn, Yn, sizeof(float) * 4 * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, launch_mlp_grad(dX, dYn, n, s->F, partial, lossp, dg, s->fmt != kActBf16, ctx->stream));
    std::vector<float> lp((size_t)(n / 64));
    HIPCHK(ctx, hipMemcpyAsync(grad, dg, sizeof(float) * kNParam, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(lp.data(), lossp, sizeof(float) * (n / 64), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (loss) {
        double acc = 0.0;
        for (float v : lp) acc += v;
        *loss = acc / (4.0 * n);
    }
    return RS_OK;
}

int rs_surrogate_eval(rs_ctx* ctx, const double* Tc, int64_t n, double* out) {
    if (!ctx || !Tc || !out || n < 1 || n > (1LL << 26)) return RS_ERR_ARG;
    if (!surr_ready(ctx)) return fail(ctx, RS_ERR_STATE, "no surrogate weights");
    SurrState* s = ctx->surr;
    Scratch sc;
    double* dTc;
    float *dX, *dO;
    HIPCHK(ctx, sc.get(&dTc, 2 * (size_t)n));
    HIPCHK(ctx, sc.get(&dX, 2 * (size_t)n));
    HIPCHK(ctx, sc.get(&dO, 4 * (size_t)n));
    HIPCHK(ctx, hipMemcpyAsync(dTc, Tc, sizeof(double) * 2 * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, launch_mlp_inputs(dTc, n, s->norm.t_mid, s->norm.t_half, dX, ctx->stream));
    HIPCHK(ctx, launch_mlp_eval(dX, n, s->F, dO, s->fmt, ctx->stream));
    std::vector<float> O((size_t)(4 * n));
    std::vector<float> Xh((size_t)(2 * n));
    HIPCHK(ctx, hipMemcpyAsync(O.data(), dO, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(Xh.data(), dX, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const double RR = ctx->p.RR;
    for (int64_t q = 0; q < n; ++q) {
        double y[4];
        for (int k = 0; k < 4; ++k) y[k] = (double)O[4 * q + k] * s->norm.y_std[k] + s->norm.y_mean[k];
        const double T = Tc[2 * q], c = Tc[2 * q + 1], om = 1.0 - c;
        double id = 0.0, lr = 0.0;
        if (c > 0.0 && om > 0.0) {
            id = RR * T * (c * log(c) + om * log(om));
            lr = RR * T * log(c / om);
        }
        double* r = out + 5 * q;
        r[1] = y[1] + id;
        r[0] = y[0] + r[1];
        r[2] = y[2] + lr;
        r[3] = r[2] + y[3];
        const float a = Xh[2 * q], b = Xh[2 * q + 1];
        bool ok = fabsf(a) <= 1.0f && fabsf(b) <= 1.0f;
        if (ok) {
            const int ia = std::min((int)((a + 1.0f) * 32.0f), kTrustN - 1);
            const int ib = std::min((int)((b + 1.0f) * 32.0f), kTrustN - 1);
            ok = (s->h_trust[ib] >> ia) & 1ULL;
        }
        r[4] = ok ? 1.0 : 0.0;
    }
    return RS_OK;
}

int rs_surrogate_time(rs_ctx* ctx, int64_t n, int32_t reps, double* avg_ms) {
    if (!ctx || !avg_ms || n < 64 || n > (1LL << 26) || reps < 1) return RS_ERR_ARG;
    if (!surr_ready(ctx)) return fail(ctx, RS_ERR_STATE, "no surrogate weights");
    SurrState* s = ctx->surr;
    Scratch sc;
    float *dX, *dO;
    HIPCHK(ctx, sc.get(&dX, 2 * (size_t)n));
    HIPCHK(ctx, sc.get(&dO, 4 * (size_t)n));
    HIPCHK(ctx, launch_mlp_spread_inputs(dX, n, ctx->stream));
    HIPCHK(ctx, launch_mlp_eval(dX, n, s->F, dO, s->fmt, ctx->stream));  // warm-up
    hipEvent_t e0, e1;
    HIPCHK(ctx, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(ctx, RS_ERR_HIP, "hipEventCreate");
    }
    hipError_t e = hipEventRecord(e0, ctx->stream);
    for (int q = 0; q < reps && e == hipSuccess; ++q)
        e = launch_mlp_eval(dX, n, s->F, dO, s->fmt, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(ctx, RS_ERR_HIP, hipGetErrorString(e));
    *avg_ms = (double)ms / reps;
    return RS_OK;
}

int rs_surrogate_stats(rs_ctx* ctx, double* out, int32_t reset) {
    if (!ctx || !out) return RS_ERR_ARG;
    if (!ctx->surr) return fail(ctx, RS_ERR_STATE, "no surrogate");
    std::vector<unsigned long long> slots(kCountBytes / sizeof(unsigned long long));
    HIPCHK(ctx, hipMemcpyAsync(slots.data(), ctx->surr->count, kCountBytes, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    unsigned long long c[4];
    sum_counts(slots.data(), c);
    out[0] = (double)c[0];
    out[1] = (double)c[1];
    out[2] = c[1] ? (double)c[0] / (double)c[1] : 0.0;
    out[3] = (double)c[2];
    out[4] = (double)c[3];
    if (reset) {
        HIPCHK(ctx, hipMemsetAsync(ctx->surr->count, 0, kCountBytes, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        ctx->surr->seen_rows = ctx->surr->seen_skip = 0;
    }
    return RS_OK;
}

}  // extern "C"
