// rs_physics.hpp -- pointwise physics shared by the HIP kernels and the host
// library: SGTE/Redlich-Kister CALPHAD free energies and chemical
// potentials, interpolation polynomials q/g/g', the anisotropic phase-field
// operator and the derived model constants.
//
// FP64 functions keep the operand order of the reference expressions they
// restate (fun.cu cited per function) so that, with floating-point
// contraction disabled, they round exactly like the reference wherever the
// math library agrees. The exceptions are marked: the FAST form of the
// CALPHAD evaluation and the anisotropic operator's reciprocals, which the
// FP64 step kernel uses to save IEEE divisions and a log per cell (each
// result within a few ulp of the reference's form).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/rsolid.h"

namespace rs {

#define RS_HD __host__ __device__ __forceinline__

// ---------------------------------------------------------------------------
// derived constants (fun.cu:58-105, 189-195)
// ---------------------------------------------------------------------------
struct Derived {
    double tau0;     // (S W0)^2 / (Gamma muk0)              fun.cu:69
    double hh;       // sigma / (a1 S W0 1e-9)                fun.cu:104
    double h0;       // RR Tmelt / v0                         fun.cu:105
    double Lcp;      // latentHeat / cp                       fun.cu:95
    double dG_den;   // 2 hh v0  (phi driving-force divisor)  fun.cu:1197
    double mu_den;   // 1 / (h0 v0)                           fun.cu:1454
    double kdc;      // dt Dl / (Gamma muk0)                  fun.cu:1462
};

inline Derived derive_constants(const rs_params& p, double dt) {
    Derived d;
    double sw = p.S * p.W0;
    d.tau0 = (sw * sw) / (p.Gamma * p.muk0);
    double a1 = 2.0 * sqrt(2.0) / 3.0;
    double lh2 = p.latentHeat * p.num_moles;
    double st = p.Gamma * 1e-9 * lh2 / p.Tmelt;
    d.hh = st / (a1 * p.S * p.W0 * 1.0e-9);
    d.h0 = p.RR * p.Tmelt / p.v0;
    d.Lcp = p.latentHeat / p.cp;
    d.dG_den = 2.0 * d.hh * p.v0;
    d.mu_den = 1.0 / (d.h0 * p.v0);
    d.kdc = dt * p.Dl / (p.Gamma * p.muk0);
    return d;
}

// Vp_ramp (fun.cu:1974-1981); t in ns
RS_HD double vp_ramp(const rs_params& p, double t) {
    if (p.if_Vp_ramp) {
        t = t * 0.001;
        return p.Vp + p.Vp_ramp_rate * t;
    }
    return p.Vp;
}

// ---------------------------------------------------------------------------
// CALPHAD, FP64.  All four SGTE unaries share one log(T) (they evaluate
// T*log(T) with the same log), so the per-cell evaluation takes log(T) as an
// argument.  NaN outside the SGTE ranges (fun.cu:1635-1665).
// ---------------------------------------------------------------------------
struct Unaries {
    double ti_bcc, ti_liq, nb_bcc, nb_liq;
};

// 1/x and 1/sqrt(x) for a normal x > 0 to about one ulp: the hardware
// estimate (the division on the host) refined by two Newton steps, ~5
// instructions against ~11 for an IEEE FP64 division and ~15 for a square
// root and a division
RS_HD double rcp_nr(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(x);
#else
    double y = 1.0 / x;
#endif
    y = fma(y, fma(-x, y, 1.0), y);
    return fma(y, fma(-x, y, 1.0), y);
}
RS_HD double rsqrt_nr(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rsq(x);
#else
    double y = 1.0 / sqrt(x);
#endif
    y = y * (1.5 - 0.5 * x * (y * y));
    return y * (1.5 - 0.5 * x * (y * y));
}

// ln r for a normal r > 0 from a 128-entry table tab[2k] = 1/F_k,
// tab[2k+1] = ln F_k, F_k = 1 + (k + 1/2)/128 (Tang's table method): with
// r = 2^(e-1) m2, m2 in [1, 2), and F the table point of m2's top 7 mantissa
// bits, ln r = (e-1) ln 2 + ln F + log1p(u), u = (m2 - F)/F, |u| < 2^-8, and
// log1p by its degree-6 Taylor polynomial (truncation below 3e-18). The sum
// (e-1) ln2_hi + ln F is carried with its rounding error (Fast2Sum: the first
// term is exact and, when non-zero, at least ln 2 > |ln F|), so the result is
// rounded once: within about half an ulp plus the table's rounding, like the
// math library's log, in ~25 instructions against its ~90.
RS_HD double ln_tab(double r, const double* tab) {
    int e;
    const double m = frexp(r, &e);  // m in [0.5, 1)
    const double m2 = m + m;        // [1, 2), exact
    const uint64_t bits = __builtin_bit_cast(uint64_t, m2);
    const uint32_t hi = (uint32_t)(bits >> 32);
    const int k = (int)((hi >> 13) & 127u);  // top 7 mantissa bits
    const double F = __builtin_bit_cast(double, (uint64_t)((hi & 0xFFFFE000u) | 0x1000u) << 32);
    const double d = m2 - F;  // exact, |d| <= 2^-8
    const double u = d * tab[2 * k];
    double q = fma(u, -1.0 / 6.0, 0.2);
    q = fma(u, q, -0.25);
    q = fma(u, q, 1.0 / 3.0);
    q = fma(u, q, -0.5);
    const double l1p = fma(u * u, q, u);  // log1p(u)
    const double ef = (double)(e - 1);
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double lnF = tab[2 * k + 1];
    const double a = ef * ln2_hi;  // exact: ln2_hi has 32 significant bits
    const double s = a + lnF;
    const double t = (a - s) + lnF;  // the rounding error of s, exactly
    return s + (t + fma(ef, ln2_lo, l1p));
}

// the natural log of the CALPHAD evaluation: the math library's, or the
// table log above (the FP64 step kernel, with the table in LDS)
struct LibLog {
    RS_HD double operator()(double x) const { return log(x); }
};
struct TabLog {
    const double* tab;
    RS_HD double operator()(double x) const { return ln_tab(x, tab); }
};

// C / T as written (FAST = false), or as C times one reciprocal of T shared
// by the four unaries (FAST = true: the FP64 step kernel, where each IEEE
// FP64 division is a sequence of about eleven instructions)
template <bool FAST>
RS_HD double over_T(double C, double T, double iT) {
    return FAST ? C * iT : C / T;
}

template <bool FAST = false>
RS_HD Unaries sgte_unaries(double T, double lT) {
    Unaries u;
    const double t2 = T * T;
    const double t3 = t2 * T;
    const double iT = FAST ? rcp_nr(T) : 0.0;
    // G_Ti_BCC fun.cu:1628-1636
    if (T >= 1155.0 && T < 1941.0)
        u.ti_bcc = 6667.385 + 105.366379 * T - 22.3771 * T * lT + 1.21707e-3 * t2 -
                   0.84534e-6 * t3 - over_T<FAST>(2002750.0, T, iT);
    else if (T >= 1941.0)
        u.ti_bcc = 26483.26 - 182.426471 * T + 19.0900905 * T * lT - 22.00832e-3 * t2 +
                   1.228863e-6 * t3 + over_T<FAST>(1400501.0, T, iT);
    else
        u.ti_bcc = NAN;
    // G_Ti_Liq fun.cu:1638-1646
    if (T >= 1300.0 && T < 1941.0)
        u.ti_liq = 369519.198 - 2554.0225 * T + 342.059267 * T * lT - 163.409355e-3 * t2 +
                   12.457117e-6 * t3 - over_T<FAST>(67034516.0, T, iT);
    else if (T >= 1941.0 && T < 4000.0)
        u.ti_liq = -19887.066 + 298.7367 * T - 46.29 * T * lT;
    else
        u.ti_liq = NAN;
    // G_Nb_BCC fun.cu:1648-1656
    if (T >= 298.15 && T < 2750.0)
        u.nb_bcc = -8519.353 + 142.045475 * T - 26.4711 * T * lT + 0.203475e-3 * t2 -
                   0.35012e-6 * t3 + over_T<FAST>(93399.0, T, iT);
    else if (T >= 2750.0)
        u.nb_bcc = -37669.3 + 271.720843 * T - 41.77 * T * lT + 1528.238e29 / ((t3 * t3) * t3);
    else
        u.nb_bcc = NAN;
    // G_Nb_Liq fun.cu:1658-1666
    if (T >= 298.15 && T < 2750.0)
        u.nb_liq = 21262.202 + 131.229057 * T - 26.4711 * T * lT + 0.203475e-3 * t2 -
                   0.35012e-6 * t3 + over_T<FAST>(93399.0, T, iT) - 306.098e-25 * (((t2 * t2) * t2) * T);
    else if (T >= 2750.0)
        u.nb_liq = -7499.398 + 260.756148 * T - 41.77 * T * lT;
    else
        u.nb_liq = NAN;
    return u;
}

struct RK {  // Redlich-Kister coefficients, J/mol
    double L0_bcc, L0_liq, L1_bcc, L1_liq, RR;
};

RS_HD RK rk_of(const rs_params& p) { return RK{p.L0_bcc, p.L0_liq, p.L1_bcc, p.L1_liq, p.RR}; }

// Per-cell thermodynamics: f_L - f_S (phi driving force numerator,
// Gl - Gs of fun.cu:1197) and the chemical potentials mu_L = Glc,
// mu_S = Gsc (fun.cu:1701-1720).
struct Thermo {
    double dG;   // Gl(T,c) - Gs(T,c)
    double muL;  // Glc(T,c)
    double muS;  // Gsc(T,c)
};

RS_HD double ideal_mix(double RR, double T, double c, double onemc, double lc, double l1c) {
    // fun.cu:1673-1680 branches at c < 1e-8 and 1-c < 1e-8
    if (c < 1.e-8) return RR * T * (onemc * l1c);
    if (onemc < 1.e-8) return RR * T * (c * lc);
    return RR * T * (onemc * l1c + c * lc);
}

// FAST (the FP64 step kernel): the SGTE terms in 1/T share one reciprocal,
// and log(c/(1-c)) is log c - log(1-c) from the two logs already taken
// (one log and one division fewer per cell; each result within a few ulp of
// the reference's form). Ln: the log taken (LibLog, or TabLog in the step
// kernel).
template <bool FAST = false, typename Ln = LibLog>
RS_HD Thermo thermo_fp64(const RK& rk, double T, double c, Ln ln = Ln{}) {
    const double lT = ln(T);
    const Unaries u = sgte_unaries<FAST>(T, lT);
    const double onemc = 1.0 - c;
    // the logs are evaluated only where the reference evaluates them
    const double lc = (onemc < 1.e-8 || c >= 1.e-8) ? ln(c) : 0.0;
    const double l1c = (c < 1.e-8 || onemc >= 1.e-8) ? ln(onemc) : 0.0;
    const double id = ideal_mix(rk.RR, T, c, onemc, lc, l1c);
    const double exl = c * onemc * (rk.L0_liq + rk.L1_liq * (1.0 - 2.0 * c));
    const double exs = c * onemc * (rk.L0_bcc + rk.L1_bcc * (1.0 - 2.0 * c));
    const double Gl = id + onemc * u.nb_liq + c * u.ti_liq + exl;
    const double Gs = id + onemc * u.nb_bcc + c * u.ti_bcc + exs;
    double lr = 0.;
    if (c > 1.e-8 && onemc > 1.e-8) lr = rk.RR * T * (FAST ? lc - l1c : ln(c / onemc));
    const double tl = rk.L0_liq * (1.0 - 2.0 * c) + rk.L1_liq * (1.0 - 6.0 * c + 6.0 * (c * c));
    const double ts = rk.L0_bcc * (1.0 - 2.0 * c) + rk.L1_bcc * (1.0 - 6.0 * c + 6.0 * (c * c));
    Thermo th;
    th.dG = Gl - Gs;
    th.muL = u.ti_liq - u.nb_liq + tl + lr;
    th.muS = u.ti_bcc - u.nb_bcc + ts + lr;
    return th;
}

// Scalar forms used by host code (common tangent, tests of the library).
RS_HD double Gl_fp64(const RK& rk, double T, double c) {
    const double lT = log(T);
    const Unaries u = sgte_unaries(T, lT);
    const double onemc = 1.0 - c;
    const double lc = (onemc < 1.e-8 || c >= 1.e-8) ? log(c) : 0.0;
    const double l1c = (c < 1.e-8 || onemc >= 1.e-8) ? log(onemc) : 0.0;
    return ideal_mix(rk.RR, T, c, onemc, lc, l1c) + onemc * u.nb_liq + c * u.ti_liq +
           c * onemc * (rk.L0_liq + rk.L1_liq * (1.0 - 2.0 * c));
}
RS_HD double Gs_fp64(const RK& rk, double T, double c) {
    const double lT = log(T);
    const Unaries u = sgte_unaries(T, lT);
    const double onemc = 1.0 - c;
    const double lc = (onemc < 1.e-8 || c >= 1.e-8) ? log(c) : 0.0;
    const double l1c = (c < 1.e-8 || onemc >= 1.e-8) ? log(onemc) : 0.0;
    return ideal_mix(rk.RR, T, c, onemc, lc, l1c) + onemc * u.nb_bcc + c * u.ti_bcc +
           c * onemc * (rk.L0_bcc + rk.L1_bcc * (1.0 - 2.0 * c));
}

// ---------------------------------------------------------------------------
// interpolation polynomials (fun.cu:189-195) with the solid-diffusivity
// floor q_s (q_s = 0 gives the reference's q exactly)
// ---------------------------------------------------------------------------
template <typename R>
RS_HD R qq(R AA, R qs, R ps) {
    const R om = R(1) - ps;
    const R q = R(0.5) * AA * om - R(0.25) * (AA - R(1)) * (om * om);
    return q * (R(1) - qs) + qs;
}
template <typename R>
RS_HD R gg(R ps) {
    const R p2 = ps * ps;
    return R(1.875) * ps - R(1.25) * (p2 * ps) + R(0.375) * ((p2 * p2) * ps);
}
template <typename R>
RS_HD R dgg(R ps) {
    const R a = R(1) - ps * ps;
    return R(1.875) * (a * a);
}

// ---------------------------------------------------------------------------
// anisotropic operator (fun.cu:1030-1139, if_1_eps). ph = 3x3 neighbourhood,
// ph[3*b + a] with a = x offset + 1, b = y offset + 1.
// ---------------------------------------------------------------------------
template <typename R>
struct Aniso {
    R ops[8];  // orientation entries x1 x2 y1 y2 z1 z2 (6 used) + pad
    R eps1, epk1, lap_w_axis, lap_w_diag;
    R inv2dx;  // unused by the FP64 path (kept for the mixed kernels)
    R dx;
};

// The divisions of fun.cu:1040-1075 by the grid constants 2 dx, dx^2,
// 4 dx^2 and 2 dx^2 are products with their reciprocals (loop invariants,
// computed once per wavefront), and the four by |grad phi| and its square
// share one reciprocal square root of |grad phi|^2 (the test |grad phi| >
// 1e-15 as |grad phi|^2 > 1e-30).
template <typename R>
RS_HD R anisotropy(const Aniso<R>& A, R pmm, R p0m, R p1m, R pm0, R p00, R p10, R pm1, R p01,
                   R p11, R& as, R& ak) {
    const R dx = A.dx;
    const R i2dx = R(1.0) / (R(2.0) * dx), idx2 = R(1.0) / (dx * dx);
    const R i4dx2 = R(1.0) / (R(4.) * dx * dx), i2dx2 = R(1.0) / (R(2.0) * dx * dx);
    const R gx = (p10 - pm0) * i2dx;
    const R gy = (p01 - p0m) * i2dx;
    const R ag2 = gx * gx + gy * gy;
    const R gxx = (p10 - R(2.0) * p00 + pm0) * idx2;
    const R gyy = (p01 - R(2.0) * p00 + p0m) * idx2;
    const R gxy = (p11 - pm1 - p1m + pmm) * i4dx2;
    R lap = A.lap_w_axis * (gxx + gyy);
    lap += A.lap_w_diag * (p11 + pm1 + p1m + pmm - R(4.0) * p00) * i2dx2;
    R thx, thy, ct, st;
    if (ag2 > R(1.0e-30)) {
        const R iag = rsqrt_nr(ag2);
        const R iag2 = iag * iag;
        thx = (gx * gxy - gy * gxx) * iag2;
        thy = (gx * gyy - gy * gxy) * iag2;
        ct = gx * iag;
        st = gy * iag;
    } else {
        thx = thy = ct = st = R(0.0);
    }
    const R nx = A.ops[0] * ct + A.ops[1] * st;
    const R ny = A.ops[2] * ct + A.ops[3] * st;
    const R nz = A.ops[4] * ct + A.ops[5] * st;
    const R ex = -A.ops[0] * st + A.ops[1] * ct;
    const R ey = -A.ops[2] * st + A.ops[3] * ct;
    const R ez = -A.ops[4] * st + A.ops[5] * ct;
    const R nx2 = nx * nx, ny2 = ny * ny, nz2 = nz * nz;
    const R nx3 = nx2 * nx, ny3 = ny2 * ny, nz3 = nz2 * nz;
    const R np4 = nx2 * nx2 + ny2 * ny2 + nz2 * nz2;
    const R d1 = R(4.0) * (nx3 * ex + ny3 * ey + nz3 * ez);
    const R xe = nx * ex, ye = ny * ey, ze = nz * ez;
    R d2 = R(3.0) * (xe * xe) + nx3 * (-nx);
    d2 += R(3.0) * (ye * ye) + ny3 * (-ny);
    d2 += R(3.0) * (ze * ze) + nz3 * (-nz);
    d2 *= R(4);
    ak = R(1.0) - R(3.0) * A.epk1 + R(4.0) * A.epk1 * np4;
    as = R(1.0) - R(3.0) * A.eps1 + R(4.0) * A.eps1 * np4;
    const R a = as;
    const R da1 = R(4.0) * A.eps1 * d1;
    const R da2 = R(4.0) * A.eps1 * d2;
    return (a * a) * lap + R(2.0) * a * da1 * (thx * gx + thy * gy) +
           (da2 * a + da1 * da1) * (thy * gx - thx * gy);
}

// set_orientation (fun.cu:750-771) with the float-precision PI of fun.cu:59
inline void orientation(double* o, double al, double be, double ga) {
    const double PI_REF = 3.1415927;
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

}  // namespace rs
