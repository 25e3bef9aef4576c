// rs_mixed.hip -- mixed-precision fused step for gfx950.
//
// Same wavefront-marching layout as the FP64 stepper (62 output columns per
// wavefront, a band of rows per wavefront, 3-row register window), with:
//   * psi (hence phi) stored and updated in FP32; anisotropy in FP32 with
//     reciprocal/rsqrt arithmetic (specialised for the identity orientation);
//   * fine T rebuilt in FP32 as T' = T - 1941 K from an FP32 copy of the
//     coarse grid, which itself is advanced in FP64;
//   * SGTE temperature dependence from a build-time polynomial table
//     (rs_sgte_poly.inc: 100 K pieces, degree 5, FP32 Horner with FP64
//     constant terms, coefficients as instruction immediates), chosen per
//     wavefront with a uniform loop over the pieces present; cells outside
//     the table use the exact FP64 routine;
//   * ln(c/(1-c)) in FP64 from a 128-entry table log (Tang's method: table
//     value ln F_k plus a degree-6 log1p of the reduced argument) and a
//     Newton-refined FP64 reciprocal, instead of an IEEE divide and libm log;
//   * conservative face fluxes: FP32 mobility weights, FP64 chemical
//     potential differences, fluxes and divergence; each face flux is
//     computed once, by one lane, and used with opposite signs by the two
//     cells it separates, so the interior scheme conserves solute to
//     rounding;
//   * i+-1 neighbours through DPP wavefront shifts (wave_shr:1 / wave_shl:1).
// The file is compiled without floating-point contraction: every FMA is
// written out, so that the prologue of a band and the loop body of the band
// above it produce bitwise-identical face fluxes.
#include <hip/hip_ext.h>
#include <math.h>

#include <algorithm>

#include "rs_mixed.hpp"
#include "rs_sgte_poly.inc"

namespace rs {

// lane l <- lane l-1 / lane l+1; the lane with no source gets 0 (it is a
// halo lane whose value is never used)
__device__ __forceinline__ int dpp_up(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ int dpp_dn(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ float shup(float v) { return __int_as_float(dpp_up(__float_as_int(v))); }
__device__ __forceinline__ float shdn(float v) { return __int_as_float(dpp_dn(__float_as_int(v))); }
__device__ __forceinline__ double shup(double v) {
    return __hiloint2double(dpp_up(__double2hiint(v)), dpp_up(__double2loint(v)));
}
__device__ __forceinline__ double shdn(double v) {
    return __hiloint2double(dpp_dn(__double2hiint(v)), dpp_dn(__double2loint(v)));
}

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// c/(1-c) with a Newton-refined reciprocal (~1 ulp)
__device__ __forceinline__ double ratio_c(double c, double om) {
    double y = __builtin_amdgcn_rcp(om);
    y = fma(y, fma(-om, y, 1.0), y);
    double r = c * y;
    r = fma(fma(-om, r, c), y, r);
    return r;
}

// coarse-T rows and the y weight for fine row j (uniform over the wavefront);
// dx*ratio is an integer number of cells, so the reference's
// int((j-1+dxT)/10) is exact integer arithmetic here. The coarse/fine ratio
// (ratio_dxT_dx, fun.cu:109) is fixed at 10 by the library.
constexpr int kRatio = 10;

// Fine T' of a lane's column (fun.cu:1580 bilinear form): T' = fmaf(A, y, B)
// with A = fmaf(a11, x, a01), B = fmaf(a10, x, f00). x and the four coarse
// values are fixed for a lane within one coarse row (10 fine rows), so A and
// B are cached and each fine row costs one FMA; the operations, hence the
// values, are those of the direct form.
// The coarse T' values a workgroup's wavefronts interpolate from, staged in
// LDS at the start of the single-run kernel (p null: read from the FP32
// coarse grid in global memory, as the batch kernel does). Reading them from
// LDS keeps the once-per-coarse-row loads off the vector-memory counter,
// which on gfx950 also counts the row's stores (surrogate kernels: 2%).
typedef const __attribute__((address_space(3))) float* LdsF;
struct TPatch {
    LdsF p;           // the patch in LDS (valid when on)
    int on;
    int iT0, jT0, W;  // first coarse column and row held, row pitch
};
constexpr int kTPatchN = 1024;  // floats: 4 bands of up to 200 rows x 9 columns

struct TCache {
    int jT;        // coarse row of the cached terms (-1: none)
    float A, B;
    float x;
    uint32_t off;  // byte offset of the lane's coarse column iT in a row (of the patch)
    TPatch tp;
};

__device__ __forceinline__ TCache tcache_init(const StepArgsMixed& a, int i, const TPatch& tp) {
    TCache tc;
    tc.jT = -1;
    tc.A = tc.B = 0.f;
    const int g = i - 1 + a.ile;
    const int iT = g / kRatio;
    tc.x = (float)(g - kRatio * iT) * 0.1f;
    tc.off = (uint32_t)(tp.on ? iT - tp.iT0 : iT) * 4u;
    tc.tp = tp;
    return tc;
}

template <typename T>
__device__ __forceinline__ T ld_at(const T* base, uint32_t byte_off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

// Raw buffer views of the row-major fields: the descriptor (base, size)
// lives in scalar registers, the row offset is a 32-bit scalar and the lane
// offset a 32-bit vector register, so an access costs no 64-bit address
// arithmetic on the vector ALU. Word 3 of the descriptor is the gfx9
// format word (32-bit data format, no swizzle). Fields must stay below
// 4 GiB (checked at context creation).
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc rsrc_of(const void* p, uint64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0,
                                             (int)(uint32_t)bytes, 0x00020000);
}
__device__ __forceinline__ float buf_ldf(Rsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ double buf_ldd(Rsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// T' of fine row j (wavefront-uniform) in the lane's column i
__device__ __forceinline__ float temp_c(const StepArgsMixed& a, TCache& tc, int j, int i) {
    if (a.tfc) {
        const int g = j - 1 + a.dxTi;
        const int jT = g / kRatio;
        if (jT != tc.jT) {
            tc.jT = jT;
            float f00, f10, f01, f11;
            if (tc.tp.on) {
                const int q = (jT - tc.tp.jT0) * tc.tp.W + (int)(tc.off >> 2);
                f00 = tc.tp.p[q];
                f10 = tc.tp.p[q + 1];
                f01 = tc.tp.p[q + tc.tp.W];
                f11 = tc.tp.p[q + tc.tp.W + 1];
            } else {
                const float* r0 = a.T32 + (size_t)jT * a.NxT_all;
                const float* r1 = r0 + a.NxT_all;
                f00 = ld_at(r0, tc.off);
                f10 = ld_at(r0, tc.off + 4u);
                f01 = ld_at(r1, tc.off);
                f11 = ld_at(r1, tc.off + 4u);
            }
            const float a10 = f10 - f00, a01 = f01 - f00, a11 = (f11 - f10) - (f01 - f00);
            tc.A = fmaf(a11, tc.x, a01);
            tc.B = fmaf(a10, tc.x, f00);
        }
        const float y = (float)(g - kRatio * jT) * 0.1f;
        return fmaf(tc.A, y, tc.B);
    }
    return (float)fma(a.fta_b, (double)i, a.fta_a);
}

// table piece of T' (-1 outside the table)
__device__ __forceinline__ int sgte_piece(float Tp) {
    if (!(Tp >= RS_SGTE_T0 && Tp < RS_SGTE_T1)) return -1;
    return min((int)((Tp - RS_SGTE_T0) * 0.01f), RS_SGTE_PIECES - 1);
}

struct Uni {
    float vU, vTi, vNb;   // non-constant parts (t^1..t^5) of U_L, dG_Ti, dG_Nb
    double c0U;
    float c0Ti, c0Nb;
};

#define RS_H5(t, c1, c2, c3, c4, c5) \
    ((t) * fmaf(fmaf(fmaf(fmaf((c5), (t), (c4)), (t), (c3)), (t), (c2)), (t), (c1)))

#define RS_SGTE_CASE(P, CEN, ISC, K0, K1, K2, a1, a2, a3, a4, a5, b1, b2, b3, b4, b5, d1, d2, d3, \
                     d4, d5)                                                                      \
    case P: {                                                                                     \
        const float t = (Tp - (CEN)) * (ISC);                                                     \
        u.vU = RS_H5(t, a1, a2, a3, a4, a5);                                                      \
        u.vTi = RS_H5(t, b1, b2, b3, b4, b5);                                                     \
        u.vNb = RS_H5(t, d1, d2, d3, d4, d5);                                                     \
        u.c0U = (K0);                                                                             \
        u.c0Ti = (float)(K1);                                                                     \
        u.c0Nb = (float)(K2);                                                                     \
        break;                                                                                    \
    }

// evaluates the piece of the wavefront's first lane (uniform: immediates);
// lanes in another piece (only wavefronts straddling a piece boundary
// isotherm) are flagged for the exact routine
__device__ __forceinline__ Uni sgte_eval(float Tp, int& piece) {
    Uni u{0.f, 0.f, 0.f, 0.0, 0.f, 0.f};
    const int pp = __builtin_amdgcn_readfirstlane(piece);
    if (piece != pp) piece = -1;
    switch (pp) { RS_SGTE_TABLE(RS_SGTE_CASE) default: break; }
    return u;
}

// per-cell thermodynamics: FP32 driving force dG = f_L - f_S (the ideal
// mixing terms cancel between the phases) and the FP64 chemical potential
// mu = mu_L + h (mu_S - mu_L)
__device__ __forceinline__ void thermo_m(const StepArgsMixed& a, const double* lntab, float Tp,
                                         double c, float h, float& dG, double& mu) {
    int piece = sgte_piece(Tp);
    const Uni u = sgte_eval(Tp, piece);
    if (piece >= 0) {
        const double onemc = 1.0 - c;
        double lr = 0.0;
        if (c > 1.e-8 && onemc > 1.e-8) lr = ln_tab(ratio_c(c, onemc), lntab);
        const float cf = (float)c;
        const float omcf = (float)onemc;
        const float om2 = 1.0f - 2.0f * cf;
        const float dGTi = u.c0Ti + u.vTi;
        const float dGNb = u.c0Nb + u.vNb;
        const double om2d = 1.0 - 2.0 * c;
        double mu64 = u.c0U + a.L0l * om2d + a.RR * (kMixedTref + (double)Tp) * lr;
        // mu_S - mu_L = (dG_Nb - dG_Ti) - dL0 (1-2c) [- dL1 (1-6c+6c^2)]
        float dsl = (dGNb - dGTi) - a.dL0 * om2;
        float ex = a.dL0;
        if (a.has_L1) {
            const double sixd = 1.0 - 6.0 * c + 6.0 * (c * c);
            mu64 += a.L1l * sixd;
            dsl -= a.dL1 * (float)sixd;
            ex = fmaf(a.dL1, om2, ex);
        }
        dG = fmaf(cf * omcf, ex, fmaf(cf, dGTi, omcf * dGNb));
        mu = mu64 + (double)fmaf(h, dsl, u.vU);
    } else {
        const Thermo th = thermo_fp64(a.rk, kMixedTref + (double)Tp, c);
        dG = (float)th.dG;
        mu = th.muL + (double)h * (th.muS - th.muL);
    }
}

// learned-surrogate thermodynamics for the 64 cells of a wavefront row
// (called with every lane active: the MFMA forward pass is wavefront-wide).
// Cells outside the training domain or in a bin the trust table rejects
// take the exact routine; nfb counts them (wavefront-uniform). So do the two
// columns at each x edge and the two rows at each y edge: a ghost copies c
// and psi but its T is interpolated at its own position (fun.cu:1220-1230,
// 1585-1606), so the edge faces carry a flux driven only by that T
// difference, which the network's rounding noise would swamp and turn into
// a solute-mass drift. Lanes past the last column (duplicates of it) keep
// the network value, which nothing reads, and are not counted.
struct SurrLds {
    const MlpInfer* F;
    const uint64_t* trust;
};

template <int FMT, bool ALWAYS = false>
__device__ __forceinline__ void thermo_s(const StepArgsMixed& a, const double* lntab,
                                         const SurrLds& sl, float Tp, double c, float h, int i,
                                         int j, bool real, float& dG, double& mu, unsigned& nfb,
                                         unsigned& nall) {
    const int lane = threadIdx.x;  // blockDim.x == 64
    const float cf = (float)c;
    const float ain = fmaf(Tp, a.s_a_scale, a.s_a_off);
    const float bin = fmaf(-5.0f, cf, 4.0f);  // (x_Nb - 0.2)/0.2, x_Nb = 1 - c
    // an opaque zero offset keeps the fragment reads inside the row: hoisted
    // out of the row loop they would pin ~160 registers
    int zoff = 0;
    asm volatile("" : "+v"(zoff));
    const MlpInfer& F =
        *reinterpret_cast<const MlpInfer*>(reinterpret_cast<const char*>(sl.F) + zoff);
    bool ok = fabsf(ain) <= 1.0f && fabsf(bin) <= 1.0f && i > 1 && i < a.Nx - 2 && j > 1 &&
              j < a.Ny - 2;
    if (ok) {
        const int ia = min((int)((ain + 1.0f) * (0.5f * kTrustN)), kTrustN - 1);
        const int ib = min((int)((bin + 1.0f) * (0.5f * kTrustN)), kTrustN - 1);
        ok = (sl.trust[ib] >> ia) & 1ULL;
    }
    // a row with no trusted cell takes the exact routine in every cell, and
    // the network runs only if some cell of the row uses it (wavefront-
    // uniform). ALWAYS drops that test, for runs where such rows are rare:
    // the branch around the network costs ~4% of the step.
    const bool any = __ballot(ok && real) != 0;
    ok = any && (ok || !real);
    // counts (wavefront-uniform): fallback cells and, from bit 20, rows
    // without the network; all cells and, from bit 20, rows
    nfb += (unsigned)__popcll(__ballot(real && !ok)) + (any ? 0u : (1u << 20));
    nall += (unsigned)__popcll(__ballot(real)) + (1u << 20);
    mlp::f4 y = {0.f, 0.f, 0.f, 0.f};
    if (ALWAYS || any) y = mlp::forward<FMT>(F, mlp::input_frag(ain, bin), lane);
    if (ok) {
        const double onemc = 1.0 - c;
        double lr = 0.0;
        if (c > 1.e-8 && onemc > 1.e-8) lr = ln_tab(ratio_c(c, onemc), lntab);
        dG = fmaf(y[0], a.y_std[0], a.y_mean[0]);
        const float muL = fmaf(y[2], a.y_std[2], a.y_mean[2]);
        const float dmu = fmaf(y[3], a.y_std[3], a.y_mean[3]);
        mu = (double)fmaf(h, dmu, muL) + a.RR * (kMixedTref + (double)Tp) * lr;
    } else {
        thermo_m(a, lntab, Tp, c, h, dG, mu);
    }
}

// adds a wavefront's packed counts (cells in bits 0-19, rows from bit 20)
// to one counter slot
__device__ __forceinline__ void add_counts(unsigned long long* q, unsigned nfb, unsigned nall) {
    atomicAdd(&q[0], (unsigned long long)(nfb & 0xFFFFFu));
    atomicAdd(&q[1], (unsigned long long)(nall & 0xFFFFFu));
    atomicAdd(&q[2], (unsigned long long)(nfb >> 20));
    atomicAdd(&q[3], (unsigned long long)(nall >> 20));
}

struct CellM {
    float psi, phi, dG, S;
    double c, mu;
};

template <int SURR>
__device__ __forceinline__ CellM make_cell_m(const StepArgsMixed& a, const double* lntab,
                                             const SurrLds& sl, float Tp, float psi,
                                             double c, int i, int j, bool real, unsigned& nfb,
                                             unsigned& nall) {
    CellM q;
    q.psi = psi;
    q.c = c;
    q.phi = fmaxf(psi, -1.0f);
    const float cf = (float)c;
    q.S = qq<float>(a.AA, a.q_solid, psi) * (cf * (1.0f - cf));
    const float h = 0.5f * (1.0f + gg<float>(psi));
    if constexpr (SURR != 0)
        thermo_s<(SURR - 1) & 3, (SURR > 4)>(a, lntab, sl, Tp, c, h, i, j, real, q.dG, q.mu, nfb,
                                            nall);
    else
        thermo_m(a, lntab, Tp, c, h, q.dG, q.mu);
    return q;
}

__device__ __forceinline__ float quad4(float a0, float a1, float a2, float a3) {
    return 0.25f * (((a0 + a1) + a2) + a3);
}

// FP32 anisotropic operator (fun.cu:1030-1139) with precomputed reciprocals
template <bool IDENT>
__device__ __forceinline__ float aniso_f32(const StepArgsMixed& A, float pmm, float p0m, float p1m,
                                           float pm0, float p00, float p10, float pm1, float p01,
                                           float p11, float& as, float& ak) {
    const float gx = (p10 - pm0) * A.inv2dx;
    const float gy = (p01 - p0m) * A.inv2dx;
    const float g2 = fmaf(gx, gx, gy * gy);
    const float gxx = ((p10 - 2.0f * p00) + pm0) * A.invdx2;
    const float gyy = ((p01 - 2.0f * p00) + p0m) * A.invdx2;
    const float gxy = (((p11 - pm1) - p1m) + pmm) * A.inv4dx2;
    float lap = A.lap_w_axis * (gxx + gyy);
    lap = fmaf(A.lap_w_diag * ((((p11 + pm1) + p1m) + pmm) - 4.0f * p00), A.inv2dx2, lap);
    float thx = 0.f, thy = 0.f, ct = 0.f, st = 0.f;
    if (g2 > 1.0e-30f) {
        const float ir = rsqrtf(g2);
        const float ir2 = ir * ir;
        thx = fmaf(gx, gxy, -gy * gxx) * ir2;
        thy = fmaf(gx, gyy, -gy * gxy) * ir2;
        ct = gx * ir;
        st = gy * ir;
    }
    float np4, d1, d2;
    if (IDENT) {
        // n' = (cos, sin, 0), dn'/dtheta = (-sin, cos, 0)
        const float c2 = ct * ct, s2 = st * st;
        np4 = fmaf(c2, c2, s2 * s2);
        d1 = 4.0f * (ct * st) * (s2 - c2);
        d2 = 4.0f * fmaf(6.0f * c2, s2, -np4);
    } else {
        const float nx = fmaf(A.ops[0], ct, A.ops[1] * st);
        const float ny = fmaf(A.ops[2], ct, A.ops[3] * st);
        const float nz = fmaf(A.ops[4], ct, A.ops[5] * st);
        const float ex = fmaf(-A.ops[0], st, A.ops[1] * ct);
        const float ey = fmaf(-A.ops[2], st, A.ops[3] * ct);
        const float ez = fmaf(-A.ops[4], st, A.ops[5] * ct);
        const float nx2 = nx * nx, ny2 = ny * ny, nz2 = nz * nz;
        const float nx3 = nx2 * nx, ny3 = ny2 * ny, nz3 = nz2 * nz;
        np4 = fmaf(nx2, nx2, fmaf(ny2, ny2, nz2 * nz2));
        d1 = 4.0f * fmaf(nx3, ex, fmaf(ny3, ey, nz3 * ez));
        const float xe = nx * ex, ye = ny * ey, ze = nz * ez;
        d2 = fmaf(3.0f, xe * xe, -nx3 * nx);
        d2 += fmaf(3.0f, ye * ye, -ny3 * ny);
        d2 += fmaf(3.0f, ze * ze, -nz3 * nz);
        d2 *= 4.0f;
    }
    ak = fmaf(4.0f * A.epk1, np4, 1.0f - 3.0f * A.epk1);
    as = fmaf(4.0f * A.eps1, np4, 1.0f - 3.0f * A.eps1);
    const float da1 = 4.0f * A.eps1 * d1;
    const float da2 = 4.0f * A.eps1 * d2;
    return fmaf(as * as, lap,
                fmaf(2.0f * as * da1, fmaf(thx, gx, thy * gy),
                     fmaf(da2, as, da1 * da1) * fmaf(thy, gx, -thx * gy)));
}

// one fine row of the 3-row window: the cell quantities and the values the
// row needs from its i-1 / i+1 neighbours
struct Slot {
    CellM q;
    float phL, phR, SR;
};

template <int SURR>
__device__ __forceinline__ void fill_slot(const StepArgsMixed& a, const double* lntab,
                                          const SurrLds& sl, TCache& tc, int j, float psi,
                                          double c, int ic, bool real, Slot& s, unsigned& nfb,
                                          unsigned& nall) {
    s.q = make_cell_m<SURR>(a, lntab, sl, temp_c(a, tc, j, ic), psi, c, ic, j, real, nfb, nall);
    s.phL = shup(s.q.phi);
    s.phR = shdn(s.q.phi);
    s.SR = shdn(s.q.S);
}

// LDS tables of the step: the FP64 log table and, with the surrogate, the
// network fragments and trust bitmap
template <int SURR>
struct StepLds;
template <>
struct StepLds<0> {
    double lntab[2 * RS_LN_TAB_N];
};
template <>
struct StepLds<1> {
    double lntab[2 * RS_LN_TAB_N];
    MlpInfer mlp;
    uint64_t trust[kTrustN];
};
template <>
struct StepLds<2> : StepLds<1> {};
template <>
struct StepLds<3> : StepLds<1> {};
template <>
struct StepLds<4> : StepLds<1> {};
template <>
struct StepLds<7> : StepLds<1> {};
template <>
struct StepLds<8> : StepLds<1> {};

template <int SURR>
__device__ __forceinline__ void load_step_lds(StepLds<SURR>& L, const MlpInfer* mlp,
                                              const uint64_t* trust, int tid, int nthr) {
    for (int q = tid; q < 2 * RS_LN_TAB_N; q += nthr) L.lntab[q] = kLnTab[q];
    if constexpr (SURR != 0) {
        const uint4* src = reinterpret_cast<const uint4*>(mlp);
        uint4* dst = reinterpret_cast<uint4*>(&L.mlp);
        for (int q = tid; q < (int)(sizeof(MlpInfer) / 16); q += nthr) dst[q] = src[q];
        for (int q = tid; q < kTrustN; q += nthr) L.trust[q] = trust[q];
    }
}

template <int SURR>
__device__ __forceinline__ SurrLds surr_lds(StepLds<SURR>& L) {
    if constexpr (SURR != 0)
        return SurrLds{&L.mlp, L.trust};
    else
        return SurrLds{nullptr, nullptr};
}

// One wavefront's work: column tile tile_x (62 output columns), band `band`
// (rows 1 + band*band_h ...), marching down the band with a 3-row window.
// BUF: the rows the window loads go through raw buffer views (no 64-bit
// address arithmetic on the vector ALU); the stores use plain pointers, as
// two more descriptors would add 8 scalar registers to a kernel that
// already spills scalars (with them: 39 spills instead of 24, 2.7% slower at
// 4096^2). The surrogate kernels and the persistent batch kernel, whose
// scalar registers already hold the network or run state, load through
// plain pointers too: measured 3% faster there. Both address the same
// elements.
// LED: the solute ledger runs (a.led set): the boundary faces' fluxes are
// recorded. A template flag, not a branch: the record costs registers in
// these register-limited kernels even when it is skipped (4% of a step).
template <bool IDENT, int SURR, bool BUF = true, bool LED = false>
__device__ __forceinline__ void step_tile(const StepArgsMixed& a, const double* lntab,
                                          const SurrLds& sl, int tile_x, int band, unsigned& nfb,
                                          unsigned& nall, const TPatch& tp) {
    const int lane = threadIdx.x;
    // wavefront-uniform (blockDim.x is the wavefront): row addresses then
    // live in scalar registers and each access adds a 32-bit lane offset
    band = __builtin_amdgcn_readfirstlane(band);
    tile_x = __builtin_amdgcn_readfirstlane(tile_x);
    if (band >= a.nbands) return;
    const int i = tile_x * kTileOut + lane;
    const int ic = i < a.Nx ? i : a.Nx - 1;
    const bool own = lane >= 1 && lane <= kTileOut && i <= a.Nx - 2;
    const int jb = 1 + band * a.band_h;
    const int je = min(jb + a.band_h - 1, a.Ny - 2);
    const int Nx = a.Nx;
    const uint64_t cells = (uint64_t)Nx * a.Ny;
    const uint32_t o4 = (uint32_t)ic * 4u, o8 = (uint32_t)ic * 8u;
    auto psi_row = [&](int r) {
        if constexpr (BUF)
            return buf_ldf(rsrc_of(a.psi, 4 * cells), o4, (uint32_t)(r * Nx) * 4u);
        else
            return a.psi[(size_t)r * Nx + ic];
    };
    auto c_row = [&](int r) {
        if constexpr (BUF)
            return buf_ldd(rsrc_of(a.c, 8 * cells), o8, (uint32_t)(r * Nx) * 8u);
        else
            return a.c[(size_t)r * Nx + ic];
    };

    Slot s0, s1, s2;
    TCache tc = tcache_init(a, ic, tp);
    const bool real = i < a.Nx;  // lanes past the last column duplicate it
    fill_slot<SURR>(a, lntab, sl, tc, jb - 1, psi_row(jb - 1), c_row(jb - 1), ic, real, s0, nfb,
                    nall);
    fill_slot<SURR>(a, lntab, sl, tc, jb, psi_row(jb), c_row(jb), ic, real, s1, nfb, nall);
    // vertex (i+1/2, jb-1/2) and the y-face flux (i, jb-1/2)
    float Vpm = quad4(s0.q.S, s0.SR, s1.q.S, s1.SR);
    const float Vmm = shup(Vpm);
    double Fy_prev = (double)quad4(s0.q.S, s1.q.S, Vpm, Vmm) * (s1.q.mu - s0.q.mu);

    float nx_psi = psi_row(jb + 1);
    double nx_c = c_row(jb + 1);
    float lh_acc = 0.0f;

    // output row j from rows M = j-1, Z = j and P = j+1 (filled here)
    auto row = [&](int j, const Slot& M, const Slot& Z, Slot& P) {
        const float cur_psi = nx_psi;
        const double cur_c = nx_c;
        // unconditional (the last row reloads row je+1): with the load under
        // a branch the compiler waited for every outstanding memory operation,
        // the previous row's stores included, at the top of each row (9%)
        {
            const int jn = min(j + 2, je + 1);
            nx_psi = psi_row(jn);
            nx_c = c_row(jn);
        }
        fill_slot<SURR>(a, lntab, sl, tc, j + 1, cur_psi, cur_c, ic, real, P, nfb, nall);
        const double muR_z = shdn(Z.q.mu);
        // vertex (i+1/2, j+1/2), its left neighbour, and the faces
        const float Vpp = quad4(Z.q.S, Z.SR, P.q.S, P.SR);
        const float Vmp = shup(Vpp);
        const double Fx = (double)quad4(Z.q.S, Z.SR, Vpp, Vpm) * (muR_z - Z.q.mu);    // (i+1/2, j)
        const double Fxm = shup(Fx);                                                  // (i-1/2, j)
        const double Fy = (double)quad4(Z.q.S, P.q.S, Vpp, Vmp) * (P.q.mu - Z.q.mu);  // (i, j+1/2)

        if (own) {
            const float p00 = Z.q.phi;
            // inside a bulk phase (phi = +-1 over the 3x3 stencil, finite
            // driving force) phi - phi^3, the anisotropic operator and g'(phi)
            // all vanish exactly and the update returns p00 bit for bit:
            // wavefronts whose own cells are all such skip it
            const bool bulk = (p00 == 1.0f || p00 == -1.0f) && M.phL == p00 && M.q.phi == p00 &&
                              M.phR == p00 && Z.phL == p00 && Z.phR == p00 && P.phL == p00 &&
                              P.q.phi == p00 && P.phR == p00 && isfinite(Z.q.dG);
            float psn = p00;
            if (__ballot(!bulk) != 0) {
                float as, ak;
                const float an = aniso_f32<IDENT>(a, M.phL, M.q.phi, M.phR, Z.phL, Z.q.phi, Z.phR,
                                                  P.phL, P.q.phi, P.phR, as, ak);
                float term = fmaf(-(p00 * p00), p00, p00) + an;
                term = fmaf(Z.q.dG * dgg<float>(p00), a.dG_den_inv, term);
                float phn = fmaf(a.dt * ak * rcpf(as * as), term, p00);
                psn = phn;
                if (psn > 1.0f) {
                    phn -= (psn - 1.0f);
                    psn = phn;
                }
            }
            const double dc = a.K * ((Fx - Fxm) + (Fy - Fy_prev));
            const double cn = fmax(Z.q.c + dc, 1.0e-6);
            if constexpr (LED) {  // the flux through the boundary faces (ghost side)
                if (i == 1) a.led[j] -= a.K * Fxm;
                if (i == Nx - 2) a.led[a.Ny + j] += a.K * Fx;
                if (j == 1) a.led[2 * a.Ny + i] -= a.K * Fy_prev;
                if (j == a.Ny - 2) a.led[2 * a.Ny + Nx + i] += a.K * Fy;
            }

            float* prow = a.psi_n + (size_t)j * Nx;
            double* crow = a.c_n + (size_t)j * Nx;
            prow[i] = psn;
            crow[i] = cn;
            if (i == 1) {
                prow[0] = psn;
                crow[0] = cn;
            }
            if (i == Nx - 2) {
                prow[Nx - 1] = psn;
                crow[Nx - 1] = cn;
            }
            if (j == a.Ny - 2 || j == 1) {
                const int g = (j == 1) ? (a.Ny - 1) * Nx : 0;
                a.psi_n[g + i] = psn;
                a.c_n[g + i] = cn;
                if (i == 1) {
                    a.psi_n[g] = psn;
                    a.c_n[g] = cn;
                }
                if (i == Nx - 2) {
                    a.psi_n[g + Nx - 1] = psn;
                    a.c_n[g + Nx - 1] = cn;
                }
            }
            if (a.lh_on && j <= a.lh_last_row) {
                lh_acc += psn - Z.q.psi;
                if (j % 10 == 0) {
                    a.lh[(size_t)(j / 10 - 1) * Nx + i] = lh_acc;
                    lh_acc = 0.0f;
                }
            }
        }
        Vpm = Vpp;
        Fy_prev = Fy;
    };

    // three rows per trip with the slot roles rotated: no register copies
    // of the window between rows
    for (int j = jb; j <= je; j += 3) {
        row(j, s0, s1, s2);
        if (j + 1 > je) break;
        row(j + 1, s1, s2, s0);
        if (j + 2 > je) break;
        row(j + 2, s2, s0, s1);
    }
}

// Loads into buf the coarse T' patch the wavefronts of a workgroup
// interpolate from: column tile `tile`, bands band0 .. band0 + nb - 1, rows
// jb - 1 .. je + 1 of each (the coarse rows and columns temp_c reaches for
// them, plus the next row and column). Not loaded (on = 0, temp_c reads the
// global grid) without coupled T or when it exceeds kTPatchN floats.
__device__ __forceinline__ TPatch tpatch_load(const StepArgsMixed& a, float* buf, int tile,
                                              int band0, int nb, int tid, int nthr) {
    TPatch t{(LdsF)buf, 0, 0, 0, 0};
    if (!a.tfc || band0 >= a.nbands) return t;
    const int b1 = min(band0 + nb, a.nbands) - 1;
    const int r_lo = band0 * a.band_h;                                    // jb - 1, first band
    const int r_hi = min(1 + (b1 + 1) * a.band_h - 1, a.Ny - 2) + 1;      // je + 1, last band
    const int jT0 = (r_lo - 1 + a.dxTi) / kRatio;
    const int jT1 = (r_hi - 1 + a.dxTi) / kRatio + 1;
    const int i_lo = tile * kTileOut;
    const int i_hi = min(tile * kTileOut + kWave - 1, a.Nx - 1);
    const int iT0 = (i_lo - 1 + a.ile) / kRatio;
    const int iT1 = (i_hi - 1 + a.ile) / kRatio + 1;
    const int W = iT1 - iT0 + 1, H = jT1 - jT0 + 1;
    if (W * H > kTPatchN) return t;
    for (int q = tid; q < W * H; q += nthr) {
        const int y = q / W;
        buf[q] = a.T32[(size_t)(jT0 + y) * a.NxT_all + iT0 + (q - y * W)];
    }
    t.on = 1;
    t.iT0 = iT0;
    t.jT0 = jT0;
    t.W = W;
    return t;
}

// Dispatch order of the single-run exact-table kernel. The dispatcher hands
// workgroup w to XCD w % kXcds; the band groups are split into kXcds strips
// of consecutive rows, one per XCD (neighbouring tiles share an L2), and a
// strip's (column tile, band group) items go out tile by tile, starting at
// the side of the grid the front is held at (a.tiles_desc: from the last
// tile down). The launch ends with the last items dispatched: at 4096^2 the
// tiles of the interface and of the solute boundary layer ahead of it take
// up to 1.8x the median tile's time, and row by row (the last band group of
// a strip holds them too) the step took 151 us against 134-135 us tile by
// tile. Every item of a strip appears once, so the order changes no result.
__device__ __forceinline__ bool strip_item(const StepArgsMixed& a, int ntx, int nbg, int w,
                                           int& tile, int& bg) {
    const int xcd = w % kXcds, k = w / kXcds;
    const int bg0 = xcd * nbg / kXcds, nb = (xcd + 1) * nbg / kXcds - bg0;
    if (nb == 0 || k >= nb * ntx) return false;
    bg = bg0 + k % nb;
    tile = a.tiles_desc ? ntx - 1 - k / nb : k / nb;
    return true;
}

// SURR: 0 exact thermodynamics, else the surrogate with activation format
// (SURR - 1) & 3 (kActBf16, kActSplit, kActFp16, kActFp16Pk); 7 and 8: FP16
// and packed FP16 with the network evaluated in every row. 4 wavefronts per SIMD (<= 128
// VGPRs, no spills) for the exact thermodynamics: at 3 (143 VGPRs) the
// 4096^2 step takes 0.191 ms instead of 0.166 ms. The surrogate variants
// hold the MFMA operands and run at 2 (3 spills 52 registers).
template <bool IDENT, int SURR, bool LED>
__global__ __launch_bounds__(256, SURR ? 2 : 4) void step_mixed_kernel(StepArgsMixed a) {
    __shared__ StepLds<SURR> L;
    // the coarse T' patch pays in the surrogate kernels (2%); the exact-table
    // kernel, whose row is short, measured no gain and reads the global grid
    __shared__ float tpatch[SURR != 0 ? kTPatchN : 1];
    const int tid = threadIdx.y * blockDim.x + threadIdx.x, nthr = blockDim.x * blockDim.y;
    if constexpr (SURR == 0) {  // 1D grid in strip order (launch_step_mixed)
        const int ntx = (a.Nx - 2 + kTileOut - 1) / kTileOut;
        const int nbg = (a.nbands + (int)blockDim.y - 1) / (int)blockDim.y;
        int tile, bg;
        if (!strip_item(a, ntx, nbg, blockIdx.x, tile, bg)) return;  // whole workgroup
        load_step_lds<SURR>(L, a.mlp, a.trust, tid, nthr);
        __syncthreads();
        unsigned nfb = 0, nall = 0;
        step_tile<IDENT, SURR, true, LED>(a, L.lntab, surr_lds<SURR>(L), tile,
                                          bg * blockDim.y + threadIdx.y, nfb, nall, TPatch{});
    } else {  // 2D grid: (column tile, band group) in XCD order
        load_step_lds<SURR>(L, a.mlp, a.trust, tid, nthr);
        const int ntx = gridDim.x;
        const int g = xcd_order(blockIdx.y * ntx + blockIdx.x, ntx * gridDim.y);
        const TPatch tp =
            tpatch_load(a, tpatch, g % ntx, (g / ntx) * blockDim.y, blockDim.y, tid, nthr);
        __syncthreads();
        unsigned nfb = 0, nall = 0;
        step_tile<IDENT, SURR, false, LED>(a, L.lntab, surr_lds<SURR>(L), g % ntx,
                                           (g / ntx) * blockDim.y + threadIdx.y, nfb, nall, tp);
        if (threadIdx.x == 0 && nall > 0)
            add_counts(a.fb_count + (size_t)(g % kCountSlots) * kCountStride, nfb, nall);
    }
}

template <bool IDENT, int SURR>
static void launch_mk(dim3 grid, dim3 block, hipStream_t s, const StepArgsMixed& a, LaunchEv ev) {
    if (a.led)
        hipExtLaunchKernelGGL((step_mixed_kernel<IDENT, SURR, true>), grid, block, 0, s, ev.start,
                              ev.stop, 0, a);
    else
        hipExtLaunchKernelGGL((step_mixed_kernel<IDENT, SURR, false>), grid, block, 0, s, ev.start,
                              ev.stop, 0, a);
}

hipError_t launch_step_mixed(const StepArgsMixed& a, int wpb, bool ident, bool surr, hipStream_t s,
                             int* variant, LaunchEv ev) {
    if (variant)
        *variant = !surr                                 ? 0
                   : a.s_fmt == kActFp16Pk && a.s_always ? 8
                   : a.s_fmt == kActFp16Pk               ? 4
                   : a.s_fmt == kActFp16 && a.s_always   ? 7
                   : a.s_fmt == kActFp16                 ? 3
                   : a.s_fmt == kActSplit                ? 2
                                                         : 1;
    if (wpb < 1) wpb = 1;
    if (wpb > 4) wpb = 4;
    dim3 block(kWave, wpb);
    dim3 grid((a.Nx - 2 + kTileOut - 1) / kTileOut, (a.nbands + wpb - 1) / wpb);
    if (!surr) {  // strip order: kXcds strips of band groups, the longest sets the grid
        int mx = 0;
        for (int x = 0; x < kXcds; ++x)
            mx = std::max(mx, ((x + 1) * (int)grid.y / kXcds - x * (int)grid.y / kXcds) * (int)grid.x);
        grid = dim3(kXcds * mx, 1);
    }
    if (surr && a.s_fmt == kActFp16Pk && a.s_always) {
        if (ident)
            launch_mk<true, 8>(grid, block, s, a, ev);
        else
            launch_mk<false, 8>(grid, block, s, a, ev);
    } else if (surr && a.s_fmt == kActFp16Pk) {
        if (ident)
            launch_mk<true, 4>(grid, block, s, a, ev);
        else
            launch_mk<false, 4>(grid, block, s, a, ev);
    } else if (surr && a.s_fmt == kActFp16 && a.s_always) {
        if (ident)
            launch_mk<true, 7>(grid, block, s, a, ev);
        else
            launch_mk<false, 7>(grid, block, s, a, ev);
    } else if (surr && a.s_fmt == kActFp16) {
        if (ident)
            launch_mk<true, 3>(grid, block, s, a, ev);
        else
            launch_mk<false, 3>(grid, block, s, a, ev);
    } else if (surr && a.s_fmt == kActSplit) {
        if (ident)
            launch_mk<true, 2>(grid, block, s, a, ev);
        else
            launch_mk<false, 2>(grid, block, s, a, ev);
    } else if (surr) {
        if (ident)
            launch_mk<true, 1>(grid, block, s, a, ev);
        else
            launch_mk<false, 1>(grid, block, s, a, ev);
    } else {
        if (ident)
            launch_mk<true, 0>(grid, block, s, a, ev);
        else
            launch_mk<false, 0>(grid, block, s, a, ev);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// batched ensemble
// ---------------------------------------------------------------------------
// Vp_ramp (fun.cu:1974-1981) as the host evaluates it
__device__ __forceinline__ double batch_vp(const BatchRun& R, double t) {
    if (R.ramp) {
        t = t * 0.001;
        return R.Vp + R.Vp_rate * t;
    }
    return R.Vp;
}

__device__ __forceinline__ int batch_run_of(const int* pre, int nruns, int q) {
    int r = 0;
    while (r + 1 < nruns && q >= pre[r + 1]) ++r;
    return r;
}

// Persistent: each workgroup loads the LDS tables once and walks the work
// list (run, column tile, band group) with a fixed stride, so the assignment
// and the results do not depend on timing. Exact thermodynamics at 3
// wavefronts per SIMD: the per-item run state spills at 4 (64-run ensemble:
// 3 waves 0.425 ms per step, 4 waves with 10 spilled registers 0.49-0.52 ms);
// plain pointers measured faster than buffer views here (0.49-0.51 ms).
template <bool IDENT, int SURR, bool LED>
__global__ __launch_bounds__(256, SURR ? 2 : 3) void batch_step_kernel(BatchLaunch L) {
    __shared__ StepLds<SURR> lds;
    load_step_lds<SURR>(lds, L.runs[0].a.mlp, L.runs[0].a.trust,
                        threadIdx.y * blockDim.x + threadIdx.x, blockDim.x * blockDim.y);
    __syncthreads();
    for (int item = blockIdx.x; item < L.items; item += gridDim.x) {
        const int r = batch_run_of(L.item0, L.nruns, item);
        const BatchRun& R = L.runs[r];
        const BatchDyn d = L.dyn[r];
        StepArgsMixed a = R.a;
        a.psi = R.psi[d.cur];
        a.c = R.c[d.cur];
        a.psi_n = R.psi[d.cur ^ 1];
        a.c_n = R.c[d.cur ^ 1];
        a.T32 = R.T32[d.tcur];
        a.ile = d.ile_interp;
        if (!R.tfc) {
            // frozen T of the current frame (fun.cu:1610-1619)
            const double dl = R.GG * (d.l_offset_Vp - R.l_off0) * R.S * R.W0 / R.deltaT0;
            a.fta_a = R.Tmelt + R.deltaT0 * (R.delta_initial - dl) + R.gstep * R.ioff - kMixedTref;
        }
        // a run's items tile by tile (its band groups of one column tile in
        // a row): 64-run ensemble 1.8% (FP32 table) and 3.6% (surrogate)
        // faster than row by row
        const int local = item - L.item0[r];
        const int tx = local / R.nbg;
        const int bg = local % R.nbg;
        unsigned nfb = 0, nall = 0;
        // runs without a ledger take the plain tile (a.led is null for them)
        if (LED && a.led)
            step_tile<IDENT, SURR, false, true>(a, lds.lntab, surr_lds<SURR>(lds), tx,
                                                bg * blockDim.y + threadIdx.y, nfb, nall, TPatch{});
        else
            step_tile<IDENT, SURR, false, false>(a, lds.lntab, surr_lds<SURR>(lds), tx,
                                                 bg * blockDim.y + threadIdx.y, nfb, nall, TPatch{});
        if constexpr (SURR != 0) {
            if (threadIdx.x == 0 && nall > 0)
                add_counts(a.fb_count + (size_t)(item % kCountSlots) * kCountStride, nfb, nall);
        }
    }
}

// coarse-T sub-step `sub` of every coupled run that takes that many
// (ratio_dt_dtT per step, fun.cu:496): sub-step s reads parity tcur ^ (s & 1)
// and writes the other, as the host loop's buffer swaps do
__global__ __launch_bounds__(256) void batch_temperature_kernel(BatchLaunch L, int sub) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < L.cells; q += gridDim.x * blockDim.x) {
        const int r = batch_run_of(L.cell0, L.nruns, q);
        const BatchRun& R = L.runs[r];
        if (sub >= R.t.ratio_dt_dtT) continue;
        const BatchDyn d = L.dyn[r];
        const int src = d.tcur ^ (sub & 1);
        TempArgs t = R.t;
        t.T = R.T[src];
        t.Tn = R.T[src ^ 1];
        t.T32n = R.T32[src ^ 1];
        t.vadv = batch_vp(R, d.step * R.tau0 * R.dt) * R.tau0 / R.SW0;
        const int local = q - L.cell0[r];
        temperature_cell(t, local % t.NxT, local / t.NxT);
    }
}

// host-loop state update of fun.cu:498-509 for every run
__global__ void batch_advance_kernel(BatchLaunch L) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= L.nruns) return;
    const BatchRun& R = L.runs[r];
    BatchDyn d = L.dyn[r];
    if (R.tfc) {
        d.tcur ^= R.t.ratio_dt_dtT & 1;
        d.ile_interp = R.t.ile;
    }
    d.cur ^= 1;
    d.l_offset_Vp += batch_vp(R, d.step * R.tau0 * R.dt) * R.dt * R.tau0 / R.SW0;
    d.step += 1;
    L.dyn[r] = d;
}

// the antitrapping pass of the runs that have it, over the step's work list
// (same bands; the band split does not change the result)
__global__ __launch_bounds__(256) void batch_antitrap_kernel(BatchLaunch L) {
    for (int item = blockIdx.x; item < L.items; item += gridDim.x) {
        const int r = batch_run_of(L.item0, L.nruns, item);
        const BatchRun& R = L.runs[r];
        if (!R.at_on) continue;
        const BatchDyn d = L.dyn[r];
        const int local = item - L.item0[r];
        const int tx = local % R.ntx;
        const int band = (local / R.ntx) * blockDim.y + threadIdx.y;
        if (band >= R.a.nbands) continue;
        const int jb = 1 + band * R.a.band_h;
        const int je = min(jb + R.a.band_h - 1, R.a.Ny - 2);
        antitrap_band<float>(R.psi[d.cur], R.psi[d.cur ^ 1], R.c[d.cur], R.c[d.cur ^ 1], R.a.Nx,
                             R.a.Ny, R.at, tx, jb, je);
    }
}

template <bool IDENT, int SURR>
static hipError_t batch_step_t(const BatchLaunch& L, hipStream_t s) {
    if (L.led_any)
        hipLaunchKernelGGL((batch_step_kernel<IDENT, SURR, true>), dim3(L.step_blocks),
                           dim3(kWave, L.wpb), 0, s, L);
    else
        hipLaunchKernelGGL((batch_step_kernel<IDENT, SURR, false>), dim3(L.step_blocks),
                           dim3(kWave, L.wpb), 0, s, L);
    return hipGetLastError();
}

hipError_t launch_batch_step(const BatchLaunch& L, hipStream_t s) {
    hipError_t e;
    switch (L.surr * 2 + (L.ident ? 1 : 0)) {
    case 0: e = batch_step_t<false, 0>(L, s); break;
    case 1: e = batch_step_t<true, 0>(L, s); break;
    case 2: e = batch_step_t<false, 1>(L, s); break;
    case 3: e = batch_step_t<true, 1>(L, s); break;
    case 4: e = batch_step_t<false, 2>(L, s); break;
    case 5: e = batch_step_t<true, 2>(L, s); break;
    case 6: e = batch_step_t<false, 3>(L, s); break;
    case 7: e = batch_step_t<true, 3>(L, s); break;
    case 8: e = batch_step_t<false, 4>(L, s); break;
    case 9: e = batch_step_t<true, 4>(L, s); break;
    case 16: e = batch_step_t<false, 8>(L, s); break;
    default: e = batch_step_t<true, 8>(L, s); break;
    }
    if (e != hipSuccess) return e;
    if (L.at_any) {
        hipLaunchKernelGGL(batch_antitrap_kernel, dim3(L.step_blocks), dim3(kWave, L.wpb), 0, s, L);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    for (int sub = 0; L.cells > 0 && sub < L.nsub; ++sub) {
        hipLaunchKernelGGL(batch_temperature_kernel, dim3(L.T_blocks), dim3(256), 0, s, L, sub);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(batch_advance_kernel, dim3((L.nruns + 63) / 64), dim3(64), 0, s, L);
    return hipGetLastError();
}

template <bool IDENT, int SURR>
static int resident_t(int wpb, bool led) {
    int n = 0;
    const hipError_t e =
        led ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_step_kernel<IDENT, SURR, true>,
                                                           kWave * wpb, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_step_kernel<IDENT, SURR, false>,
                                                           kWave * wpb, 0);
    if (e != hipSuccess) return 1;
    return n > 0 ? n : 1;
}

int batch_step_resident(const BatchLaunch& L) {
    switch (L.surr * 2 + (L.ident ? 1 : 0)) {
    case 0: return resident_t<false, 0>(L.wpb, L.led_any != 0);
    case 1: return resident_t<true, 0>(L.wpb, L.led_any != 0);
    case 2: return resident_t<false, 1>(L.wpb, L.led_any != 0);
    case 3: return resident_t<true, 1>(L.wpb, L.led_any != 0);
    case 4: return resident_t<false, 2>(L.wpb, L.led_any != 0);
    case 5: return resident_t<true, 2>(L.wpb, L.led_any != 0);
    case 6: return resident_t<false, 3>(L.wpb, L.led_any != 0);
    case 7: return resident_t<true, 3>(L.wpb, L.led_any != 0);
    case 8: return resident_t<false, 4>(L.wpb, L.led_any != 0);
    case 9: return resident_t<true, 4>(L.wpb, L.led_any != 0);
    case 16: return resident_t<false, 8>(L.wpb, L.led_any != 0);
    default: return resident_t<true, 8>(L.wpb, L.led_any != 0);
    }
}

void sgte_table_info(double* T_lo, double* T_hi, double* max_err) {
    *T_lo = kMixedTref + (double)RS_SGTE_T0;
    *T_hi = kMixedTref + (double)RS_SGTE_T1;
    *max_err = RS_SGTE_MAXERR;
}

// ---------------------------------------------------------------------------
// small kernels of the mixed path
// ---------------------------------------------------------------------------
__global__ void t32_copy_kernel(const double* T, float* T32, int64_t n, double Tref) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) T32[q] = (float)(T[q] - Tref);
}

hipError_t launch_T32_copy(const double* T, float* T32, int64_t n, double Tref, hipStream_t s) {
    hipLaunchKernelGGL(t32_copy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, T, T32,
                       n, Tref);
    return hipGetLastError();
}

__global__ void init_fields_f32_kernel(float* psi, double* c, const double* x0, int Nx, int Ny,
                                       double dx, int mode, double cinf, double c_liq,
                                       double c_sol) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= Nx || j >= Ny) return;
    const double itip = x0[j];
    const double ph = -tanh((i - itip) * dx / sqrt(2.0));
    double cv;
    if (mode == 0) {
        const double cc = 0.3184 * tanh(((i - itip) * dx - 0.08) / 0.67) + 0.6816;
        cv = cc * cinf;
    } else {
        cv = 0.5 * (1.0 + ph) * c_sol + 0.5 * (1.0 - ph) * c_liq;
    }
    psi[(int64_t)j * Nx + i] = (float)ph;
    c[(int64_t)j * Nx + i] = cv;
}

hipError_t launch_init_fields_f32(float* psi, double* c, const double* x0, int Nx, int Ny,
                                  double dx, int mode, double cinf, double c_liq, double c_sol,
                                  hipStream_t s) {
    dim3 block(64, 4);
    dim3 grid((Nx + 63) / 64, (Ny + 3) / 4);
    hipLaunchKernelGGL(init_fields_f32_kernel, grid, block, 0, s, psi, c, x0, Nx, Ny, dx, mode,
                       cinf, c_liq, c_sol);
    return hipGetLastError();
}

__global__ void pull_back_f32_kernel(const float* psi, const double* c, float* psi_n, double* c_n,
                                     int Nx, int Ny, int off, double cinf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= Nx || j >= Ny) return;
    int in = i + off;
    in = max(in, 1);
    in = min(in, Nx - 2);
    const int64_t row = (int64_t)j * Nx;
    c_n[row + i] = (in < Nx - 2) ? c[row + in] : cinf;
    psi_n[row + i] = psi[row + in];
}

hipError_t launch_pull_back_f32(const float* psi, const double* c, float* psi_n, double* c_n,
                                int Nx, int Ny, int off, double cinf, hipStream_t s) {
    dim3 block(64, 4);
    dim3 grid((Nx + 63) / 64, (Ny + 3) / 4);
    hipLaunchKernelGGL(pull_back_f32_kernel, grid, block, 0, s, psi, c, psi_n, c_n, Nx, Ny, off,
                       cinf);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void tip_rows_f32_kernel(const float* psi, int Nx, int Ny,
                                                           float pin, int* rowmax) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (j >= Ny) return;
    const float* row = psi + (int64_t)j * Nx;
    int best = 1;
    for (int i = 1 + lane; i < Nx; i += 64)
        if (row[i] > pin) best = i;
    for (int o = 32; o >= 1; o >>= 1) best = max(best, __shfl_xor(best, o, 64));
    if (lane == 0) rowmax[j] = best;
}

__global__ __launch_bounds__(256) void max_int_f32_kernel(const int* v, int n, int* out) {
    __shared__ int red[256];
    int best = 1;
    for (int q = threadIdx.x; q < n; q += blockDim.x) best = max(best, v[q]);
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

hipError_t launch_tip_column_f32(const float* psi, int Nx, int Ny, float pin, int* rowbuf,
                                 int* out, hipStream_t s) {
    hipLaunchKernelGGL(tip_rows_f32_kernel, dim3((Ny + 3) / 4), dim3(256), 0, s, psi, Nx, Ny, pin,
                       rowbuf);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(max_int_f32_kernel, dim3(1), dim3(256), 0, s, rowbuf, Ny, out);
    return hipGetLastError();
}

}  // namespace rs
