// rs_antitrap.hpp -- the antitrapping / solute-trapping correction of c
// (an extension absent from the reference): face flux, the row window and
// the band routine shared by the single-run pass (rs_kernels.hip) and the
// batched one (rs_mixed.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "rs_kernels.hpp"

namespace rs {

// Antitrapping current with the solute-trapping factor (an extension the
// reference does not have): face flux between cells P and Q,
//   J = a_t (1-k) c_L dphidt (-g_n/|g|) / (1 + |V|/V_D),
// g_n = (phi_Q - phi_P)/dx, g_t from the four cells beside the face,
// dphidt the P/Q mean of (phi_next - phi)/dt, V = dphidt/|g|,
// c_L = c_f/(1 - h_f (1-k)) with face means c_f and h_f = (1 + phi_f)/2.
// Arithmetic: reciprocals of dx, 4 dx, dt and V_D precomputed, 1/|g| once,
// and c_L and the trapping factor under one division: on the FP64 path two
// IEEE divisions and one square root per face (the CPU restatement uses the
// same form), on the FP32 path the hardware reciprocal and reciprocal
// square root.
struct AtArgs {
    double idx, i4dx, idt, ivdn;  // 1/dx, 1/(4 dx), 1/dt, 1/V_D (0: no trapping)
    double omk, coef;             // 1 - k, a_t (1 - k)
    double dt;
};

// 1/sqrt and division: IEEE in FP64; in FP32 the hardware reciprocal square
// root and reciprocal (~1 ulp), the flux being a small correction
__device__ __forceinline__ double at_rsqrt(double x) { return 1.0 / sqrt(x); }
__device__ __forceinline__ float at_rsqrt(float x) { return __frsqrt_rn(x); }
__device__ __forceinline__ double at_div(double a, double b) { return a / b; }
__device__ __forceinline__ float at_div(float a, float b) { return a * __frcp_rn(b); }

// R: arithmetic of the flux (FP64 on the FP64 path, FP32 on the mixed path)
template <typename R>
struct AtConst {
    R idx, i4dx, idt, ivdn, omk, coef;
};
template <typename R>
__device__ __forceinline__ AtConst<R> at_const(const AtArgs& a) {
    return AtConst<R>{(R)a.idx, (R)a.i4dx, (R)a.idt, (R)a.ivdn, (R)a.omk, (R)a.coef};
}

template <typename R>
__device__ __forceinline__ R at_flux(const AtConst<R>& a, R pP, R pQ, R nP, R nQ, R cP, R cQ, R t1,
                                     R t2, R t3, R t4) {
    const R gn = (pQ - pP) * a.idx;
    const R gt = ((t1 + t2) - (t3 + t4)) * a.i4dx;
    const R g2 = gn * gn + gt * gt;
    if (!(g2 > R(1e-20))) return R(0);
    const R ig = at_rsqrt(g2);
    const R dpt = R(0.5) * ((nP - pP) + (nQ - pQ)) * a.idt;
    const R cf = R(0.5) * (cP + cQ);
    const R hf = R(0.5) * (R(1) + R(0.5) * (pP + pQ));
    R den = R(1) - hf * a.omk;
    if (a.ivdn > R(0)) den = den * (R(1) + fabs(dpt * ig) * a.ivdn);
    return a.coef * at_div(cf, den) * dpt * (-gn * ig);
}

namespace at {
// lane l <- lane l-1 / lane l+1 (the lane without a source gets 0: a halo
// lane whose value is not used)
__device__ __forceinline__ int at_up(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ int at_dn(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ float at_up(float v) { return __int_as_float(at_up(__float_as_int(v))); }
__device__ __forceinline__ float at_dn(float v) { return __int_as_float(at_dn(__float_as_int(v))); }
__device__ __forceinline__ double at_up(double v) {
    return __hiloint2double(at_up(__double2hiint(v)), at_up(__double2loint(v)));
}
__device__ __forceinline__ double at_dn(double v) {
    return __hiloint2double(at_dn(__double2hiint(v)), at_dn(__double2loint(v)));
}

// one row of the window: phi (= max(psi, -1)) with its i-1 / i+1
// neighbours, phi_next and c with their i+1 neighbours, and the row's
// uncorrected c_next
template <typename R>
struct AtRow {
    R f, fl, fr, g, gr, c, cr;
    double cn;
};
// the raw loads of one row, issued two rows ahead of the row's update (c_next
// included: the update then waits on no load)
template <typename T>
struct AtRaw {
    T f, g;
    double c, cn;
};
template <typename T>
__device__ __forceinline__ AtRaw<T> at_load(const T* psi, const T* psi_n, const double* c,
                                            const double* c_n, int Nx, int j, int ic) {
    const int64_t q = (int64_t)j * Nx + ic;
    return AtRaw<T>{psi[q], psi_n[q], c[q], c_n[q]};
}
template <typename R, typename T>
__device__ __forceinline__ AtRow<R> at_row(const AtRaw<T>& w) {
    AtRow<R> r;
    r.f = fmax((R)w.f, R(-1));
    r.g = fmax((R)w.g, R(-1));
    r.c = (R)w.c;
    r.cn = w.cn;
    r.fl = at_up(r.f);
    r.fr = at_dn(r.f);
    r.gr = at_dn(r.g);
    r.cr = at_dn(r.c);
    return r;
}
// y face between rows Z (below) and P (above) at the lane's column
template <typename R>
__device__ __forceinline__ R at_yface(const AtConst<R>& a, const AtRow<R>& Z, const AtRow<R>& P) {
    return at_flux(a, Z.f, P.f, Z.g, P.g, Z.c, P.c, Z.fr, P.fr, Z.fl, P.fl);
}
}  // namespace at

// One wavefront's band of the correction: column tile `tile` (62 output
// columns, lanes 1..62), rows jb..je, marching down with a 3-row window.
// Every face flux is evaluated once -- the x face (i+1/2) by lane i, its
// i-1/2 face taken from lane i-1, the y face (j+1/2) carried to row j+1 --
// and the divergence is formed in FP64, so the correction conserves solute
// exactly; a band's first y face is the last one of the band below,
// recomputed from the same operands, so any band split gives the same
// bits. The flux arithmetic follows the field precision T: FP64 with FP64
// psi, FP32 with FP32 psi. c_n holds the step's new c and is corrected in
// place (clamp at 1e-6, ghost images by the owning cell); an interior cell's
// c_n is read and written only by the lane that owns it, so the window's
// early loads of the rows below read uncorrected values (halo lanes load
// values they never use).
template <typename T>
__device__ __forceinline__ void antitrap_band(const T* psi, const T* psi_n, const double* c,
                                              double* c_n, int Nx, int Ny, const AtArgs& a,
                                              int tile, int jb, int je) {
    using namespace at;
    typedef T R;
    const AtConst<R> k = at_const<R>(a);
    const int lane = threadIdx.x;
    const int i = tile * kTileOut + lane;
    const int ic = i < Nx ? i : Nx - 1;
    const bool own = lane >= 1 && lane <= kTileOut && i <= Nx - 2;
    AtRow<R> M = at_row<R>(at_load(psi, psi_n, c, c_n, Nx, jb - 1, ic));
    AtRow<R> Z = at_row<R>(at_load(psi, psi_n, c, c_n, Nx, jb, ic));
    AtRaw<T> nx = at_load(psi, psi_n, c, c_n, Nx, jb + 1, ic);
    double Jy_prev = (double)at_yface(k, M, Z);
    for (int j = jb; j <= je; ++j) {
        const AtRow<R> P = at_row<R>(nx);
        // unconditional (the last row reloads row je+1, whose values are not
        // used), so the compiler need not wait on the stores before it
        nx = at_load(psi, psi_n, c, c_n, Nx, min(j + 2, je + 1), ic);
        const double Jx = (double)at_flux(k, Z.f, Z.fr, Z.g, Z.gr, Z.c, Z.cr, P.f, P.fr, M.f, M.fr);
        const double Jxm = at_up(Jx);
        const double Jy = (double)at_yface(k, Z, P);
        if (own) {
            const double dc = -a.dt * ((Jx - Jxm) + (Jy - Jy_prev)) * a.idx;
            const int64_t row = (int64_t)j * Nx;
            const double v = fmax(Z.cn + dc, 1.0e-6);
            c_n[row + i] = v;
            if (i == 1) c_n[row] = v;
            if (i == Nx - 2) c_n[row + Nx - 1] = v;
            if (j == 1 || j == Ny - 2) {
                const int64_t g = (j == 1) ? (int64_t)(Ny - 1) * Nx : 0;
                c_n[g + i] = v;
                if (i == 1) c_n[g] = v;
                if (i == Nx - 2) c_n[g + Nx - 1] = v;
            }
        }
        Jy_prev = Jy;
        M = Z;
        Z = P;
    }
}

// flux constants of a run: dx, dt, a_t, k and V_D in W/tau0 units (0: no
// trapping factor)
inline AtArgs at_args(double dx, double dt, double at, double k, double vdn) {
    AtArgs a;
    a.idx = 1.0 / dx;
    a.i4dx = 1.0 / (4.0 * dx);
    a.idt = 1.0 / dt;
    a.ivdn = vdn > 0.0 ? 1.0 / vdn : 0.0;
    a.omk = 1.0 - k;
    a.coef = at * (1.0 - k);
    a.dt = dt;
    return a;
}

}  // namespace rs
