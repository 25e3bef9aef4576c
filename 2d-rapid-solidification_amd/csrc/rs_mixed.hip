// rs_mixed.hip -- mixed-precision fused step for gfx950.
//
// Same wavefront-marching layout as the FP64 stepper (62 output columns per
// wavefront, a band of rows per wavefront, 3-row register window, cross-lane
// shuffles), with:
//   * psi (hence phi) stored and updated in FP32, anisotropy in FP32 with
//     reciprocal/rsqrt arithmetic;
//   * fine T rebuilt in FP32 as T' = T - Tref from an FP32 copy of the
//     coarse grid, which itself is advanced in FP64;
//   * SGTE temperature dependence from the two-piece polynomial table (FP32
//     Horner, FP64 constant term), ln(c/(1-c)) and the chemical potential
//     assembled in FP64; cells outside the table window use the exact FP64
//     routine;
//   * conservative face fluxes in FP64: each face flux is computed once, by
//     one lane, and used with opposite signs by the two cells it separates,
//     so the interior scheme conserves solute to rounding.
// The file is compiled without floating-point contraction: every FMA is
// written out, so that the prologue of a band and the loop body of the band
// above it produce bitwise-identical face fluxes.
#include <math.h>

#include <algorithm>
#include <vector>

#include "rs_mixed.hpp"

namespace rs {

template <typename T>
__device__ __forceinline__ T shup(T v) {
    return __shfl_up(v, 1, kWave);
}
template <typename T>
__device__ __forceinline__ T shdn(T v) {
    return __shfl_down(v, 1, kWave);
}

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// fine temperature relative to Tref, FP32
struct RowT {
    const float* r0;
    const float* r1;
    float y;
};

__device__ __forceinline__ RowT row_t(const StepArgsMixed& a, int j) {
    RowT rt;
    const int jT = (int)((j - 1 + (double)a.dxT) / a.ratio);
    const double y = (j - 1.0 + (double)a.dxT) / a.ratio - jT;
    rt.r0 = a.T32 + (int64_t)jT * a.NxT_all;
    rt.r1 = rt.r0 + a.NxT_all;
    rt.y = (float)y;
    return rt;
}

__device__ __forceinline__ float temp_m(const StepArgsMixed& a, const RowT& rt, int i) {
    if (a.tfc) {
        const int g = i - 1 + a.ile;
        const int iT = g / a.ratio;
        const float x = (float)(g - a.ratio * iT) * (1.0f / 10.0f);
        const float f00 = rt.r0[iT], f10 = rt.r0[iT + 1];
        const float f01 = rt.r1[iT], f11 = rt.r1[iT + 1];
        const float a10 = f10 - f00, a01 = f01 - f00, a11 = (f11 - f10) - (f01 - f00);
        return fmaf(fmaf(a11, x, a01), rt.y, fmaf(a10, x, f00));
    }
    return (float)fma(a.fta_b, (double)i, a.fta_a);
}

template <int PIECE>
__device__ __forceinline__ void eval_poly(const PolyTable& P, float Tp, float v[4]) {
    const float t = (Tp - P.cen[PIECE]) * P.isc[PIECE];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        float acc = P.c[PIECE][f][kPolyDeg - 1];
#pragma unroll
        for (int k = kPolyDeg - 2; k >= 0; --k) acc = fmaf(acc, t, P.c[PIECE][f][k]);
        v[f] = acc * t;  // sum_{k>=1} c_k t^k
    }
}

// per-cell thermodynamics: FP32 driving force dG = f_L - f_S and FP64
// chemical potential mu = mu_L + h (mu_S - mu_L)
__device__ __forceinline__ void thermo_m(const StepArgsMixed& a, float Tp, double c, float h,
                                         float& dG, double& mu) {
    const PolyTable& P = a.poly;
    const double onemc = 1.0 - c;
    double lr = 0.0;
    if (c > 1.e-8 && onemc > 1.e-8) lr = log(c / onemc);
    if (Tp >= P.lo && Tp < P.hi) {
        float v[4];
        double c0[4];
        if (Tp < P.tb) {
            eval_poly<0>(P, Tp, v);
#pragma unroll
            for (int f = 0; f < 4; ++f) c0[f] = P.c0[0][f];
        } else {
            eval_poly<1>(P, Tp, v);
#pragma unroll
            for (int f = 0; f < 4; ++f) c0[f] = P.c0[1][f];
        }
        const float cf = (float)c;
        const float omcf = (float)onemc;
        const float om2 = 1.0f - 2.0f * cf;
        const float dGTi = (float)c0[2] + v[2];
        const float dGNb = (float)c0[3] + v[3];
        const float ex = fmaf(a.dL1, om2, a.dL0);
        dG = fmaf(cf * omcf, ex, fmaf(cf, dGTi, omcf * dGNb));
        const double om2d = 1.0 - 2.0 * c;
        const double six = 1.0 - 6.0 * c + 6.0 * (c * c);
        const double mu64 = c0[0] + a.L0l * om2d + a.L1l * six + a.RR * (P.Tref + (double)Tp) * lr;
        const float dsl = (float)(c0[1] - c0[0]) + (v[1] - v[0]) +
                          fmaf(a.dL0, -om2, -a.dL1 * (float)six);
        // mu_S - mu_L = (U_S - U_L) + (L0s - L0l)(1-2c) + (L1s - L1l)(..)
        const float mu32 = fmaf(h, dsl, v[0]);
        mu = mu64 + (double)mu32;
    } else {
        const double T = P.Tref + (double)Tp;
        const Thermo th = thermo_fp64(a.rk, T, c);
        dG = (float)th.dG;
        mu = th.muL + (double)h * (th.muS - th.muL);
    }
}

struct CellM {
    float psi, phi, dG;
    double c, S, mu;
};

__device__ __forceinline__ CellM make_cell_m(const StepArgsMixed& a, const RowT& rt, float psi,
                                             double c, int i) {
    CellM q;
    q.psi = psi;
    q.c = c;
    q.phi = fmaxf(psi, -1.0f);
    const float qv = qq<float>(a.AA, a.q_solid, psi);
    q.S = (double)qv * (c * (1.0 - c));
    const float Tp = temp_m(a, rt, i);
    const float g = gg<float>(psi);
    const float h = 0.5f * (1.0f + g);
    thermo_m(a, Tp, c, h, q.dG, q.mu);
    return q;
}

__device__ __forceinline__ double vertex4(double a0, double a1, double a2, double a3) {
    return 0.25 * (((a0 + a1) + a2) + a3);
}

// FP32 anisotropic operator (fun.cu:1030-1139) with precomputed reciprocals
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
    const float nx = fmaf(A.ops[0], ct, A.ops[1] * st);
    const float ny = fmaf(A.ops[2], ct, A.ops[3] * st);
    const float nz = fmaf(A.ops[4], ct, A.ops[5] * st);
    const float ex = fmaf(-A.ops[0], st, A.ops[1] * ct);
    const float ey = fmaf(-A.ops[2], st, A.ops[3] * ct);
    const float ez = fmaf(-A.ops[4], st, A.ops[5] * ct);
    const float nx2 = nx * nx, ny2 = ny * ny, nz2 = nz * nz;
    const float nx3 = nx2 * nx, ny3 = ny2 * ny, nz3 = nz2 * nz;
    const float np4 = fmaf(nx2, nx2, fmaf(ny2, ny2, nz2 * nz2));
    const float d1 = 4.0f * fmaf(nx3, ex, fmaf(ny3, ey, nz3 * ez));
    const float xe = nx * ex, ye = ny * ey, ze = nz * ez;
    float d2 = fmaf(3.0f, xe * xe, -nx3 * nx);
    d2 += fmaf(3.0f, ye * ye, -ny3 * ny);
    d2 += fmaf(3.0f, ze * ze, -nz3 * nz);
    d2 *= 4.0f;
    ak = fmaf(4.0f * A.epk1, np4, 1.0f - 3.0f * A.epk1);
    as = fmaf(4.0f * A.eps1, np4, 1.0f - 3.0f * A.eps1);
    const float da1 = 4.0f * A.eps1 * d1;
    const float da2 = 4.0f * A.eps1 * d2;
    return fmaf(as * as, lap,
                fmaf(2.0f * as * da1, fmaf(thx, gx, thy * gy),
                     fmaf(da2, as, da1 * da1) * fmaf(thy, gx, -thx * gy)));
}

__global__ __launch_bounds__(256) void step_mixed_kernel(StepArgsMixed a) {
    const int lane = threadIdx.x;
    const int band = blockIdx.y * blockDim.y + threadIdx.y;
    if (band >= a.nbands) return;
    const int i = blockIdx.x * kTileOut + lane;
    const int ic = i < a.Nx ? i : a.Nx - 1;
    const bool own = lane >= 1 && lane <= kTileOut && i <= a.Nx - 2;
    const int jb = 1 + band * a.band_h;
    const int je = min(jb + a.band_h - 1, a.Ny - 2);
    const int Nx = a.Nx;
    const double K = a.kdc_dx2 * a.mu_den;

    CellM m = make_cell_m(a, row_t(a, jb - 1), a.psi[(int64_t)(jb - 1) * Nx + ic],
                          a.c[(int64_t)(jb - 1) * Nx + ic], ic);
    CellM z = make_cell_m(a, row_t(a, jb), a.psi[(int64_t)jb * Nx + ic], a.c[(int64_t)jb * Nx + ic],
                          ic);
    float phL_m = shup(m.phi), phR_m = shdn(m.phi);
    float phL_z = shup(z.phi), phR_z = shdn(z.phi);
    const double SR_m = shdn(m.S);
    double SR_z = shdn(z.S);
    // vertex (i+1/2, jb-1/2) and the y-face flux (i, jb-1/2)
    double Vpm = vertex4(m.S, SR_m, z.S, SR_z);
    double Vmm = shup(Vpm);
    double Fy_prev = 0.25 * (((m.S + z.S) + Vpm) + Vmm) * (z.mu - m.mu);

    float nx_psi = a.psi[(int64_t)(jb + 1) * Nx + ic];
    double nx_c = a.c[(int64_t)(jb + 1) * Nx + ic];
    float lh_acc = 0.0f;

    for (int j = jb; j <= je; ++j) {
        const float cur_psi = nx_psi;
        const double cur_c = nx_c;
        if (j + 2 <= je + 1) {
            nx_psi = a.psi[(int64_t)(j + 2) * Nx + ic];
            nx_c = a.c[(int64_t)(j + 2) * Nx + ic];
        }
        const CellM p = make_cell_m(a, row_t(a, j + 1), cur_psi, cur_c, ic);
        const float phL_p = shup(p.phi), phR_p = shdn(p.phi);
        const double SR_p = shdn(p.S);
        const double muR_z = shdn(z.mu);
        // vertex (i+1/2, j+1/2), its left neighbour, and the faces
        const double Vpp = vertex4(z.S, SR_z, p.S, SR_p);
        const double Vmp = shup(Vpp);
        const double Fx = 0.25 * (((z.S + SR_z) + Vpp) + Vpm) * (muR_z - z.mu);  // (i+1/2, j)
        const double Fxm = shup(Fx);                                              // (i-1/2, j)
        const double Fy = 0.25 * (((z.S + p.S) + Vpp) + Vmp) * (p.mu - z.mu);    // (i, j+1/2)

        if (own) {
            float as, ak;
            const float an = aniso_f32(a, phL_m, m.phi, phR_m, phL_z, z.phi, phR_z, phL_p, p.phi,
                                       phR_p, as, ak);
            const float p00 = z.phi;
            float term = fmaf(-(p00 * p00), p00, p00) + an;
            term = fmaf(z.dG * dgg<float>(p00), a.dG_den_inv, term);
            float phn = fmaf(a.dt * ak * rcpf(as * as), term, p00);
            float psn = phn;
            if (psn > 1.0f) {
                phn -= (psn - 1.0f);
                psn = phn;
            }
            const double dc = K * ((Fx - Fxm) + (Fy - Fy_prev));
            const double cn = fmax(z.c + dc, 1.0e-6);

            const int64_t row = (int64_t)j * Nx;
            a.psi_n[row + i] = psn;
            a.c_n[row + i] = cn;
            if (i == 1) {
                a.psi_n[row] = psn;
                a.c_n[row] = cn;
            }
            if (i == Nx - 2) {
                a.psi_n[row + Nx - 1] = psn;
                a.c_n[row + Nx - 1] = cn;
            }
            if (j == a.Ny - 2 || j == 1) {
                const int64_t g = (j == 1) ? (int64_t)(a.Ny - 1) * Nx : 0;
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
                lh_acc += psn - z.psi;
                if (j % 10 == 0) {
                    a.lh[(int64_t)(j / 10 - 1) * Nx + i] = (double)lh_acc;
                    lh_acc = 0.0f;
                }
            }
        }
        m = z;
        phL_m = phL_z;
        phR_m = phR_z;
        z = p;
        phL_z = phL_p;
        phR_z = phR_p;
        SR_z = SR_p;
        Vpm = Vpp;
        Fy_prev = Fy;
    }
}

hipError_t launch_step_mixed(const StepArgsMixed& a, int wpb, hipStream_t s) {
    if (wpb < 1) wpb = 1;
    if (wpb > 4) wpb = 4;
    dim3 block(kWave, wpb);
    dim3 grid((a.Nx - 2 + kTileOut - 1) / kTileOut, (a.nbands + wpb - 1) / wpb);
    hipLaunchKernelGGL(step_mixed_kernel, grid, block, 0, s, a);
    return hipGetLastError();
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

__global__ void f32_to_f64_kernel(const float* a, double* b, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) b[q] = (double)a[q];
}
__global__ void f64_to_f32_kernel(const double* a, float* b, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) b[q] = (float)a[q];
}
hipError_t launch_f32_to_f64(const float* a, double* b, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(f32_to_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b,
                       n);
    return hipGetLastError();
}
hipError_t launch_f64_to_f32(const double* a, float* b, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b,
                       n);
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

// ---------------------------------------------------------------------------
// host: least-squares fit of the SGTE temperature dependence
// ---------------------------------------------------------------------------
namespace {

void unary_funcs(double T, double f[4]) {
    const Unaries u = sgte_unaries(T, log(T));
    f[0] = u.ti_liq - u.nb_liq;
    f[1] = u.ti_bcc - u.nb_bcc;
    f[2] = u.ti_liq - u.ti_bcc;
    f[3] = u.nb_liq - u.nb_bcc;
}

// fit degree kPolyDeg in t = (T - mid) / half on [a, b] (long double normal
// equations on Chebyshev nodes)
void fit_piece(double a, double b, double Tref, PolyTable* P, int piece) {
    const int D = kPolyDeg;
    const double mid = 0.5 * (a + b), half = 0.5 * (b - a);
    P->cen[piece] = (float)(mid - Tref);
    P->isc[piece] = half > 0 ? (float)(1.0 / half) : 0.0f;
    const double hw = half > 0 ? 1.0 / (double)P->isc[piece] : 0.0;  // the scale the kernel uses
    if (half <= 0) {
        for (int f = 0; f < 4; ++f) {
            P->c0[piece][f] = 0.0;
            for (int k = 0; k < D; ++k) P->c[piece][f][k] = 0.0f;
        }
        return;
    }
    const int N = 64;
    long double A[kPolyDeg + 1][kPolyDeg + 1] = {};
    long double B[4][kPolyDeg + 1] = {};
    for (int q = 0; q < N; ++q) {
        const double tq = cos(M_PI * (q + 0.5) / N);
        // the kernel forms t from T' and cen in FP32; fit against that t
        const double T = mid + tq * hw;
        double f[4];
        unary_funcs(T, f);
        long double tp[2 * kPolyDeg + 1];
        tp[0] = 1.0L;
        for (int k = 1; k <= 2 * D; ++k) tp[k] = tp[k - 1] * (long double)tq;
        for (int r = 0; r <= D; ++r) {
            for (int cc = 0; cc <= D; ++cc) A[r][cc] += tp[r + cc];
            for (int g = 0; g < 4; ++g) B[g][r] += tp[r] * (long double)f[g];
        }
    }
    // Gaussian elimination (SPD, no pivoting needed), four right-hand sides
    for (int k = 0; k <= D; ++k) {
        for (int r = k + 1; r <= D; ++r) {
            const long double fct = A[r][k] / A[k][k];
            for (int cc = k; cc <= D; ++cc) A[r][cc] -= fct * A[k][cc];
            for (int g = 0; g < 4; ++g) B[g][r] -= fct * B[g][k];
        }
    }
    for (int g = 0; g < 4; ++g) {
        long double x[kPolyDeg + 1];
        for (int r = D; r >= 0; --r) {
            long double acc = B[g][r];
            for (int cc = r + 1; cc <= D; ++cc) acc -= A[r][cc] * x[cc];
            x[r] = acc / A[r][r];
        }
        P->c0[piece][g] = (double)x[0];
        for (int k = 1; k <= D; ++k) P->c[piece][g][k - 1] = (float)x[k];
    }
}

float eval_host(const PolyTable& P, int piece, int f, float Tp) {
    const float t = (Tp - P.cen[piece]) * P.isc[piece];
    float acc = P.c[piece][f][kPolyDeg - 1];
    for (int k = kPolyDeg - 2; k >= 0; --k) acc = fmaf(acc, t, P.c[piece][f][k]);
    return acc * t;
}

}  // namespace

double fit_poly_table(const rs_params& p, double T_lo, double T_hi, double Tref, PolyTable* P) {
    (void)p;
    const double TB = 1941.0;  // G_Ti breakpoint (fun.cu:1629, 1639)
    // stay inside the SGTE branches that do not change inside the window:
    // G_Ti_Liq low branch starts at 1300 K, G_Nb changes at 2750 K
    T_lo = std::max(T_lo, 1300.0);
    T_hi = std::min(T_hi, 2749.0);
    if (T_hi <= T_lo) T_hi = T_lo;
    P->Tref = Tref;
    P->lo = (float)(T_lo - Tref);
    P->hi = (float)(T_hi - Tref);
    const double a0 = T_lo, b0 = std::min(std::max(TB, T_lo), T_hi);
    const double a1 = b0, b1 = T_hi;
    P->tb = (float)(b0 - Tref);
    fit_piece(a0, b0, Tref, P, 0);
    fit_piece(a1, b1, Tref, P, 1);
    // check on a dense grid exactly as the kernel evaluates (FP32 T')
    double worst = 0.0;
    for (int q = 0; q <= 4000; ++q) {
        const float Tp = P->lo + (P->hi - P->lo) * (float)q / 4000.0f;
        if (!(Tp >= P->lo && Tp < P->hi)) continue;
        const int piece = Tp < P->tb ? 0 : 1;
        double f[4];
        unary_funcs(Tref + (double)Tp, f);
        for (int g = 0; g < 4; ++g) {
            const double v = P->c0[piece][g] + (double)eval_host(*P, piece, g, Tp);
            worst = std::max(worst, fabs(v - f[g]));
        }
    }
    return worst;
}

}  // namespace rs
