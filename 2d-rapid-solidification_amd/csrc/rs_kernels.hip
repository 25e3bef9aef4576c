// rs_kernels.hip -- FP64 HIP kernels of the time-stepper for gfx950.
//
// The reference launches five fine-grid kernels per step (compute_phi,
// compute_psi, compute_c_CALPHAD, interpolate_temperature, integrate_phi;
// fun.cu:466-495), each a full HBM round trip, and re-evaluates the CALPHAD
// chemical potential five times per cell. Here one kernel does the whole
// fine-grid update:
//   * persistent state is psi and c only (phi = max(psi,-1) exactly, since
//     compute_psi clamps phi but not psi, fun.cu:1243-1267); fine T is
//     rebuilt on the fly from the coarse grid (or the frozen gradient);
//   * a wavefront owns a 62-column strip and marches down a band of rows,
//     keeping a 3-row window of phi, S = q(psi)c(1-c) and mu in registers;
//     the i+-1 neighbours come from cross-lane shuffles, so every cell's
//     CALPHAD evaluation (3 table logs) is done once and shared with its
//     neighbours;
//   * ghost cells are written by the lane that owns their source cell, so the
//     boundary copies of compute_boundary_individual (fun.cu:1220) are
//     race-free;
//   * latent heat (integrate_phi, fun.cu:1494) is accumulated per column
//     over each 10-row block band and written as one FP64 value per column
//     and band; the coarse T kernel finishes the 10x10 block sums.
#include <hip/hip_ext.h>

#include "rs_kernels.hpp"
#include "rs_antitrap.hpp"
#include "rs_sgte_poly.inc"

namespace rs {

// lane l <- lane l-1 / lane l+1 through DPP wavefront shifts (wave_shr:1 /
// wave_shl:1), one VALU move per 32-bit half instead of an LDS permute; the
// lane with no source gets 0, a halo lane whose result is not stored
__device__ __forceinline__ int dpp_up1(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ int dpp_dn1(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ double shfl_up1(double v) {
    return __hiloint2double(dpp_up1(__double2hiint(v)), dpp_up1(__double2loint(v)));
}
__device__ __forceinline__ double shfl_dn1(double v) {
    return __hiloint2double(dpp_dn1(__double2hiint(v)), dpp_dn1(__double2loint(v)));
}

// fine temperature at (i, j) as the reference's T_fine / get_delta give it
__device__ __forceinline__ double temp_at(const TempSrc& t, int i, int j) {
    if (t.tfc) {
        const int r = t.ratio;
        const int iT = (i - 1 + t.ile) / r;
        const int jT = (int)((j - 1 + t.dxT) / r);
        const double x = (i - 1.0 + t.ile) / r - iT;
        const double y = (j - 1.0 + t.dxT) / r - jT;
        const double* row0 = t.T + (int64_t)jT * t.NxT_all;
        const double* row1 = row0 + t.NxT_all;
        const double f00 = row0[iT], f10 = row0[iT + 1];
        const double f01 = row1[iT], f11 = row1[iT + 1];
        const double a10 = f10 - f00, a01 = f01 - f00, a11 = f11 - f10 - f01 + f00;
        return f00 + a10 * x + a01 * y + a11 * x * y;
    }
    const double delta = t.d0 + t.GG * (i + t.fta_ioff) * t.dx * t.S * t.W0 / t.dT0 - t.dl;
    return delta * t.dT0 + t.Tmelt;
}

struct Cell {
    double psi, c, phi, S, mu, dG;
};

__device__ __forceinline__ Cell make_cell(const StepArgsF64& a, const double* lntab, double psi,
                                          double c, int i, int j) {
    Cell q;
    q.psi = psi;
    q.c = c;
    q.phi = fmax(psi, -1.0);
    q.S = qq<double>(a.AA, a.q_solid, psi) * c * (1.0 - c);
    const double T = temp_at(a.ts, i, j);
    const Thermo th = thermo_fp64<true>(a.rk, T, c, TabLog{lntab});
    const double g = gg<double>(psi);
    q.mu = 0.5 * (1.0 + g) * th.muS + 0.5 * (1.0 - g) * th.muL;
    q.dG = th.dG;
    return q;
}

// three wavefronts per SIMD: with every division and log as the reference
// writes them, 537 us for a 4096^2 step on a developed front, against 598 us
// at two (181 VGPRs) and 720 us at four (128 VGPRs, 50 registers spilled)
__global__ __launch_bounds__(256, 3) void step_fp64_kernel(StepArgsF64 a) {
    // the log table of the CALPHAD evaluation (ln_tab), staged in LDS
    __shared__ double lntab[2 * RS_LN_TAB_N];
    for (int q = threadIdx.y * blockDim.x + threadIdx.x; q < 2 * RS_LN_TAB_N;
         q += blockDim.x * blockDim.y)
        lntab[q] = kLnTab[q];
    __syncthreads();
    const int lane = threadIdx.x;
    // logical workgroup (column tile, band group) in XCD order
    const int ntx = gridDim.x;
    const int g = xcd_order(blockIdx.y * ntx + blockIdx.x, ntx * gridDim.y);
    const int band = (g / ntx) * blockDim.y + threadIdx.y;
    if (band >= a.nbands) return;  // whole wavefront leaves together
    const int i = (g % ntx) * kTileOut + lane;
    const int ic = i < a.Nx ? i : a.Nx - 1;  // loads of lanes past the grid
    const bool own = lane >= 1 && lane <= kTileOut && i <= a.Nx - 2;
    const int jb = 1 + band * a.band_h;
    const int je = min(jb + a.band_h - 1, a.Ny - 2);
    const int Nx = a.Nx;
    // the solute ledger's x faces: does this wavefront own column 1 or Nx-2
    const bool led_wave = a.led && __ballot(own && (i == 1 || i == Nx - 2)) != 0;
    // divisions by the run's constants as products with their reciprocals
    const double idG = 1.0 / a.dG_den, idx2 = 1.0 / a.dx2;

    // rows jb-1 and jb
    Cell m = make_cell(a, lntab, a.psi[(int64_t)(jb - 1) * Nx + ic],
                       a.c[(int64_t)(jb - 1) * Nx + ic], ic, jb - 1);
    Cell z = make_cell(a, lntab, a.psi[(int64_t)jb * Nx + ic], a.c[(int64_t)jb * Nx + ic], ic, jb);
    double phL_m = shfl_up1(m.phi), phR_m = shfl_dn1(m.phi);
    double SL_m = shfl_up1(m.S), SR_m = shfl_dn1(m.S);
    double phL_z = shfl_up1(z.phi), phR_z = shfl_dn1(z.phi);
    double SL_z = shfl_up1(z.S), SR_z = shfl_dn1(z.S);

    // prefetch of row jb+1
    double nx_psi = a.psi[(int64_t)(jb + 1) * Nx + ic];
    double nx_c = a.c[(int64_t)(jb + 1) * Nx + ic];
    double lh_acc = 0.0;

    for (int j = jb; j <= je; ++j) {
        const double cur_psi = nx_psi, cur_c = nx_c;
        {  // unconditional, the last row reloading row je+1 (no wait on the stores)
            const int jn = min(j + 2, je + 1);
            nx_psi = a.psi[(int64_t)jn * Nx + ic];
            nx_c = a.c[(int64_t)jn * Nx + ic];
        }
        const Cell p = make_cell(a, lntab, cur_psi, cur_c, ic, j + 1);
        const double phL_p = shfl_up1(p.phi), phR_p = shfl_dn1(p.phi);
        const double SL_p = shfl_up1(p.S), SR_p = shfl_dn1(p.S);
        const double muL = shfl_up1(z.mu), muR = shfl_dn1(z.mu);

        if (own) {
            // ---- phase field (compute_phi, fun.cu:1143-1197) ----
            double as, ak;
            const double an = anisotropy<double>(a.an, phL_m, m.phi, phR_m, phL_z, z.phi, phR_z,
                                                 phL_p, p.phi, phR_p, as, ak);
            const double p00 = z.phi;
            double term = p00 - (p00 * p00) * p00 + an;
            term += z.dG * dgg<double>(p00) * idG;
            double phn = p00 + a.dt * ak * rcp_nr(as * as) * term;  // tau = 1 (one orientation)
            // ---- psi <= 1 fix (compute_psi, fun.cu:1243-1261) ----
            double psn = phn;
            if (psn > 1.0) {
                phn -= (psn - 1.0);
                psn = phn;
            }
            // ---- solute (compute_c_CALPHAD, fun.cu:1422-1464) ----
            const double S00 = z.S, S10 = SR_z, S01 = p.S, Sm0 = SL_z, S0m = m.S;
            const double S11 = SR_p, Sm1 = SL_p, S1m = SR_m, Smm = SL_m;
            const double Shh = 0.25 * (S11 + S10 + S01 + S00);
            const double Shn = 0.25 * (S10 + S00 + S1m + S0m);
            const double Snh = 0.25 * (Sm1 + S01 + Sm0 + S00);
            const double Snn = 0.25 * (Sm0 + Smm + S0m + S00);
            const double f00 = z.mu;
            const double F10 = 0.25 * (S10 + S00 + Shh + Shn) * (muR - f00) * a.mu_den;
            const double F01 = 0.25 * (S01 + S00 + Shh + Snh) * (p.mu - f00) * a.mu_den;
            const double Fm0 = 0.25 * (S00 + Sm0 + Snh + Snn) * (muL - f00) * a.mu_den;
            const double F0m = 0.25 * (S00 + S0m + Shn + Snn) * (m.mu - f00) * a.mu_den;
            const double dc = a.kdc * (F10 + F01 + Fm0 + F0m) * idx2;
            const double cn = fmax(z.c + dc, 1.0e-6);
            // ledger: the flux through the boundary faces (ghost side); the
            // test is wavefront-uniform, taken by the edge tiles and rows only
            if (led_wave || (a.led && (j == 1 || j == a.Ny - 2))) {
                if (i == 1) a.led[j] += a.kdc * Fm0 * idx2;
                if (i == Nx - 2) a.led[a.Ny + j] += a.kdc * F10 * idx2;
                if (j == 1) a.led[2 * a.Ny + i] += a.kdc * F0m * idx2;
                if (j == a.Ny - 2) a.led[2 * a.Ny + Nx + i] += a.kdc * F01 * idx2;
            }

            // ---- stores + ghost images ----
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
            // ---- latent heat (integrate_phi, fun.cu:1526-1531) ----
            if (a.lh_on && j <= a.lh_last_row) {
                lh_acc += psn - z.psi;
                if (j % 10 == 0) {
                    a.lh[(int64_t)(j / 10 - 1) * Nx + i] = lh_acc;
                    lh_acc = 0.0;
                }
            }
        }
        // slide the window
        m = z;
        phL_m = phL_z; phR_m = phR_z; SL_m = SL_z; SR_m = SR_z;
        z = p;
        phL_z = phL_p; phR_z = phR_p; SL_z = SL_p; SR_z = SR_p;
    }
}

hipError_t launch_step_fp64(const StepArgsF64& a, int wpb, hipStream_t s, LaunchEv ev) {
    if (wpb < 1) wpb = 1;
    if (wpb > 4) wpb = 4;
    dim3 block(kWave, wpb);
    dim3 grid((a.Nx - 2 + kTileOut - 1) / kTileOut, (a.nbands + wpb - 1) / wpb);
    hipExtLaunchKernelGGL(step_fp64_kernel, grid, block, 0, s, ev.start, ev.stop, 0, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// coarse temperature: integrate_phi's block sums finished from the column
// sums, then compute_temperature (fun.cu:1537-1578) with race-free
// boundaries (the thread owning a source cell writes its periodic image).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void temperature_kernel(TempArgs a) {
    const int iT = blockIdx.x * blockDim.x + threadIdx.x;
    const int jT = blockIdx.y * blockDim.y + threadIdx.y;
    if (iT >= a.NxT || jT >= a.NyT) return;
    temperature_cell(a, iT, jT);
}

hipError_t launch_temperature(const TempArgs& a, hipStream_t s, LaunchEv ev) {
    dim3 block(64, 4);
    dim3 grid((a.NxT + 63) / 64, (a.NyT + 3) / 4);
    hipExtLaunchKernelGGL(temperature_kernel, grid, block, 0, s, ev.start, ev.stop, 0, a);
    return hipGetLastError();
}

// initial_temperature (fun.cu:1023-1028) on the padded grid, both buffers
__global__ void init_T_kernel(double* T0, double* T1, int NxT_all, int NyT_all, double Tcold,
                              int S, double W0, double dx, int r, double GG) {
    const int iT = blockIdx.x * blockDim.x + threadIdx.x;
    const int jT = blockIdx.y * blockDim.y + threadIdx.y;
    if (iT >= NxT_all || jT >= NyT_all) return;
    const double v = Tcold + (iT - 1) * S * W0 * dx * r * GG;
    T0[(int64_t)jT * NxT_all + iT] = v;
    T1[(int64_t)jT * NxT_all + iT] = v;
}

hipError_t launch_init_T(double* T0, double* T1, int NxT_all, int NyT_all, double Tcold, int S,
                         double W0, double dx, int r, double GG, hipStream_t s) {
    dim3 block(64, 4);
    dim3 grid((NxT_all + 63) / 64, (NyT_all + 3) / 4);
    hipLaunchKernelGGL(init_T_kernel, grid, block, 0, s, T0, T1, NxT_all, NyT_all, Tcold, S, W0,
                       dx, r, GG);
    return hipGetLastError();
}

// initial (fun.cu:983-1012) [mode 0] or the seeded planar front [mode 1];
// x0[j] is the front position of row j
__global__ void init_fields_kernel(double* psi, double* c, const double* x0, int Nx, int Ny,
                                   double dx, int mode, double cinf, double c_liq, double c_sol) {
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
    psi[(int64_t)j * Nx + i] = 0.0 + ph;
    c[(int64_t)j * Nx + i] = cv;
}

hipError_t launch_init_fields(double* psi, double* c, const double* x0, int Nx, int Ny,
                              double dx, int mode, double cinf, double c_liq, double c_sol,
                              hipStream_t s) {
    dim3 block(64, 4);
    dim3 grid((Nx + 63) / 64, (Ny + 3) / 4);
    hipLaunchKernelGGL(init_fields_kernel, grid, block, 0, s, psi, c, x0, Nx, Ny, dx, mode, cinf,
                       c_liq, c_sol);
    return hipGetLastError();
}

// pull_back (fun.cu:1469-1492)
__global__ void pull_back_kernel(const double* psi, const double* c, double* psi_n, double* c_n,
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

hipError_t launch_pull_back(const double* psi, const double* c, double* psi_n, double* c_n,
                            int Nx, int Ny, int off, double cinf, hipStream_t s) {
    dim3 block(64, 4);
    dim3 grid((Nx + 63) / 64, (Ny + 3) / 4);
    hipLaunchKernelGGL(pull_back_kernel, grid, block, 0, s, psi, c, psi_n, c_n, Nx, Ny, off,
                       cinf);
    return hipGetLastError();
}

// get_itip_all_grain (fun.cu:1745-1754): the reference's running scan
// reduces to max(1, max{i >= 1 : psi(i,j) > phi_interface}) over all rows.
// One wavefront per row, then one block over the row maxima.
__global__ __launch_bounds__(256) void tip_rows_kernel(const double* psi, int Nx, int Ny,
                                                       double pin, int* rowmax) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (j >= Ny) return;
    const double* row = psi + (int64_t)j * Nx;
    int best = 1;
    for (int i = 1 + lane; i < Nx; i += 64)
        if (row[i] > pin) best = i;
    for (int o = 32; o >= 1; o >>= 1) best = max(best, __shfl_xor(best, o, 64));
    if (lane == 0) rowmax[j] = best;
}

__global__ __launch_bounds__(256) void max_int_kernel(const int* v, int n, int* out) {
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

hipError_t launch_tip_column(const double* psi, int Nx, int Ny, double pin, int* rowbuf,
                             int* out, hipStream_t s) {
    hipLaunchKernelGGL(tip_rows_kernel, dim3((Ny + 3) / 4), dim3(256), 0, s, psi, Nx, Ny, pin,
                       rowbuf);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(max_int_kernel, dim3(1), dim3(256), 0, s, rowbuf, Ny, out);
    return hipGetLastError();
}

// FP64 solute mass over the interior: per-row sums (one wavefront per row,
// fixed shuffle tree), then a single block sums the rows in a fixed order.
__global__ __launch_bounds__(256) void mass_rows_kernel(const double* c, int Nx, int Ny,
                                                        double* rowsum) {
    const int lane = threadIdx.x & 63;
    const int j = 1 + blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (j > Ny - 2) return;
    const double* row = c + (int64_t)j * Nx;
    double acc = 0.0;
    for (int i = 1 + lane; i <= Nx - 2; i += 64) acc += row[i];
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) rowsum[j] = acc;
}

__global__ __launch_bounds__(256) void sum_rows_kernel(const double* v, int lo, int hi,
                                                       double* out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int q = lo + threadIdx.x; q <= hi; q += blockDim.x) acc += v[q];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

hipError_t launch_mass(const double* c, int Nx, int Ny, double* rowbuf, double* out,
                       hipStream_t s) {
    hipLaunchKernelGGL(mass_rows_kernel, dim3((Ny - 2 + 3) / 4), dim3(256), 0, s, c, Nx, Ny,
                       rowbuf);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sum_rows_kernel, dim3(1), dim3(256), 0, s, rowbuf, 1, Ny - 2, out);
    return hipGetLastError();
}

hipError_t launch_sum(const double* v, int lo, int hi, double* out, hipStream_t s) {
    hipLaunchKernelGGL(sum_rows_kernel, dim3(1), dim3(256), 0, s, v, lo, hi, out);
    return hipGetLastError();
}

__global__ void ledger_pull_kernel(const double* m, double* acc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) acc[0] += m[1] - m[0];
}

hipError_t launch_ledger_pull(const double* m, double* acc, hipStream_t s) {
    hipLaunchKernelGGL(ledger_pull_kernel, dim3(1), dim3(64), 0, s, m, acc);
    return hipGetLastError();
}

// T_fine as the stepper sees it (interpolate_temperature, fun.cu:1580-1608)
__global__ void interp_kernel(TempSrc ts, double* Tf, int Nx, int Ny) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= Nx || j >= Ny) return;
    Tf[(int64_t)j * Nx + i] = temp_at(ts, i, j);
}

hipError_t launch_interp(const TempSrc& ts, double* Tf, int Nx, int Ny, hipStream_t s) {
    dim3 block(64, 4);
    dim3 grid((Nx + 63) / 64, (Ny + 3) / 4);
    hipLaunchKernelGGL(interp_kernel, grid, block, 0, s, ts, Tf, Nx, Ny);
    return hipGetLastError();
}

// plain streaming copy (profiler byte-counter calibration)
__global__ __launch_bounds__(256) void copy_kernel(const double* __restrict__ a,
                                                   double* __restrict__ b, int64_t n) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; q < n; q += stride) b[q] = a[q];
}

hipError_t launch_copy(const double* a, double* b, int64_t n, hipStream_t s) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, b, n);
    return hipGetLastError();
}

TempSrc make_tempsrc(const rs_params& p, const rs_constants& k, const double* Tc, int ile) {
    TempSrc t;
    t.T = Tc;
    t.tfc = p.ifTFC;
    t.ratio = p.ratio_dxT_dx;
    t.NxT_all = k.NxT_all;
    t.ile = ile;
    t.dxT = p.dx * p.ratio_dxT_dx;
    t.fta_ioff = -p.i_tip_target + k.i_offset_history - k.i_offset_history_initial;
    t.S = p.S;
    t.d0 = k.delta_initial;
    t.GG = p.GG;
    t.dx = p.dx;
    t.W0 = p.W0;
    t.dT0 = p.deltaT0;
    t.dl = p.GG * (k.l_offset_Vp - k.l_offset_Vp_initial) * p.S * p.W0 / p.deltaT0;
    t.Tmelt = p.Tmelt;
    return t;
}


}  // namespace rs

namespace rs {

// finite-value sentinel: count and first index of NaN / Inf values (the
// count and the minimum do not depend on the order of the atomics)
template <typename T>
__global__ __launch_bounds__(256) void nonfinite_kernel(const T* f, int64_t n,
                                                        unsigned long long* out) {
    unsigned long long cnt = 0, first = ~0ull;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        if (!isfinite(f[q])) {
            ++cnt;
            if ((unsigned long long)q < first) first = (unsigned long long)q;
        }
    }
    if (cnt) {
        atomicAdd(&out[0], cnt);
        atomicMin(&out[1], first);
    }
}

template <typename T>
static hipError_t launch_nf(const T* f, int64_t n, unsigned long long* out, hipStream_t s) {
    const int64_t want = (n + 255) / 256;
    const int blocks = (int)(want < 2048 ? (want > 0 ? want : 1) : 2048);
    hipLaunchKernelGGL(nonfinite_kernel<T>, dim3(blocks), dim3(256), 0, s, f, n, out);
    return hipGetLastError();
}
hipError_t launch_nonfinite(const double* f, int64_t n, unsigned long long* out, hipStream_t s) {
    return launch_nf(f, n, out, s);
}
hipError_t launch_nonfinite_f32(const float* f, int64_t n, unsigned long long* out, hipStream_t s) {
    return launch_nf(f, n, out, s);
}

// antitrapping pass of one run (the band routine is in rs_antitrap.hpp)
constexpr int kAtBand = 64;

template <typename T>
__global__ __launch_bounds__(256) void antitrap_kernel(const T* psi, const T* psi_n, const double* c,
                                                       double* c_n, int Nx, int Ny, AtArgs a) {
    // logical workgroup (column tile, band group) in XCD order
    const int ntx = gridDim.x;
    const int g = xcd_order(blockIdx.y * ntx + blockIdx.x, ntx * gridDim.y);
    const int band = (g / ntx) * blockDim.y + threadIdx.y;
    const int jb = 1 + band * kAtBand;
    if (jb > Ny - 2) return;
    antitrap_band<T>(psi, psi_n, c, c_n, Nx, Ny, a, g % ntx, jb, min(jb + kAtBand - 1, Ny - 2));
}

template <typename P>
static hipError_t launch_at(const P* psi, const P* psi_n, const double* c, double* c_n, int Nx,
                            int Ny, const AtArgs& a, hipStream_t s, LaunchEv ev) {
    const int nb = (Ny - 2 + kAtBand - 1) / kAtBand;
    dim3 block(kWave, 4);
    dim3 grid((Nx - 2 + kTileOut - 1) / kTileOut, (nb + 3) / 4);
    hipExtLaunchKernelGGL(antitrap_kernel<P>, grid, block, 0, s, ev.start, ev.stop, 0, psi, psi_n, c,
                          c_n, Nx, Ny, a);
    return hipGetLastError();
}


hipError_t launch_antitrap(const double* psi, const double* psi_n, const double* c, double* c_n,
                           int Nx, int Ny, double dx, double dt, double at, double k, double vdn,
                           hipStream_t s, LaunchEv ev) {
    return launch_at(psi, psi_n, c, c_n, Nx, Ny, at_args(dx, dt, at, k, vdn), s, ev);
}

hipError_t launch_antitrap_f32(const float* psi, const float* psi_n, const double* c, double* c_n,
                               int Nx, int Ny, double dx, double dt, double at, double k, double vdn,
                               hipStream_t s, LaunchEv ev) {
    return launch_at(psi, psi_n, c, c_n, Nx, Ny, at_args(dx, dt, at, k, vdn), s, ev);
}

}  // namespace rs

namespace rs {

// Per-row data of the tip fit (one wavefront per row): the last column in
// [0, Nx-2] with psi > pin (0 if none) by a wavefront max, then psi at that
// column +-2 (clamped) and at columns 0..2. The host fit needs nothing else
// (precise_tip_rows), so a tip output moves 36 B per row instead of the field.
template <typename P>
__global__ __launch_bounds__(256) void tip_detail_kernel(const P* psi, int Nx, int Ny, double pin,
                                                         int* irow, double* vals) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (j >= Ny) return;
    const P* row = psi + (int64_t)j * Nx;
    int best = -1;
    for (int i = lane; i <= Nx - 2; i += 64)
        if ((double)row[i] > pin) best = i;
    for (int o = 32; o >= 1; o >>= 1) best = max(best, __shfl_xor(best, o, 64));
    if (lane < 8) {
        const int i0 = best < 0 ? 0 : best;
        int i = lane < 5 ? i0 - 2 + lane : lane - 5;
        i = min(max(i, 0), Nx - 1);
        vals[8 * (int64_t)j + lane] = (double)row[i];
        if (lane == 0) irow[j] = i0;
    }
}

hipError_t launch_tip_detail(const double* psi, int Nx, int Ny, double pin, int* irow,
                             double* vals, hipStream_t s) {
    hipLaunchKernelGGL(tip_detail_kernel<double>, dim3((Ny + 3) / 4), dim3(256), 0, s, psi, Nx, Ny,
                       pin, irow, vals);
    return hipGetLastError();
}

hipError_t launch_tip_detail_f32(const float* psi, int Nx, int Ny, double pin, int* irow,
                                 double* vals, hipStream_t s) {
    hipLaunchKernelGGL(tip_detail_kernel<float>, dim3((Ny + 3) / 4), dim3(256), 0, s, psi, Nx, Ny,
                       pin, irow, vals);
    return hipGetLastError();
}

}  // namespace rs
