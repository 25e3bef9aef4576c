// rs_step64.hip -- the fused FP64 step kernel of the time-stepper for gfx950.
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
// This file is built with FMA contraction on (the compiler fuses multiplies
// and adds, as the reference's CUDA build does by default); the support
// kernels in rs_kernels.hip round every operation as written.
#include <hip/hip_ext.h>

#include "rs_kernels.hpp"
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

// three wavefronts per SIMD (168 VGPRs, 16 spilled): 303 us for a 4096^2
// step on a developed front, 371 us at two (191 VGPRs, none spilled); with
// every division and log as the reference writes them and no contraction,
// 537 us at three against 598 us at two and 720 us at four
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
    Cell s0 = make_cell(a, lntab, a.psi[(int64_t)(jb - 1) * Nx + ic],
                        a.c[(int64_t)(jb - 1) * Nx + ic], ic, jb - 1);
    Cell s1 = make_cell(a, lntab, a.psi[(int64_t)jb * Nx + ic], a.c[(int64_t)jb * Nx + ic], ic, jb);
    Cell s2;

    // prefetch of row jb+1
    double nx_psi = a.psi[(int64_t)(jb + 1) * Nx + ic];
    double nx_c = a.c[(int64_t)(jb + 1) * Nx + ic];
    double lh_acc = 0.0;

    // output row j from rows m = j-1, z = j and p = j+1 (filled here)
    auto row = [&](int j, const Cell& m, const Cell& z, Cell& p) {
        const double cur_psi = nx_psi, cur_c = nx_c;
        {  // unconditional, the last row reloading row je+1 (no wait on the stores)
            const int jn = min(j + 2, je + 1);
            nx_psi = a.psi[(int64_t)jn * Nx + ic];
            nx_c = a.c[(int64_t)jn * Nx + ic];
        }
        p = make_cell(a, lntab, cur_psi, cur_c, ic, j + 1);
        // the i+-1 neighbours of all three rows, shifted anew each row (eight
        // more DPP moves; carried from row to row they held 16 more VGPRs,
        // which spilled)
        const double phL_m = shfl_up1(m.phi), phR_m = shfl_dn1(m.phi);
        const double SL_m = shfl_up1(m.S), SR_m = shfl_dn1(m.S);
        const double phL_z = shfl_up1(z.phi), phR_z = shfl_dn1(z.phi);
        const double SL_z = shfl_up1(z.S), SR_z = shfl_dn1(z.S);
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
    };
    // three rows per trip with the window's roles rotated: no register
    // copies of the window between rows
    for (int j = jb; j <= je; j += 3) {
        row(j, s0, s1, s2);
        if (j + 1 > je) break;
        row(j + 1, s1, s2, s0);
        if (j + 2 > je) break;
        row(j + 2, s2, s0, s1);
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

}  // namespace rs
