// rs_kernels.hpp -- launch-side declarations of the HIP kernels.
#pragma once

#include "rs_physics.hpp"

namespace rs {

// Column tile of the fused step kernel: one wavefront spans 64 consecutive
// columns; lanes 1..62 own output columns, lanes 0 and 63 are the halo.
constexpr int kWave = 64;
constexpr int kTileOut = kWave - 2;

// The dispatcher hands workgroup w (in launch order) to XCD w % 8, each XCD
// with its own L2. xcd_order(w, n) is the position of w in the order that
// gives each XCD a contiguous run of logical indices (a bijection of
// [0, n)), so workgroups with neighbouring logical indices -- neighbouring
// column tiles, which read the same 128-byte lines at their shared edges --
// share an L2 and run at about the same time.
constexpr int kXcds = 8;
__host__ __device__ __forceinline__ int xcd_order(int w, int n) {
    const int xcd = w % kXcds, k = w / kXcds;
    const int full = n / kXcds, rem = n % kXcds;
    return xcd * full + (xcd < rem ? xcd : rem) + k;
}

// Fine-grid temperature source: coupled coarse grid (bilinear, fun.cu:1580)
// or frozen gradient (get_delta, fun.cu:1610).
struct TempSrc {
    const double* T;  // coarse T (NxT_all pitch), TFC only
    int tfc;
    int ratio;        // ratio_dxT_dx
    int NxT_all;
    int ile;          // i_left_end used for the interpolation
    double dxT;       // dx*ratio (y-index offset quirk, fun.cu:1590)
    // frozen: delta(i) = (d0 + ((((GG*(i+ioff))*dx)*S)*W0)/dT0) - dl
    int fta_ioff;
    int S;
    double d0, GG, dx, W0, dT0, dl, Tmelt;
};

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

struct StepArgsF64 {
    const double* psi;
    const double* c;
    double* psi_n;
    double* c_n;
    double* lh;       // latent-heat column sums, (NyT-2) x Nx
    int Nx, Ny;
    int band_h;       // output rows per wavefront band (multiple of ratio)
    int nbands;
    int lh_last_row;  // last fine row that contributes latent heat, 10*(NyT-2)
    int lh_on;
    TempSrc ts;
    double dt, AA, q_solid, dG_den, mu_den, kdc, dx2;
    RK rk;
    Aniso<double> an;
    // solute ledger (or null): the solute the edge cells gain through the
    // boundary faces, accumulated over the steps (led[j] x = 1/2 face of row
    // j, led[Ny + j] x = Nx - 3/2 face, led[2 Ny + i] y = 1/2 face of column
    // i, led[2 Ny + Nx + i] y = Ny - 3/2 face)
    double* led;
};

struct TempArgs {
    const double* T;
    double* Tn;
    const double* lh;     // latent-heat column sums (FP64 path), or null
    const float* lh32;    // the same in FP32 (FP32 paths), or null
    int Nx, NxT, NyT, NxT_all, ratio, ile, ratio_dt_dtT;
    double vadv;      // Vp_ramp(t) tau0 / (S W0)
    double dxT, dtT, DT, Gmu /* Gamma*muk0 */, Lcp, Tcold, Thot;
    float* T32n;      // optional FP32 copy of T_next - Tref (mixed path), or null
    double Tref;
};

// integrate_phi (fun.cu:1494-1535) from the column sums + compute_temperature
// (fun.cu:1537-1578) for coarse cell (iT, jT) < (NxT, NyT), race-free
// boundaries: the owner of row NyT-2 writes row 0, the owner of row 1 writes
// row NyT-1, Dirichlet columns by their own threads
// sum of one row's latent-heat column sums i0..in, in column order, in FP64
template <typename L>
__device__ __forceinline__ double lh_block_sum(const L* row, int i0, int in) {
    double integral = 0.0;
    if (in - i0 == 9) {
        // a full block of 10 columns (the library's ratio): all loads in
        // flight before the in-order sum
        double v[10];
#pragma unroll
        for (int q = 0; q < 10; ++q) v[q] = (double)row[i0 + q];
#pragma unroll
        for (int q = 0; q < 10; ++q) integral += v[q];
    } else {
        for (int i = i0; i <= in; ++i) integral += (double)row[i];
    }
    return integral;
}

__device__ __forceinline__ void temperature_cell(const TempArgs& a, int iT, int jT) {
    const int64_t P = a.NxT_all;
    if (iT == 0 || iT == a.NxT - 1) {
        const double v = (iT == 0) ? a.Tcold : a.Thot;
        a.Tn[jT * P + iT] = v;
        if (a.T32n) a.T32n[jT * P + iT] = (float)(v - a.Tref);
        return;
    }
    if (jT == 0 || jT == a.NyT - 1) return;
    const double* T = a.T;
    const double T00 = T[jT * P + iT];
    const double T10 = T[jT * P + iT + 1];
    const double Tm0 = T[jT * P + iT - 1];
    const double T01 = T[(jT + 1) * P + iT];
    const double T0m = T[(jT - 1) * P + iT];
    const int r = a.ratio;
    const int iT1 = a.ile / r;
    const int iTn = (a.ile + a.Nx - 3) / r;
    double dphidt = 0.0;
    if (iT >= iT1 && iT <= iTn) {
        int i0 = r * iT + 1 - a.ile;
        int in = i0 + r - 1;
        if (i0 < 1) i0 = 1;
        if (in > a.Nx - 2) in = a.Nx - 2;
        const int64_t off = (int64_t)(jT - 1) * a.Nx;
        const double integral =
            a.lh32 ? lh_block_sum(a.lh32 + off, i0, in) : lh_block_sum(a.lh + off, i0, in);
        dphidt = integral * 0.5;
    }
    const double lap = (T10 + Tm0 + T01 + T0m - 4.0 * T00) / (a.dxT * a.dxT);
    const double dTdt = a.vadv * (T10 - Tm0) / (2.0 * a.dxT) + lap * a.DT / a.Gmu;
    const double tn = T00 + a.dtT * dTdt + a.Lcp * dphidt / (a.ratio_dt_dtT * (r * r));
    a.Tn[jT * P + iT] = tn;
    if (jT == a.NyT - 2) a.Tn[iT] = tn;                      // row 0 <- row NyT-2
    if (jT == 1) a.Tn[(int64_t)(a.NyT - 1) * P + iT] = tn;  // row NyT-1 <- row 1
    if (a.T32n) {
        const float t32 = (float)(tn - a.Tref);
        a.T32n[jT * P + iT] = t32;
        if (jT == a.NyT - 2) a.T32n[iT] = t32;
        if (jT == 1) a.T32n[(int64_t)(a.NyT - 1) * P + iT] = t32;
    }
}

// Timing of a launch by its own dispatch: hipExtLaunchKernel writes the
// kernel's start and end times into these events (null: untimed), so the
// interval is the kernel alone, with no marker packets around it.
struct LaunchEv {
    hipEvent_t start = nullptr, stop = nullptr;
};

hipError_t launch_step_fp64(const StepArgsF64& a, int waves_per_block, hipStream_t s,
                            LaunchEv ev = {});
hipError_t launch_temperature(const TempArgs& a, hipStream_t s, LaunchEv ev = {});
hipError_t launch_init_T(double* T0, double* T1, int NxT_all, int NyT_all, double Tcold, int S,
                         double W0, double dx, int r, double GG, hipStream_t s);
hipError_t launch_init_fields(double* psi, double* c, const double* x0, int Nx, int Ny,
                              double dx, int mode, double cinf, double c_liq, double c_sol,
                              hipStream_t s);
hipError_t launch_pull_back(const double* psi, const double* c, double* psi_n, double* c_n,
                            int Nx, int Ny, int i_offset, double cinf, hipStream_t s);
hipError_t launch_tip_column(const double* psi, int Nx, int Ny, double pin, int* rowbuf,
                             int* out, hipStream_t s);
hipError_t launch_mass(const double* c, int Nx, int Ny, double* rowbuf, double* out,
                       hipStream_t s);
hipError_t launch_interp(const TempSrc& ts, double* Tfine, int Nx, int Ny, hipStream_t s);
hipError_t launch_copy(const double* a, double* b, int64_t n, hipStream_t s);
// fixed-order FP64 sum of v[lo..hi] into out[0] (one workgroup)
hipError_t launch_sum(const double* v, int lo, int hi, double* out, hipStream_t s);
// solute ledger: acc[0] += m[1] - m[0] (the mass change of one pull-back)
hipError_t launch_ledger_pull(const double* m, double* acc, hipStream_t s);
// number of non-finite values of f[0, n) and the smallest index holding one
// (out[0] += count, out[1] = min(out[1], index); out must be initialised to
// 0 and ~0)
hipError_t launch_nonfinite(const double* f, int64_t n, unsigned long long* out, hipStream_t s);
hipError_t launch_nonfinite_f32(const float* f, int64_t n, unsigned long long* out, hipStream_t s);
// per-row data of the tip fit: irow[Ny], vals[8 Ny] (see precise_tip_rows)
hipError_t launch_tip_detail(const double* psi, int Nx, int Ny, double pin, int* irow,
                             double* vals, hipStream_t s);
hipError_t launch_tip_detail_f32(const float* psi, int Nx, int Ny, double pin, int* irow,
                                 double* vals, hipStream_t s);
// antitrapping + solute-trapping correction of c_n (extension; psi in FP64
// or FP32); vdn = V_D in W/tau0 units (0: no trapping factor)
hipError_t launch_antitrap(const double* psi, const double* psi_n, const double* c, double* c_n,
                           int Nx, int Ny, double dx, double dt, double at, double k, double vdn,
                           hipStream_t s, LaunchEv ev = {});
hipError_t launch_antitrap_f32(const float* psi, const float* psi_n, const double* c, double* c_n,
                               int Nx, int Ny, double dx, double dt, double at, double k, double vdn,
                               hipStream_t s, LaunchEv ev = {});

// host-side helpers that fill the temperature source from the run state
TempSrc make_tempsrc(const rs_params& p, const rs_constants& k, const double* Tcoarse, int ile);

}  // namespace rs
