// rs_kernels.hip -- FP64 support kernels of the time-stepper for gfx950:
// coarse temperature (integrate_phi's block sums + compute_temperature),
// initialisers, pull-back, tip-column and solute-mass reductions, the fine-T
// interpolation, the finite-value sentinel, the antitrapping pass and the
// tip detail rows. The fused FP64 step kernel is in rs_step64.hip.
#include <hip/hip_ext.h>

#include "rs_kernels.hpp"
#include "rs_antitrap.hpp"

namespace rs {

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
