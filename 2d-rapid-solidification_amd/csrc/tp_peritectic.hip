// tp_peritectic.hip -- 3D three-phase peritectic Ti-Ag stepper on one GPU
// (include/rs_peritectic.h): the fused step kernel, initial condition,
// reductions, and the C ABI.
//
// One step of the reference (TiAg.cu:288-345) is computeMu + boundaryMu
// {X,Y,Z}, then diffuse + boundary{X,Y,Z}, then a pointer swap: 88 B per
// cell move through HBM (mu written and read back). Here it is one launch:
// mu is a pointwise function of (phi1, phi2, c), so the kernel evaluates it
// for the cells it loads and never stores it (48 B per cell: read and write
// phi1, phi2, c). A workgroup owns a 64 x 8 column tile and marches up a run
// of z planes: the current plane's phi1, phi2 and mu (tile plus a one-cell
// ring) sit in LDS for the x/y neighbours, the planes below and above in
// registers for the z neighbours. Ghost cells are written by the thread
// that owns their source cell, in the reference's boundary order (periodic
// x, then periodic y, then no-flux z: every ghost takes the new value of
// its image), so no second pass and no race.
//
// Operation order follows the reference expressions and the file is built
// with -ffp-contract=off; there are no transcendental functions, so the
// kernel reproduces the CPU restatement bit for bit.
#include <stdio.h>
#include <string.h>

#include <math.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "../../include/rs_peritectic.h"
#include "rs_kernels.hpp"

namespace tp {

// model coefficients of one run (TiAg.cu:36-62 and the r_i, s_i of
// computeMu / diffuse, TiAg.cu:454, 496-497)
struct Coef {
    double DT, DXin, DYin, DZin;
    double a1, a2, a3, b, lambda, alpha, Da, Db;
    double r1, r2, r3, s1, s2, s3;
};

// interpolation g_i = g(p_i, p_j, p_k) and its partial derivatives with
// respect to p_i (gp1) and p_j (gp2), TiAg.cu:71-73
__device__ __forceinline__ double g_of(double pi, double pj, double pk) {
    const double d = pk - pj;
    return ((pi * pi) * 0.25) *
           (((15.0 * (1.0 - pi)) * ((1.0 + pi) - (d * d))) + (pi * ((9.0 * (pi * pi)) - 5.0)));
}
__device__ __forceinline__ double gp1_of(double pi, double pj, double pk) {
    const double d = pj - pk;
    const double pi2 = pi * pi;
    return (3.75 * pi) *
           (((((-4.0 * pi2) + (3.0 * (pi2 * pi))) + 2.0) - pi) + (((3.0 * pi) - 2.0) * (d * d)));
}
__device__ __forceinline__ double gp2_of(double pi, double pj, double pk) {
    return ((7.5 * (pi - 1.0)) * (pi * pi)) * (pj - pk);
}

// computeMu (TiAg.cu:446-456)
__device__ __forceinline__ double mu_of(const Coef& k, double p1, double p2, double c) {
    const double p3 = (1.0 - p1) - p2;
    return ((c - k.r1 * g_of(p1, p2, p3)) - k.r2 * g_of(p2, p3, p1)) - k.r3 * g_of(p3, p1, p2);
}

struct Nb {
    double q1, q2, mu;
};

// one face of the solute flux (TiAg.cu:526-531): mobility of the face times
// the mu difference (toward the neighbour for the + face, from it for the -)
template <bool MOB23>
__device__ __forceinline__ double face_mob(const Coef& k, double p1, double p2, double p3, const Nb& q) {
    double m = 0.5 * (p1 + q.q1);
    if (MOB23)
        m = (m + (k.Da * 0.5) * (p2 + q.q2)) + (k.Db * 0.5) * (((p3 + 1.0) - q.q1) - q.q2);
    return m;
}

// diffuse (TiAg.cu:486-534) for one interior cell. nb: x+, x-, y+, y-, z+, z-.
template <bool MOB23>
__device__ __forceinline__ void update(const Coef& k, double p1, double p2, double muu, double c,
                                       const Nb* nb, double& o1, double& o2, double& oc) {
    const double p3 = (1.0 - p1) - p2;
    const double e1 = k.s1 - k.r1 * muu, e2 = k.s2 - k.r2 * muu, e3 = k.s3 - k.r3 * muu;
    const double fc1 = k.lambda * ((gp1_of(p1, p2, p3) * e1 + gp2_of(p2, p1, p3) * e2) + gp2_of(p3, p1, p2) * e3);
    const double fc2 = k.lambda * ((gp2_of(p1, p2, p3) * e1 + gp1_of(p2, p3, p1) * e2) + gp2_of(p3, p2, p1) * e3);
    const double fc3 = k.lambda * ((gp2_of(p1, p3, p2) * e1 + gp2_of(p2, p3, p1) * e2) + gp1_of(p3, p1, p2) * e3);

    const double l1 = ((((nb[0].q1 + nb[1].q1) - 2.0 * p1) * k.DXin) * k.DXin +
                       (((nb[2].q1 + nb[3].q1) - 2.0 * p1) * k.DYin) * k.DYin) +
                      (((nb[4].q1 + nb[5].q1) - 2.0 * p1) * k.DZin) * k.DZin;
    const double l2 = ((((nb[0].q2 + nb[1].q2) - 2.0 * p2) * k.DXin) * k.DXin +
                       (((nb[2].q2 + nb[3].q2) - 2.0 * p2) * k.DYin) * k.DYin) +
                      (((nb[4].q2 + nb[5].q2) - 2.0 * p2) * k.DZin) * k.DZin;
    const double l3 = (-l1) - l2;

    const double tb = 2.0 * k.b;
    const double f1 = (((((2.0 * p1) * (1.0 - p1)) * (1.0 - 2.0 * p1) +
                         (k.a1 * ((p2 * p3) * (p2 * p3))) * (3.0 + tb * p1)) +
                        (((2.0 * k.a2) * p1) * (p3 * p3)) * ((((3.0 * p1) * p3) + 3.0 * p2) + k.b * (p2 * p2))) +
                       (((2.0 * k.a3) * p1) * (p2 * p2)) * ((((3.0 * p1) * p2) + 3.0 * p3) + k.b * (p3 * p3))) +
                      fc1;
    const double f2 = (((((2.0 * p2) * (1.0 - p2)) * (1.0 - 2.0 * p2) +
                         (((2.0 * k.a1) * p2) * (p3 * p3)) * ((((3.0 * p2) * p3) + 3.0 * p1) + k.b * (p1 * p1))) +
                        (k.a2 * ((p1 * p3) * (p1 * p3))) * (3.0 + tb * p2)) +
                       (((2.0 * k.a3) * p2) * (p1 * p1)) * ((((3.0 * p1) * p2) + 3.0 * p3) + k.b * (p3 * p3))) +
                      fc2;
    const double f3 = (((((2.0 * p3) * (1.0 - p3)) * (1.0 - 2.0 * p3) +
                         (((2.0 * k.a1) * p3) * (p2 * p2)) * ((((3.0 * p2) * p3) + 3.0 * p1) + k.b * (p1 * p1))) +
                        (((2.0 * k.a2) * p3) * (p1 * p1)) * ((((3.0 * p1) * p3) + 3.0 * p2) + k.b * (p2 * p2))) +
                       (k.a3 * ((p1 * p2) * (p1 * p2))) * (3.0 + tb * p3)) +
                      fc3;
    const double m1 = f1 - l1, m2 = f2 - l2, m3 = f3 - l3;
    const double third = 0.333333333333333333333333333;
    const double t1 = (-(((2.0 * m1) - m2) - m3)) * third;
    const double t2 = (-(((2.0 * m2) - m1) - m3)) * third;
    o1 = p1 + k.DT * t1;
    o2 = p2 + k.DT * t2;

    const double hxp = (face_mob<MOB23>(k, p1, p2, p3, nb[0]) * (nb[0].mu - muu)) * k.DXin;
    const double hxn = (face_mob<MOB23>(k, p1, p2, p3, nb[1]) * (muu - nb[1].mu)) * k.DXin;
    const double hyp = (face_mob<MOB23>(k, p1, p2, p3, nb[2]) * (nb[2].mu - muu)) * k.DYin;
    const double hyn = (face_mob<MOB23>(k, p1, p2, p3, nb[3]) * (muu - nb[3].mu)) * k.DYin;
    const double hzp = (face_mob<MOB23>(k, p1, p2, p3, nb[4]) * (nb[4].mu - muu)) * k.DZin;
    const double hzn = (face_mob<MOB23>(k, p1, p2, p3, nb[5]) * (muu - nb[5].mu)) * k.DZin;
    oc = c + (k.DT * k.alpha) * ((((hxp - hxn) * k.DXin) + ((hyp - hyn) * k.DYin)) + ((hzp - hzn) * k.DZin));
}

constexpr int kBX = 64, kBY = 8;              // output tile (one wavefront per row)
constexpr int kLX = kBX + 2, kLY = kBY + 2;   // with the one-cell ring
constexpr int kRing = 2 * kLX + 2 * kBY;      // ring cells, one per thread tid < kRing

struct StepArgs {
    const double* p1;
    const double* p2;
    const double* c;
    double* n1;
    double* n2;
    double* nc;
    int Nx, Ny, Nz;
    int zseg, ntx, nty, nseg;
    Coef k;
};

// boundary image of a ghost index (periodic x/y, TiAg.cu:458-474; no-flux
// z, TiAg.cu:476-483)
__device__ __forceinline__ int img_p(int i, int n) { return i == 0 ? n - 2 : (i == n - 1 ? 1 : i); }
__device__ __forceinline__ int img_z(int k, int n) { return k == 0 ? 1 : (k == n - 1 ? n - 2 : k); }

// A cell as the stencil sees it: phi1, phi2 and c as stored, mu of the
// cell's boundary image (computeMu runs on every cell and boundaryMu then
// overwrites the ghosts with their images' mu). Cells outside the grid are
// never read by an interior cell's stencil; they load zeros.
__device__ __forceinline__ void load_cell(const StepArgs& a, int i, int j, int k, double& q1,
                                          double& q2, double& qc, double& mu) {
    if (i < 0 || i >= a.Nx || j < 0 || j >= a.Ny) {
        q1 = q2 = qc = mu = 0.0;
        return;
    }
    const size_t o = ((size_t)k * a.Ny + j) * a.Nx + i;
    q1 = a.p1[o];
    q2 = a.p2[o];
    qc = a.c[o];
    const int ii = img_p(i, a.Nx), jj = img_p(j, a.Ny), kk = img_z(k, a.Nz);
    if (ii == i && jj == j && kk == k) {
        mu = mu_of(a.k, q1, q2, qc);
    } else {
        const size_t m = ((size_t)kk * a.Ny + jj) * a.Nx + ii;
        mu = mu_of(a.k, a.p1[m], a.p2[m], a.c[m]);
    }
}

// the new value of interior cell (i, j, k) and of every ghost whose image
// it is (boundaryX, then boundaryY, then boundaryZ: TiAg.cu:538-579)
__device__ __forceinline__ void store_cell(const StepArgs& a, int i, int j, int k, double o1,
                                           double o2, double oc) {
    const int xs1 = i == a.Nx - 2 ? 0 : (i == 1 ? a.Nx - 1 : -1);
    const int ys1 = j == a.Ny - 2 ? 0 : (j == 1 ? a.Ny - 1 : -1);
    const int zs1 = k == a.Nz - 2 ? a.Nz - 1 : (k == 1 ? 0 : -1);
    const int o = (k * a.Ny + j) * a.Nx + i;
    a.n1[o] = o1;
    a.n2[o] = o2;
    a.nc[o] = oc;
    if (xs1 < 0 && ys1 < 0 && zs1 < 0) return;
    const int xs[2] = {i, xs1}, ys[2] = {j, ys1}, zs[2] = {k, zs1};
    for (int q = 1; q < 8; ++q) {
        const int x = xs[q & 1], y = ys[(q >> 1) & 1], z = zs[q >> 2];
        if (x < 0 || y < 0 || z < 0) continue;
        const int g = (z * a.Ny + y) * a.Nx + x;
        a.n1[g] = o1;
        a.n2[g] = o2;
        a.nc[g] = oc;
    }
}

// phi1, phi2, c of a cell as stored (zeros outside the grid: such cells
// are never read by an interior cell's stencil). 32-bit cell offsets: the
// grid holds fewer than 2^31 cells (checked at tp_create).
__device__ __forceinline__ void raw_load(const StepArgs& a, int i, int j, int k, double& q1,
                                         double& q2, double& qc) {
    if (i < 0 || i >= a.Nx || j < 0 || j >= a.Ny) {
        q1 = q2 = qc = 0.0;
        return;
    }
    const int o = (k * a.Ny + j) * a.Nx + i;
    q1 = a.p1[o];
    q2 = a.p2[o];
    qc = a.c[o];
}

// mu of a cell whose stored values are (q1, q2, qc): a ghost takes its
// image's mu, read here (ghosts are rare: the grid's faces)
__device__ __forceinline__ double mu_at(const StepArgs& a, int i, int j, int k, double q1, double q2,
                                        double qc) {
    if (i < 0 || i >= a.Nx || j < 0 || j >= a.Ny) return 0.0;
    const int ii = img_p(i, a.Nx), jj = img_p(j, a.Ny), kk = img_z(k, a.Nz);
    if (ii != i || jj != j || kk != k) {
        const int m = (kk * a.Ny + jj) * a.Nx + ii;
        q1 = a.p1[m];
        q2 = a.p2[m];
        qc = a.c[m];
    }
    return mu_of(a.k, q1, q2, qc);
}

// Software pipeline over the planes: while plane k is updated from the LDS
// copy of plane k and the own column's planes k - 1 and k + 1 in registers,
// the loads of plane k + 2 are in flight; plane k + 1's mu and LDS copy are
// finished at the top of the next trip, one plane after its loads were
// issued. One barrier per plane (double-buffered LDS).
template <bool MOB23>
__global__ __launch_bounds__(kBX * kBY, 2) void step_kernel(StepArgs a) {
    __shared__ double s1[2][kLY][kLX], s2[2][kLY][kLX], sm[2][kLY][kLX];
    const int L = rs::xcd_order(blockIdx.x, gridDim.x);
    const int txt = L % a.ntx, rest = L / a.ntx;
    const int tyt = rest % a.nty, seg = rest / a.nty;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * kBX + tx;
    const int i = txt * kBX + tx, j = tyt * kBY + ty;
    const int k0 = 1 + seg * a.zseg;
    const int k1 = min(k0 + a.zseg, a.Nz - 1);
    // this thread's ring cell (LDS coordinates hx, hy)
    const bool ring = tid < kRing;
    int hx = 0, hy = 0;
    if (tid < kLX) {
        hx = tid;
        hy = 0;
    } else if (tid < 2 * kLX) {
        hx = tid - kLX;
        hy = kLY - 1;
    } else if (tid < 2 * kLX + kBY) {
        hx = 0;
        hy = 1 + tid - 2 * kLX;
    } else if (ring) {
        hx = kLX - 1;
        hy = 1 + tid - 2 * kLX - kBY;
    }
    const int hi = txt * kBX + hx - 1, hj = tyt * kBY + hy - 1;
    const bool inter = i >= 1 && i <= a.Nx - 2 && j >= 1 && j <= a.Ny - 2;

    double m1, m2, mc, mm;  // own cell, plane k - 1
    double c1, c2, cc, cm;  // own cell, plane k
    double p1, p2, pc;      // own cell, plane k + 1 (as loaded)
    double h1 = 0.0, h2 = 0.0, hc = 0.0;  // ring cell, plane k + 1 (as loaded)
    raw_load(a, i, j, k0 - 1, m1, m2, mc);
    mm = mu_at(a, i, j, k0 - 1, m1, m2, mc);
    raw_load(a, i, j, k0, c1, c2, cc);
    cm = mu_at(a, i, j, k0, c1, c2, cc);
    s1[0][ty + 1][tx + 1] = c1;
    s2[0][ty + 1][tx + 1] = c2;
    sm[0][ty + 1][tx + 1] = cm;
    if (ring) {
        double q1, q2, qc;
        raw_load(a, hi, hj, k0, q1, q2, qc);
        s1[0][hy][hx] = q1;
        s2[0][hy][hx] = q2;
        sm[0][hy][hx] = mu_at(a, hi, hj, k0, q1, q2, qc);
        raw_load(a, hi, hj, k0 + 1, h1, h2, hc);
    }
    raw_load(a, i, j, k0 + 1, p1, p2, pc);
    __syncthreads();
    int b = 0;
    for (int k = k0; k < k1; ++k) {
        // plane k + 1: mu of the own and ring cells, its LDS copy
        const double pm = mu_at(a, i, j, k + 1, p1, p2, pc);
        s1[b ^ 1][ty + 1][tx + 1] = p1;
        s2[b ^ 1][ty + 1][tx + 1] = p2;
        sm[b ^ 1][ty + 1][tx + 1] = pm;
        if (ring) {
            s1[b ^ 1][hy][hx] = h1;
            s2[b ^ 1][hy][hx] = h2;
            sm[b ^ 1][hy][hx] = mu_at(a, hi, hj, k + 1, h1, h2, hc);
        }
        // plane k + 2 in flight during the update of plane k
        double n1 = 0.0, n2 = 0.0, nc = 0.0;
        if (k + 2 <= k1) {
            raw_load(a, i, j, k + 2, n1, n2, nc);
            if (ring) raw_load(a, hi, hj, k + 2, h1, h2, hc);
        }
        if (inter) {
            Nb nb[6];
            nb[0] = Nb{s1[b][ty + 1][tx + 2], s2[b][ty + 1][tx + 2], sm[b][ty + 1][tx + 2]};
            nb[1] = Nb{s1[b][ty + 1][tx], s2[b][ty + 1][tx], sm[b][ty + 1][tx]};
            nb[2] = Nb{s1[b][ty + 2][tx + 1], s2[b][ty + 2][tx + 1], sm[b][ty + 2][tx + 1]};
            nb[3] = Nb{s1[b][ty][tx + 1], s2[b][ty][tx + 1], sm[b][ty][tx + 1]};
            nb[4] = Nb{p1, p2, pm};
            nb[5] = Nb{m1, m2, mm};
            double o1, o2, oc;
            update<MOB23>(a.k, c1, c2, cm, cc, nb, o1, o2, oc);
            store_cell(a, i, j, k, o1, o2, oc);
        }
        __syncthreads();
        b ^= 1;
        m1 = c1;
        m2 = c2;
        mm = cm;
        c1 = p1;
        c2 = p2;
        cc = pc;
        cm = pm;
        p1 = n1;
        p2 = n2;
        pc = nc;
    }
}

// mu of every cell (computeMu + boundaryMu{X,Y,Z}: ghosts take their
// images' mu)
__global__ __launch_bounds__(256) void mu_kernel(StepArgs a, double* mu) {
    const size_t n = (size_t)a.Nx * a.Ny * a.Nz;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % a.Nx);
        const int j = (int)((q / a.Nx) % a.Ny);
        const int k = (int)(q / ((size_t)a.Nx * a.Ny));
        double q1, q2, qc, qm;
        load_cell(a, i, j, k, q1, q2, qc, qm);
        mu[q] = qm;
    }
}

// initial() (TiAg.cu:411-444): radius, roff and the compositions come from
// the host (the same libm as the CPU restatement)
struct InitArgs {
    double* p1;
    double* p2;
    double* c;
    int Nx, Ny, Nz, zcut, ncaps;
    double radius, roff, ca, cb, clb;
    double cx[4], cy[4];
};

__global__ __launch_bounds__(256) void init_kernel(InitArgs a) {
    const size_t n = (size_t)a.Nx * a.Ny * a.Nz;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % a.Nx);
        const int j = (int)((q / a.Nx) % a.Ny);
        const int k = (int)(q / ((size_t)a.Nx * a.Ny));
        bool cap = false;
        const double dz = k - (a.zcut - a.roff);
        for (int s = 0; s < a.ncaps; ++s) {
            const double dx = i - a.cx[s], dy = j - a.cy[s];
            cap = cap || ((dx * dx + dy * dy) + dz * dz <= a.radius * a.radius);
        }
        double v1, v2, vc;
        if (k >= a.zcut && cap) {
            v1 = 0.0;
            v2 = 1.0;
            vc = a.ca;
        } else if (k < a.zcut) {
            v1 = 0.0;
            v2 = 0.0;
            vc = a.cb;
        } else {
            v1 = 1.0;
            v2 = 0.0;
            vc = a.clb;
        }
        a.p1[q] = v1;
        a.p2[q] = v2;
        a.c[q] = vc;
    }
}

// deterministic FP64 reductions: per-plane sums of the interior c (fixed
// per-thread order, fixed tree), then the planes in order; largest |c| over
// the interior of one plane
__global__ __launch_bounds__(256) void plane_sum_kernel(const double* c, int Nx, int Ny, double* part) {
    __shared__ double red[256];
    const int k = blockIdx.x + 1;
    const int w = Nx - 2, h = Ny - 2;
    double s = 0.0;
    for (int q = threadIdx.x; q < w * h; q += 256) {
        const int i = 1 + q % w, j = 1 + q / w;
        s += c[((size_t)k * Ny + j) * Nx + i];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void report_kernel(const double* c, const double* part, int nplanes,
                                                     int Nx, int Ny, int kc, double* out) {
    __shared__ double red[256];
    const int w = Nx - 2, h = Ny - 2;
    double mx = 0.0;
    for (int q = threadIdx.x; q < w * h; q += 256) {
        const int i = 1 + q % w, j = 1 + q / w;
        mx = fmax(mx, fabs(c[((size_t)kc * Ny + j) * Nx + i]));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = 0; q < nplanes; ++q) s += part[q];
        out[0] = s;
        out[1] = red[0];
    }
}

}  // namespace tp

using namespace tp;

struct tp_ctx {
    tp_params p;
    Coef k{};
    hipStream_t stream = nullptr;
    double* f[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    double* scratch = nullptr;  // mu download / reduction partials
    double* outd = nullptr;
    double* h_pin = nullptr;
    int cur = 0;
    int64_t step = 0;
    int zseg = 32;
    std::string err;
    std::vector<hipEvent_t> ev;
    int nev = 0;
};

namespace {

thread_local std::string g_tp_create_err;

int tfail(tp_ctx* ctx, int code, const char* what, hipError_t e = hipSuccess) {
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

#define TCHK(ctx, call)                                                  \
    do {                                                                 \
        hipError_t e_ = (call);                                          \
        if (e_ != hipSuccess) return tfail((ctx), RS_ERR_HIP, #call, e_); \
    } while (0)

Coef coef_of(const tp_params& p) {
    Coef k;
    k.DT = p.DT;
    k.DXin = p.DXin;
    k.DYin = p.DYin;
    k.DZin = p.DZin;
    k.a1 = p.a1;
    k.a2 = p.a2;
    k.a3 = p.a3;
    k.b = p.b;
    k.lambda = p.lambda;
    k.alpha = p.alpha;
    k.Da = p.Da;
    k.Db = p.Db;
    const double B1 = p.B11 + p.B12 * p.T, B2 = p.B21 + p.B22 * p.T;
    k.r1 = p.A1;
    k.r2 = -p.A1 - p.A2;
    k.r3 = -p.A1 + p.A2;
    k.s1 = B1;
    k.s2 = -B1 - B2;
    k.s3 = -B1 + B2;
    return k;
}

StepArgs step_args(tp_ctx* ctx) {
    StepArgs a;
    const int c0 = ctx->cur;
    a.p1 = ctx->f[c0][0];
    a.p2 = ctx->f[c0][1];
    a.c = ctx->f[c0][2];
    a.n1 = ctx->f[c0 ^ 1][0];
    a.n2 = ctx->f[c0 ^ 1][1];
    a.nc = ctx->f[c0 ^ 1][2];
    a.Nx = ctx->p.Nx;
    a.Ny = ctx->p.Ny;
    a.Nz = ctx->p.Nz;
    a.zseg = ctx->zseg;
    a.ntx = (a.Nx + kBX - 1) / kBX;
    a.nty = (a.Ny + kBY - 1) / kBY;
    a.nseg = (a.Nz - 2 + a.zseg - 1) / a.zseg;
    a.k = ctx->k;
    return a;
}

int check(const tp_params& p, std::string* why) {
    if (p.Nx < 4 || p.Ny < 4 || p.Nz < 4) {
        *why = "grid must be at least 4 x 4 x 4 (one interior cell between the ghosts on each axis)";
        return RS_ERR_ARG;
    }
    if ((int64_t)p.Nx * p.Ny * p.Nz >= ((int64_t)1 << 31)) {
        *why = "grid must hold fewer than 2^31 cells (32-bit cell offsets)";
        return RS_ERR_ARG;
    }
    if (p.ncaps < 0 || p.ncaps > 4) {
        *why = "ncaps must be 0..4";
        return RS_ERR_ARG;
    }
    if (!(p.DT > 0.0) || !(p.DXin > 0.0) || !(p.DYin > 0.0) || !(p.DZin > 0.0)) {
        *why = "DT and the inverse spacings must be positive";
        return RS_ERR_ARG;
    }
    return RS_OK;
}

}  // namespace

extern "C" {

int tp_abi_info(int32_t* size_of_params) {
    if (!size_of_params) return RS_ERR_ARG;
    *size_of_params = (int32_t)sizeof(tp_params);
    return RS_OK;
}

int tp_params_reference(tp_params* p, int32_t Nx, int32_t Ny, int32_t Nz) {
    if (!p) return RS_ERR_ARG;
    memset(p, 0, sizeof *p);
    p->Nx = Nx;
    p->Ny = Ny;
    p->Nz = Nz;
    p->zcut_from_top = 100;
    p->DT = 0.015;
    p->DXin = p->DYin = p->DZin = 2.0;
    p->a1 = 0.0;
    p->a2 = 4.43;
    p->a3 = 6.09;
    p->b = 12.0;
    p->lambda = 30.0;
    p->alpha = 2.0;
    p->Da = 0.0;
    p->Db = 0.0;
    p->A1 = 0.391701;
    p->A2 = 0.216598;
    p->B11 = 0.0685881;
    p->B21 = 0.037927;
    p->B12 = -0.466867;
    p->B22 = -0.0662661;
    p->T = 60.0 * 0.000321787;
    p->R = 30.0;
    p->theta = 1.33;
    p->ncaps = 2;
    p->cap_x[0] = 128.0;
    p->cap_y[0] = 128.0;
    p->cap_x[1] = 300.0;
    p->cap_y[1] = 240.0;
    return RS_OK;
}

int tp_derived(const tp_params* p, double* out, int32_t n) {
    if (!p || !out) return RS_ERR_ARG;
    const double B1 = p->B11 + p->B12 * p->T, B2 = p->B21 + p->B22 * p->T;
    const double v[6] = {((B1 + 0.5 * B2) / (p->A1 + 0.5 * p->A2) - p->A1) - p->A2,
                         ((B1 - 0.5 * B2) / (p->A1 - 0.5 * p->A2) - p->A1) + p->A2,
                         (B1 + 0.5 * B2) / (p->A1 + 0.5 * p->A2) + p->A1,
                         (B1 - 0.5 * B2) / (p->A1 - 0.5 * p->A2) + p->A1, B1, B2};
    for (int q = 0; q < n && q < 6; ++q) out[q] = v[q];
    return RS_OK;
}

int tp_create(const tp_params* p, tp_ctx** out) {
    if (!p || !out) {
        g_tp_create_err = "null argument";
        return RS_ERR_ARG;
    }
    *out = nullptr;
    std::string why;
    const int rc = check(*p, &why);
    if (rc) {
        g_tp_create_err = why;
        return rc;
    }
    tp_ctx* ctx = new (std::nothrow) tp_ctx;
    if (!ctx) return RS_ERR_NOMEM;
    ctx->p = *p;
    ctx->k = coef_of(*p);
    if (const char* e = getenv("TP_ZSEG")) {
        const int v = atoi(e);
        if (v >= 1 && v <= 4096) ctx->zseg = v;
    }
    const size_t nb = sizeof(double) * (size_t)p->Nx * p->Ny * p->Nz;
    bool ok = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess;
    for (int b = 0; b < 2 && ok; ++b)
        for (int f = 0; f < 3 && ok; ++f)
            ok = hipMalloc(&ctx->f[b][f], nb) == hipSuccess &&
                 hipMemsetAsync(ctx->f[b][f], 0, nb, ctx->stream) == hipSuccess;
    ok = ok && hipMalloc(&ctx->outd, sizeof(double) * 8) == hipSuccess &&
         hipHostMalloc(&ctx->h_pin, 64) == hipSuccess &&
         hipStreamSynchronize(ctx->stream) == hipSuccess;
    if (!ok) {
        g_tp_create_err = "device allocation failed";
        tp_destroy(ctx);
        return RS_ERR_NOMEM;
    }
    g_tp_create_err.clear();
    *out = ctx;
    return RS_OK;
}

int tp_destroy(tp_ctx* ctx) {
    if (!ctx) return RS_ERR_ARG;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (int b = 0; b < 2; ++b)
        for (int f = 0; f < 3; ++f)
            if (ctx->f[b][f]) (void)hipFree(ctx->f[b][f]);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->outd) (void)hipFree(ctx->outd);
    if (ctx->h_pin) (void)hipHostFree(ctx->h_pin);
    for (hipEvent_t e : ctx->ev) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RS_OK;
}

const char* tp_last_error(const tp_ctx* ctx) {
    if (ctx) return ctx->err.c_str();
    return g_tp_create_err.empty() ? "null context" : g_tp_create_err.c_str();
}

int tp_init_reference(tp_ctx* ctx) {
    if (!ctx) return RS_ERR_ARG;
    const tp_params& p = ctx->p;
    double d[6];
    tp_derived(&p, d, 6);
    InitArgs a;
    ctx->cur = 0;
    a.p1 = ctx->f[0][0];
    a.p2 = ctx->f[0][1];
    a.c = ctx->f[0][2];
    a.Nx = p.Nx;
    a.Ny = p.Ny;
    a.Nz = p.Nz;
    a.zcut = p.Nz - p.zcut_from_top;
    a.ncaps = p.ncaps;
    a.radius = p.R / sin(p.theta);
    a.roff = p.R / tan(p.theta);
    a.ca = d[0];
    a.cb = d[1];
    a.clb = d[3];
    for (int q = 0; q < 4; ++q) {
        a.cx[q] = p.cap_x[q];
        a.cy[q] = p.cap_y[q];
    }
    hipLaunchKernelGGL(init_kernel, dim3(2048), dim3(256), 0, ctx->stream, a);
    TCHK(ctx, hipGetLastError());
    TCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->step = 0;
    return RS_OK;
}

int tp_upload(tp_ctx* ctx, int32_t field, const double* host, int64_t count) {
    if (!ctx || !host) return RS_ERR_ARG;
    const int64_t n = (int64_t)ctx->p.Nx * ctx->p.Ny * ctx->p.Nz;
    if (field < TP_FIELD_PHI1 || field > TP_FIELD_C) return tfail(ctx, RS_ERR_ARG, "field cannot be uploaded");
    if (count != n) return tfail(ctx, RS_ERR_ARG, "field size mismatch");
    TCHK(ctx, hipMemcpyAsync(ctx->f[ctx->cur][field], host, sizeof(double) * n, hipMemcpyHostToDevice,
                             ctx->stream));
    TCHK(ctx, hipStreamSynchronize(ctx->stream));
    return RS_OK;
}

int tp_download(tp_ctx* ctx, int32_t field, double* host, int64_t count) {
    if (!ctx || !host) return RS_ERR_ARG;
    const int64_t n = (int64_t)ctx->p.Nx * ctx->p.Ny * ctx->p.Nz;
    if (count != n) return tfail(ctx, RS_ERR_ARG, "field size mismatch");
    const double* src = nullptr;
    if (field >= TP_FIELD_PHI1 && field <= TP_FIELD_C) {
        src = ctx->f[ctx->cur][field];
    } else if (field == TP_FIELD_MU) {
        if (!ctx->scratch) TCHK(ctx, hipMalloc(&ctx->scratch, sizeof(double) * n));
        hipLaunchKernelGGL(mu_kernel, dim3(2048), dim3(256), 0, ctx->stream, step_args(ctx), ctx->scratch);
        TCHK(ctx, hipGetLastError());
        src = ctx->scratch;
    } else {
        return tfail(ctx, RS_ERR_ARG, "bad field");
    }
    TCHK(ctx, hipMemcpyAsync(host, src, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    TCHK(ctx, hipStreamSynchronize(ctx->stream));
    return RS_OK;
}

int tp_step(tp_ctx* ctx, int64_t nsteps) {
    if (!ctx || nsteps < 0) return RS_ERR_ARG;
    const bool mob23 = ctx->p.Da != 0.0 || ctx->p.Db != 0.0;
    for (int64_t q = 0; q < nsteps; ++q) {
        const StepArgs a = step_args(ctx);
        // timed steps: the kernel's dispatch writes its start and end times
        // into an event pair (hipExtLaunchKernel), no marker packets
        const bool timed = ctx->nev >= 0 && ctx->nev + 2 <= 4096;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timed) {
            while ((int)ctx->ev.size() < ctx->nev + 2) {
                hipEvent_t e;
                TCHK(ctx, hipEventCreate(&e));
                ctx->ev.push_back(e);
            }
            e0 = ctx->ev[ctx->nev];
            e1 = ctx->ev[ctx->nev + 1];
        }
        const dim3 grid(a.ntx * a.nty * a.nseg), block(kBX, kBY);
        if (mob23)
            hipExtLaunchKernelGGL(step_kernel<true>, grid, block, 0, ctx->stream, e0, e1, 0, a);
        else
            hipExtLaunchKernelGGL(step_kernel<false>, grid, block, 0, ctx->stream, e0, e1, 0, a);
        TCHK(ctx, hipGetLastError());
        if (timed) ctx->nev += 2;
        ctx->cur ^= 1;
        ++ctx->step;
    }
    return RS_OK;
}

int tp_get_step(tp_ctx* ctx, int64_t* step) {
    if (!ctx || !step) return RS_ERR_ARG;
    *step = ctx->step;
    return RS_OK;
}

int tp_set_step(tp_ctx* ctx, int64_t step) {
    if (!ctx || step < 0) return RS_ERR_ARG;
    ctx->step = step;
    return RS_OK;
}

int tp_report(tp_ctx* ctx, int32_t k_check, double* out, int32_t n) {
    if (!ctx || !out) return RS_ERR_ARG;
    const tp_params& p = ctx->p;
    if (k_check < 1 || k_check > p.Nz - 2) return tfail(ctx, RS_ERR_ARG, "k_check must be an interior plane");
    if (!ctx->scratch)
        TCHK(ctx, hipMalloc(&ctx->scratch, sizeof(double) * (size_t)p.Nx * p.Ny * p.Nz));
    const double* c = ctx->f[ctx->cur][2];
    hipLaunchKernelGGL(plane_sum_kernel, dim3(p.Nz - 2), dim3(256), 0, ctx->stream, c, p.Nx, p.Ny,
                       ctx->scratch);
    TCHK(ctx, hipGetLastError());
    hipLaunchKernelGGL(report_kernel, dim3(1), dim3(256), 0, ctx->stream, c, ctx->scratch, p.Nz - 2, p.Nx,
                       p.Ny, k_check, ctx->outd);
    TCHK(ctx, hipGetLastError());
    TCHK(ctx, hipMemcpyAsync(ctx->h_pin, ctx->outd, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TCHK(ctx, hipStreamSynchronize(ctx->stream));
    const double cells = (double)(p.Nx - 2) * (p.Ny - 2) * (p.Nz - 2);
    const double v[2] = {ctx->h_pin[0] / cells, ctx->h_pin[1]};
    for (int q = 0; q < n && q < 2; ++q) out[q] = v[q];
    return RS_OK;
}

int tp_sync(tp_ctx* ctx) {
    if (!ctx) return RS_ERR_ARG;
    TCHK(ctx, hipStreamSynchronize(ctx->stream));
    return RS_OK;
}

int tp_timing(tp_ctx* ctx, int32_t reset, double* out, int32_t n) {
    if (!ctx) return RS_ERR_ARG;
    TCHK(ctx, hipStreamSynchronize(ctx->stream));
    double tot = 0.0;
    int cnt = 0;
    for (int q = 0; q + 1 < ctx->nev; q += 2) {
        float ms = 0.f;
        TCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[q], ctx->ev[q + 1]));
        tot += ms;
        ++cnt;
    }
    if (out) {
        const double v[2] = {cnt ? tot / cnt : 0.0, (double)cnt};
        for (int q = 0; q < n && q < 2; ++q) out[q] = v[q];
    }
    if (reset) ctx->nev = 0;
    return RS_OK;
}

// ---- files ----------------------------------------------------------------
int tp_write_cross_section(const char* path, const double* data, int32_t Nx, int32_t Ny, int32_t Nz,
                           int32_t plane, int32_t index) {
    if (!path || !data || Nx < 1 || Ny < 1 || Nz < 1) return RS_ERR_ARG;
    if (plane != 0 && plane != 1) return RS_ERR_ARG;
    if (index < 0 || index >= (plane == 0 ? Ny : Nx)) return RS_ERR_ARG;
    FILE* f = fopen(path, "w");
    if (!f) return RS_ERR_ARG;
    const int Ni = plane == 0 ? Nx : Ny;
    for (int k = 0; k < Nz; ++k) {
        for (int i = 0; i < Ni; ++i) {
            const size_t o = plane == 0 ? ((size_t)k * Ny + index) * Nx + i : ((size_t)k * Ny + i) * Nx + index;
            fprintf(f, "%.4f\t", data[o]);
        }
        fprintf(f, "\n");
    }
    return fclose(f) == 0 ? RS_OK : RS_ERR_ARG;
}

int tp_write_break_point(const char* path, const double* data, int32_t Nx, int32_t Ny, int32_t Nz) {
    if (!path || !data || Nx < 1 || Ny < 1 || Nz < 1) return RS_ERR_ARG;
    FILE* f = fopen(path, "wb");
    if (!f) return RS_ERR_ARG;
    const size_t n = (size_t)Nx * Ny * Nz;
    fprintf(f,
            "# vtk DataFile Version 3.0\nconcentration\nbinary\nDATASET STRUCTURED_POINTS\n"
            "DIMENSIONS %d %d %d\nASPECT_RATIO 1 1 1\nORIGIN 0 0 0\nPOINT_DATA %zu\n"
            "SCALARS c double 1\nLOOKUP_TABLE default\n",
            Nx, Ny, Nz, n);
    const size_t w = fwrite(data, sizeof(double), n, f);
    const int rc = fclose(f);
    return (w == n && rc == 0) ? RS_OK : RS_ERR_ARG;
}

int tp_read_break_point(const char* path, double* data, int32_t Nx, int32_t Ny, int32_t Nz) {
    if (!path || !data || Nx < 1 || Ny < 1 || Nz < 1) return RS_ERR_ARG;
    FILE* f = fopen(path, "rb");
    if (!f) return RS_ERR_ARG;
    // the header is 10 lines (load_break_point skips them, TiAg.cu:734-737)
    int lines = 0, ch;
    while (lines < 10 && (ch = fgetc(f)) != EOF)
        if (ch == '\n') ++lines;
    const size_t n = (size_t)Nx * Ny * Nz;
    const size_t r = lines == 10 ? fread(data, sizeof(double), n, f) : 0;
    fclose(f);
    return r == n ? RS_OK : RS_ERR_ARG;
}

}  // extern "C"
