// rs_tip.hip -- host tip fit (get_precise_tip and helpers, fun.cu:2628-2954).
//
// Fit chain: integer tip from a running row scan whose sub-cell estimate
// multiplies by the psi jump instead of dividing (fun.cu:2884-2886, kept),
// a least-squares quartic x(psi) on five columns around the last psi > 0 cell
// of five rows, solved by Gaussian elimination with a pivot permutation,
// and a least-squares parabola x(y) whose vertex and curvature give
// (x0, y0, R).
#include <math.h>

#include <vector>

#include "rs_host.hpp"

namespace rs {
namespace {

int wrap_row(int x, int a, int b) {  // periodic deal_with_boundary (fun.cu:2628-2644)
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    if (x <= lo - 1) return x + (hi - lo + 1);
    if (x >= hi + 1) return x - (hi - lo + 1);
    return x;
}

int last_above(const rs_params& p, const double* psi, int start, int j) {
    int it = 0;
    const double* row = psi + (int64_t)j * p.Nx;
    for (int i = start; i < p.Nx - 1; ++i)
        if (row[i] > p.phi_interface) it = i;
    return it;
}

// solve A c = b, elimination with partial pivoting through a permutation
void solve5(double A[5][5], double b[5], double x[5]) {
    int perm[5] = {0, 1, 2, 3, 4};
    for (int k = 0; k < 4; ++k) {
        double best = fabs(A[perm[k]][k]);
        int arg = k;
        for (int i = k + 1; i < 5; ++i) {
            const double v = fabs(A[perm[i]][k]);
            if (v > best) {
                best = v;
                arg = i;
            }
        }
        const int t = perm[k];
        perm[k] = perm[arg];
        perm[arg] = t;
        const double* pk = A[perm[k]];
        for (int i = k + 1; i < 5; ++i) {
            double* pi = A[perm[i]];
            const double f = pi[k] / pk[k];
            for (int c = k; c < 5; ++c) pi[c] -= f * pk[c];
            b[perm[i]] -= f * b[perm[k]];
        }
    }
    for (int i = 4; i >= 0; --i) {
        double acc = 0.0;
        for (int c = i + 1; c < 5; ++c) acc += A[perm[i]][c] * x[c];
        x[i] = (b[perm[i]] - acc) / A[perm[i]][i];
    }
}

double quartic_at(const double* xs, const double* ys, double yq) {
    double A[5][5] = {}, b[5] = {}, cf[5] = {};
    for (int q = 0; q < 5; ++q) {
        double yp[9];
        yp[0] = 1.0;
        for (int k = 1; k <= 8; ++k) yp[k] = yp[k - 1] * ys[q];
        for (int r = 0; r < 5; ++r) {
            b[r] += xs[q] * yp[r];
            for (int c = 0; c < 5; ++c) A[r][c] += yp[r + c];
        }
    }
    solve5(A, b, cf);
    return cf[0] + cf[1] * yq + cf[2] * yq * yq + cf[3] * yq * yq * yq + cf[4] * yq * yq * yq * yq;
}

int solve3(double A[3][3], double b[3], double x[3]) {
    for (int k = 0; k < 3; ++k) {
        int pv = k;
        double mv = fabs(A[k][k]);
        for (int i = k + 1; i < 3; ++i)
            if (fabs(A[i][k]) > mv) {
                mv = fabs(A[i][k]);
                pv = i;
            }
        if (mv < 1e-14) return -1;
        if (pv != k) {
            for (int c = 0; c < 3; ++c) {
                const double t = A[k][c];
                A[k][c] = A[pv][c];
                A[pv][c] = t;
            }
            const double t = b[k];
            b[k] = b[pv];
            b[pv] = t;
        }
        const double piv = A[k][k];
        for (int i = k + 1; i < 3; ++i) {
            const double f = A[i][k] / piv;
            A[i][k] = 0.0;
            for (int c = k + 1; c < 3; ++c) A[i][c] -= f * A[k][c];
            b[i] -= f * b[k];
        }
    }
    for (int i = 2; i >= 0; --i) {
        double acc = b[i];
        for (int c = i + 1; c < 3; ++c) acc -= A[i][c] * x[c];
        x[i] = acc / A[i][i];
    }
    return 0;
}

int parabola(const double* xs, const double* ys, double* x0, double* y0, double* R) {
    double M[3][3] = {}, v[3] = {}, a[3] = {};
    for (int q = 0; q < 5; ++q) {
        const double f[3] = {1.0, ys[q], ys[q] * ys[q]};
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) M[r][c] += f[r] * f[c];
            v[r] += f[r] * xs[q];
        }
    }
    if (solve3(M, v, a) != 0) return -2;
    if (fabs(a[2]) < 1e-14) return -3;
    const double Rl = 1.0 / (2.0 * a[2]);
    const double yl = -a[1] / (2.0 * a[2]);
    *x0 = a[0] - (yl * yl) / (2.0 * Rl);
    *y0 = yl;
    *R = Rl;
    return 0;
}

}  // namespace

// Per-row data of a psi field for the fit: irow[j] = last i in [0, Nx-2]
// with psi > phi_interface (0 if none); vals[8j..8j+4] = psi at
// irow-2..irow+2 (clamped to the row), vals[8j+5..8j+7] = psi at 0, 1, 2.
void tip_rows_host(const rs_params& p, const double* psi, int* irow, double* vals) {
    const int Nx = p.Nx;
    for (int j = 0; j < p.Ny; ++j) {
        const double* row = psi + (int64_t)j * Nx;
        const int i0 = last_above(p, psi, 0, j);
        irow[j] = i0;
        for (int m = 0; m < 5; ++m) {
            int i = i0 - 2 + m;
            if (i < 0) i = 0;
            if (i > Nx - 1) i = Nx - 1;
            vals[8 * j + m] = row[i];
        }
        for (int m = 0; m < 3; ++m) vals[8 * j + 5 + m] = row[m];
    }
}

// The fit from the per-row data. A scan started at column s finds the
// row's last psi > phi_interface at or after s, i.e. irow when irow >= s
// and column 0 otherwise, so the running scans of fun.cu:2867-2954 are
// reproduced exactly without the field.
int precise_tip_rows(const rs_params& p, const int* irow, const double* vals, TipFit* out) {
    const int Ny = p.Ny;
    const double pin = p.phi_interface;
    auto above = [&](int j, int start) {
        return (irow[j] >= start && vals[8 * j + 2] > pin) ? irow[j] : 0;
    };
    // psi at column i0 - 2 + m of row j for the i0 the scan found
    auto window = [&](int j, int i0, int m) {
        if (i0 == irow[j]) return vals[8 * j + m];
        int i = i0 - 2 + m;  // i0 == 0
        if (i < 0) i = 0;
        return vals[8 * j + 5 + i];
    };
    double itip = 0.0;
    int bi = 0, bj = 0;
    for (int j = 0; j < Ny; ++j) {
        const int i0 = above(j, (int)itip);
        const double a = window(j, i0, 2), b = window(j, i0, 3);
        const double est = (b - a) * (pin - a) + i0;
        if (est > itip) {
            itip = est;
            bi = (int)itip;
            bj = j;
        }
    }
    double xd[5], yd[5];
    for (int q = 0; q < 5; ++q) {
        const int j0 = wrap_row(bj - 2 + q, Ny - 2, 1);
        const int start = bi - 5 < 0 ? 0 : bi - 5;
        const int i0 = above(j0, start);
        double id[5], pd[5];
        for (int m = 0; m < 5; ++m) {
            id[m] = i0 - 2 + m;
            pd[m] = window(j0, i0, m);
        }
        xd[q] = quartic_at(id, pd, pin);
        yd[q] = j0;
    }
    double x0 = NAN, y0 = NAN, R = NAN;
    const int st = parabola(xd, yd, &x0, &y0, &R);
    out->i = bi;
    out->j = bj;
    out->x = x0;
    out->y = y0;
    double Rn = fabs(R) * p.dx * p.W0 * p.S;
    if (Rn > 1e5) Rn = 1e5;
    if (Rn < 1.0) Rn = 1.0;
    out->R = Rn;
    return st;
}

int precise_tip(const rs_params& p, const double* psi, TipFit* out) {
    std::vector<int> irow((size_t)p.Ny);
    std::vector<double> vals((size_t)p.Ny * 8);
    tip_rows_host(p, psi, irow.data(), vals.data());
    return precise_tip_rows(p, irow.data(), vals.data(), out);
}

}  // namespace rs
