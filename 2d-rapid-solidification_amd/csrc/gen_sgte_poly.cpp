// gen_sgte_poly.cpp -- build-time generator of the mixed stepper's SGTE
// polynomial table (writes rs_sgte_poly.inc).
//
// The temperature dependence of the SGTE unaries (rs_physics.hpp,
// fun.cu:1628-1666) is fitted piecewise on [1541 K, 2749 K] in 100 K pieces
// split at the 1941 K Ti breakpoint, as polynomials of degree kDeg in
// t = (T - Tmid)/Thalf. The kernel evaluates them in FP32 with the constant
// term in FP64; the coefficients become instruction immediates. Fit: long
// double least squares on 64 Chebyshev nodes per piece. The generator also
// prints the largest error of the FP32 evaluation against the FP64 routine.
#include <math.h>
#include <stdio.h>

#include "rs_physics.hpp"

namespace {

constexpr int kDeg = 5;
constexpr int kFun = 3;
constexpr double kTref = 1941.0;  // T' = T - kTref
const double kBreaks[] = {1541.0, 1641.0, 1741.0, 1841.0, 1941.0, 2041.0, 2141.0,
                          2241.0, 2341.0, 2441.0, 2541.0, 2641.0, 2749.0};
constexpr int kPieces = sizeof(kBreaks) / sizeof(kBreaks[0]) - 1;

void funcs(double T, double f[kFun]) {
    const rs::Unaries u = rs::sgte_unaries(T, log(T));
    f[0] = u.ti_liq - u.nb_liq;  // U_L
    f[1] = u.ti_liq - u.ti_bcc;  // dG_Ti
    f[2] = u.nb_liq - u.nb_bcc;  // dG_Nb
}

struct Piece {
    float cen, isc;  // in T'
    double c0[kFun];
    float c[kFun][kDeg];
};

void fit(double a, double b, Piece* P) {
    const double mid = 0.5 * (a + b), half = 0.5 * (b - a);
    P->cen = (float)(mid - kTref);
    P->isc = (float)(1.0 / half);
    const double hw = 1.0 / (double)P->isc;
    long double A[kDeg + 1][kDeg + 1] = {};
    long double B[kFun][kDeg + 1] = {};
    const int N = 64;
    for (int q = 0; q < N; ++q) {
        const double tq = cos(M_PI * (q + 0.5) / N);
        double f[kFun];
        funcs(kTref + (double)P->cen + tq * hw, f);
        long double tp[2 * kDeg + 1];
        tp[0] = 1.0L;
        for (int k = 1; k <= 2 * kDeg; ++k) tp[k] = tp[k - 1] * (long double)tq;
        for (int r = 0; r <= kDeg; ++r) {
            for (int c = 0; c <= kDeg; ++c) A[r][c] += tp[r + c];
            for (int g = 0; g < kFun; ++g) B[g][r] += tp[r] * (long double)f[g];
        }
    }
    for (int k = 0; k <= kDeg; ++k)
        for (int r = k + 1; r <= kDeg; ++r) {
            const long double m = A[r][k] / A[k][k];
            for (int c = k; c <= kDeg; ++c) A[r][c] -= m * A[k][c];
            for (int g = 0; g < kFun; ++g) B[g][r] -= m * B[g][k];
        }
    for (int g = 0; g < kFun; ++g) {
        long double x[kDeg + 1];
        for (int r = kDeg; r >= 0; --r) {
            long double acc = B[g][r];
            for (int c = r + 1; c <= kDeg; ++c) acc -= A[r][c] * x[c];
            x[r] = acc / A[r][r];
        }
        P->c0[g] = (double)x[0];
        for (int k = 1; k <= kDeg; ++k) P->c[g][k - 1] = (float)x[k];
    }
}

double eval(const Piece& P, int g, float Tp) {
    const float t = (Tp - P.cen) * P.isc;
    float acc = P.c[g][kDeg - 1];
    for (int k = kDeg - 2; k >= 0; --k) acc = fmaf(acc, t, P.c[g][k]);
    return P.c0[g] + (double)(acc * t);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: gen_sgte_poly OUT.inc\n");
        return 2;
    }
    Piece P[kPieces];
    double worst = 0.0;
    for (int p = 0; p < kPieces; ++p) {
        fit(kBreaks[p], kBreaks[p + 1], &P[p]);
        for (int q = 0; q < 2000; ++q) {
            const double T = kBreaks[p] + (kBreaks[p + 1] - kBreaks[p]) * q / 2000.0;
            const float Tp = (float)(T - kTref);
            double f[kFun];
            funcs(kTref + (double)Tp, f);
            for (int g = 0; g < kFun; ++g) worst = fmax(worst, fabs(eval(P[p], g, Tp) - f[g]));
        }
    }
    FILE* o = fopen(argv[1], "w");
    if (!o) return 1;
    fprintf(o, "// generated by gen_sgte_poly.cpp -- do not edit\n");
    fprintf(o, "// SGTE temperature polynomials, %d pieces on [%.0f, %.0f] K, degree %d\n",
            kPieces, kBreaks[0], kBreaks[kPieces], kDeg);
    fprintf(o, "// largest FP32-evaluation error vs the FP64 routine: %.3e J/mol\n", worst);
    fprintf(o, "#define RS_SGTE_TREF %.17g\n", kTref);
    fprintf(o, "#define RS_SGTE_PIECES %d\n#define RS_SGTE_DEG %d\n", kPieces, kDeg);
    fprintf(o, "#define RS_SGTE_MAXERR %.17g\n", worst);
    fprintf(o, "#define RS_SGTE_T0 %.9ef\n#define RS_SGTE_T1 %.9ef\n", (float)(kBreaks[0] - kTref),
            (float)(kBreaks[kPieces] - kTref));
    // one X-macro line per piece: index, cen, isc, 3 FP64 constants, then the
    // t^1..t^5 coefficients of U_L, dG_Ti, dG_Nb
    fprintf(o, "#define RS_SGTE_TABLE(X) \\\n");
    for (int p = 0; p < kPieces; ++p) {
        fprintf(o, "    X(%d, %.9ef, %.9ef, %.17e, %.17e, %.17e", p, P[p].cen, P[p].isc, P[p].c0[0],
                P[p].c0[1], P[p].c0[2]);
        for (int g = 0; g < kFun; ++g)
            for (int k = 0; k < kDeg; ++k) fprintf(o, ", %.9ef", P[p].c[g][k]);
        fprintf(o, ")%s\n", p < kPieces - 1 ? " \\" : "");
    }
    // natural-log table: for F_k = 1 + (k + 1/2)/128, {1/F_k, ln F_k}
    fprintf(o, "#define RS_LN_TAB_N 128\n");
    fprintf(o, "static __device__ const double kLnTab[2 * RS_LN_TAB_N] = {\n");
    for (int k = 0; k < 128; ++k) {
        const long double F = 1.0L + (k + 0.5L) / 128.0L;
        fprintf(o, "    %.17e, %.17e,\n", (double)(1.0L / F), (double)logl(F));
    }
    fprintf(o, "};\n");
    fclose(o);
    printf("gen_sgte_poly: %d pieces, worst FP32 evaluation error %.3e J/mol\n", kPieces, worst);
    return 0;
}
