// host_checks.cpp -- drives the library's host-side code and the CPU
// oracle under AddressSanitizer and UndefinedBehaviorSanitizer (built by
// tests/sanitize/Makefile, run by tests/test_host_sanitizers.py). No GPU:
// only calls that do no device work. Exits non-zero on the first failed
// check; the sanitizers abort on the first finding.
//
// Covered: parameter presets and set_parameters (rs_params_*, rs_derive),
// the tip fit (rs::precise_tip: row scans, quartic and parabola solves),
// every file format (field text, tip rows, offset and history records,
// checkpoints, parameters.txt write/read, Ti-Ag cross sections and break
// points), argument rejections, the oracle's whole-run paths (coupled
// and frozen T, pull-backs of both signs, antitrapping, tip outputs, the
// Ti-Ag restatement), the FP64 step kernel's form of the CALPHAD
// evaluation against the reference's form, and the table log against the C
// library's.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../2d-rapid-solidification_amd/csrc/rs_host.hpp"
#include "../../include/rs_peritectic.h"
#include "../../include/rsolid.h"
#include "../../oracle/rs_oracle.h"
#include "../../oracle/tp_oracle.h"

static int g_fail = 0;
#define CHECK(cond)                                                     \
    do {                                                                \
        if (!(cond)) {                                                  \
            fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
            g_fail = 1;                                                 \
        }                                                               \
    } while (0)

static std::string tmp_path(const char* dir, const char* name) { return std::string(dir) + "/" + name; }

static void params_and_derive() {
    const int grids[][2] = {{768, 832}, {512, 272}, {300, 290}};
    for (auto& g : grids) {
        rs_params p;
        CHECK(rs_params_reference(&p, g[0], g[1]) == RS_OK);
        rs_constants a, b;
        CHECK(rs_derive(&p, &a) == RS_OK);
        CHECK(or_derive(&p, &b) == RS_OK);
        CHECK(memcmp(&a, &b, sizeof a) == 0);
    }
    for (int n : {8, 22, 64, 256, 1000}) {
        for (int tfc : {0, 1}) {
            rs_params p;
            if (rs_params_config(&p, n, tfc, 100.0) != RS_OK) continue;
            rs_constants a, b;
            CHECK(rs_derive(&p, &a) == RS_OK);
            CHECK(or_derive(&p, &b) == RS_OK);
            CHECK(a.dt == b.dt && a.NxT_all == b.NxT_all && a.step_pull_back == b.step_pull_back);
        }
    }
    CHECK(rs_params_reference(nullptr, 8, 8) == RS_ERR_ARG);
    CHECK(rs_derive(nullptr, nullptr) == RS_ERR_ARG);
}

// an oracle run with pull-backs and tip outputs; returns the final psi
static std::vector<double> oracle_run(rs_params p, int steps, int spb, uint64_t seed, bool ref) {
    or_state* s = or_create(&p);
    CHECK(s != nullptr);
    if (!s) return {};
    CHECK((ref ? or_init_reference(s) : or_init_config(s, seed)) == RS_OK);
    rs_constants k;
    or_get_constants(s, &k);
    k.step_pull_back = spb;
    or_set_constants(s, &k);
    rs_tip_row t;
    for (int q = 0; q < steps; q += 25) {
        CHECK(or_step(s, 25) == RS_OK);
        (void)or_tip(s, &t);
    }
    // pull-backs of both signs and beyond the grid
    (void)or_pull_back(s, 3);
    (void)or_pull_back(s, -2);
    (void)or_pull_back(s, 0);
    const double m = or_solute_mass(s);
    CHECK(isfinite(m));
    std::vector<double> psi((size_t)p.Nx * p.Ny), c(psi.size()), Tf(psi.size());
    CHECK(or_get(s, RS_FIELD_PSI, psi.data(), (int64_t)psi.size()) == RS_OK);
    CHECK(or_get(s, RS_FIELD_C, c.data(), (int64_t)c.size()) == RS_OK);
    CHECK(or_get(s, RS_FIELD_TFINE, Tf.data(), (int64_t)Tf.size()) == RS_OK);
    CHECK(or_get(s, RS_FIELD_C, c.data(), 3) == RS_ERR_ARG);
    or_destroy(s);
    return psi;
}

static void oracle_and_tip() {
    or_set_threads(2);
    rs_params p;
    rs_params_config(&p, 96, 1, 100.0);
    std::vector<double> psi = oracle_run(p, 200, 50, 7, false);
    rs::TipFit f;
    (void)rs::precise_tip(p, psi.data(), &f);
    int32_t ti, tj;
    double tx, ty, tR;
    (void)or_precise_tip(&p, psi.data(), &ti, &tj, &tx, &ty, &tR);
    CHECK(f.i == ti && f.j == tj);
    // frozen T, antitrapping with trapping, a rotated grain
    rs_params q;
    rs_params_config(&q, 64, 0, 100.0);
    q.at_coef = 1.0 / (2.0 * sqrt(2.0));
    q.V_D = q.Vp / 2.0;
    q.alpha0 = 10.0;
    q.beta0 = 15.0;
    q.gamma0 = 25.0;
    (void)oracle_run(q, 100, 40, 11, false);
    // reference preset, reference initialiser, ragged grid
    rs_params r;
    rs_params_reference(&r, 300, 97);
    std::vector<double> psi_r = oracle_run(r, 50, 25, 0, true);
    (void)rs::precise_tip(r, psi_r.data(), &f);
    // a field with no interface at all and one fully solid
    std::vector<double> flat((size_t)r.Nx * r.Ny, -1.0);
    (void)rs::precise_tip(r, flat.data(), &f);
    std::fill(flat.begin(), flat.end(), 1.0);
    (void)rs::precise_tip(r, flat.data(), &f);
}

static void formats(const char* dir) {
    const int Nx = 13, Ny = 7;
    std::vector<double> a((size_t)Nx * Ny);
    for (size_t q = 0; q < a.size(); ++q) a[q] = sin(0.37 * (double)q) * 1e3;
    CHECK(rs_write_field_txt(tmp_path(dir, "f.txt").c_str(), a.data(), Nx, Ny, 2, 3) == RS_OK);
    CHECK(rs_write_field_txt(tmp_path(dir, "f.txt").c_str(), a.data(), Nx, Ny, 0, 1) == RS_ERR_ARG);
    rs_tip_row t{};
    t.step = 5;
    t.V_tip = NAN;
    CHECK(rs_write_tip_row(tmp_path(dir, "tip.txt").c_str(), &t, 1) == RS_OK);
    CHECK(rs_write_tip_row(tmp_path(dir, "tip.txt").c_str(), &t, 0) == RS_OK);
    CHECK(rs_append_offset(tmp_path(dir, "offset.txt").c_str(), 1025, 3, 7, 1025) == RS_OK);
    int32_t ig = 0;
    CHECK(rs_append_history(tmp_path(dir, "hist.txt").c_str(), a.data(), 3, Ny, 9, 1, 2, &ig) == RS_OK);
    CHECK(rs_append_history(tmp_path(dir, "hist.txt").c_str(), a.data(), 0, Ny, 9, 1, 2, &ig) == RS_OK);
    std::vector<double> b(a.size());
    CHECK(rs_write_checkpoint(tmp_path(dir, "ck.bin").c_str(), a.data(), (int64_t)a.size()) == RS_OK);
    CHECK(rs_read_checkpoint(tmp_path(dir, "ck.bin").c_str(), b.data(), (int64_t)b.size()) == RS_OK);
    CHECK(a == b);
    CHECK(rs_read_checkpoint(tmp_path(dir, "ck.bin").c_str(), b.data(), (int64_t)b.size() + 1) != RS_OK);
    CHECK(rs_read_checkpoint(tmp_path(dir, "missing.bin").c_str(), b.data(), 4) != RS_OK);
    rs_params p;
    rs_params_config(&p, 256, 1, 100.0);
    rs_constants k, k2;
    rs_derive(&p, &k);
    k.step = 12345;
    k.l_offset_Vp = 3.25;
    CHECK(rs_write_parameters(tmp_path(dir, "parameters.txt").c_str(), &p, &k, 77) == RS_OK);
    rs_derive(&p, &k2);
    CHECK(rs_read_parameters(tmp_path(dir, "parameters.txt").c_str(), &p, &k2) == RS_OK);
    CHECK(k2.step == k.step);
    CHECK(rs_read_parameters(tmp_path(dir, "missing.txt").c_str(), &p, &k2) != RS_OK);
    // a truncated parameters file
    FILE* f = fopen(tmp_path(dir, "short.txt").c_str(), "w");
    fputs("12\n", f);
    fclose(f);
    (void)rs_read_parameters(tmp_path(dir, "short.txt").c_str(), &p, &k2);
}

static void peritectic(const char* dir) {
    tp_params p;
    CHECK(tp_params_reference(&p, 24, 20, 18) == RS_OK);
    p.zcut_from_top = 12;
    p.R = 5.0;
    p.ncaps = 1;
    p.cap_x[0] = 12.0;
    p.cap_y[0] = 10.0;
    double d[6];
    CHECK(tp_derived(&p, d, 6) == RS_OK);
    int32_t sz = 0;
    CHECK(tp_abi_info(&sz) == RS_OK && sz == (int32_t)sizeof(tp_params));
    tpo_state* s = tpo_create(&p);
    CHECK(s != nullptr);
    tpo_init_reference(s);
    tpo_step(s, 20);
    std::vector<double> c((size_t)p.Nx * p.Ny * p.Nz), back(c.size());
    CHECK(tpo_get(s, TP_FIELD_C, c.data(), (int64_t)c.size()) == 0);
    CHECK(tpo_get(s, TP_FIELD_MU, back.data(), (int64_t)back.size()) == 0);
    double rep[2];
    tpo_report(s, 5, rep);
    CHECK(isfinite(rep[0]));
    tpo_destroy(s);
    CHECK(tp_write_cross_section(tmp_path(dir, "xz.txt").c_str(), c.data(), p.Nx, p.Ny, p.Nz, 0, 3) == RS_OK);
    CHECK(tp_write_cross_section(tmp_path(dir, "yz.txt").c_str(), c.data(), p.Nx, p.Ny, p.Nz, 1, 23) == RS_OK);
    CHECK(tp_write_cross_section(tmp_path(dir, "bad.txt").c_str(), c.data(), p.Nx, p.Ny, p.Nz, 0, 20) == RS_ERR_ARG);
    CHECK(tp_write_break_point(tmp_path(dir, "bp.vtk").c_str(), c.data(), p.Nx, p.Ny, p.Nz) == RS_OK);
    CHECK(tp_read_break_point(tmp_path(dir, "bp.vtk").c_str(), back.data(), p.Nx, p.Ny, p.Nz) == RS_OK);
    CHECK(back == c);
    CHECK(tp_read_break_point(tmp_path(dir, "bp.vtk").c_str(), back.data(), p.Nx, p.Ny, p.Nz + 1) != RS_OK);
    CHECK(tp_read_break_point(tmp_path(dir, "xz.txt").c_str(), back.data(), p.Nx, p.Ny, p.Nz) != RS_OK);
}

// thermo_fp64<true> (one reciprocal of T for the SGTE terms in 1/T, and
// log(c/(1-c)) as log c - log(1-c)) against the form the reference writes
// (fun.cu:1628-1720) over the SGTE ranges and the composition branches: the
// two agree to rounding (at most 6e-11 J/mol on values of 1e2-1e5 J/mol), and
// both are NaN outside the ranges. A wrong coefficient or sign in either
// form moves a value by 1e-3 J/mol or more.
static void thermo_forms() {
    rs_params p;
    CHECK(rs_params_reference(&p, 512, 272) == RS_OK);
    const rs::RK rk = rs::rk_of(p);
    const double cs[] = {0.0, 1e-9, 1e-8, 2e-8, 1e-6, 0.01, 0.1, 0.3, 0.5,
                         0.6325, 0.9, 0.999999, 1.0 - 2e-8, 1.0 - 1e-9, 1.0};
    double worst = 0.0;
    int n = 0;
    for (double T = 250.0; T <= 4100.0; T += 3.7) {
        for (double c : cs) {
            const rs::Thermo a = rs::thermo_fp64<false>(rk, T, c);
            const rs::Thermo b = rs::thermo_fp64<true>(rk, T, c);
            const double va[3] = {a.dG, a.muL, a.muS}, vb[3] = {b.dG, b.muL, b.muS};
            for (int q = 0; q < 3; ++q) {
                CHECK(isnan(va[q]) == isnan(vb[q]));
                if (isnan(va[q])) continue;
                const double d = fabs(va[q] - vb[q]);
                CHECK(d <= 1e-11 + 1e-13 * fabs(va[q]));
                if (d > worst) worst = d;
                ++n;
            }
        }
    }
    CHECK(n > 1000);
    printf("thermo forms: %d values, largest difference %.3g J/mol\n", n, worst);
}

// ln_tab (the table log of the FP64 step kernel and of the FP32 paths'
// ln(c/(1-c))) against the C library's log, over the arguments the steppers
// give it: within one ulp or 6e-17 (the rounding of the table's ln F, which
// is what is left where the result cancels to near 0, r near 1), and equal
// to the library's value for at least 97% of the temperatures (log T carries
// the largest multipliers of the SGTE terms). The table is rebuilt here as
// the build's generator writes it: tab[2k] = 1/F_k, tab[2k+1] = ln F_k.
static void table_log() {
    double tab[256];
    for (int k = 0; k < 128; ++k) {
        const double F = 1.0 + (k + 0.5) / 128.0;
        tab[2 * k] = (double)(1.0L / (long double)F);
        tab[2 * k + 1] = (double)logl((long double)F);
    }
    uint64_t st = 0x9E3779B97F4A7C15ull;
    auto next = [&st]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return (double)(st >> 11) * 0x1p-53;
    };
    long n[2] = {0, 0}, same[2] = {0, 0};
    auto one = [&](double r, int g) {
        const double a = rs::ln_tab(r, tab), b = log(r);
        const double ulp = nextafter(fabs(b), INFINITY) - fabs(b);
        CHECK(fabs(a - b) <= ulp + 6e-17);
        same[g] += a == b;
        ++n[g];
    };
    for (int q = 0; q < 400000; ++q) {
        one(exp(log(1e-8) * next()), 0);  // compositions (1e-8, 1], log-uniform
        one(1.0 - 0.999 * next(), 0);     // their complements, near 1
        one(250.0 + 3850.0 * next(), 1);  // temperatures
    }
    for (int b = -30; b <= 12; ++b)
        for (double f : {1.0, 1.0 + 0x1p-52, 2.0 - 0x1p-52, 1.00390625, 1.99609375})
            one(ldexp(f, b), 0);
    CHECK(same[1] >= 0.97 * n[1]);
    printf("table log: %ld arguments, equal to libm for %.4f of the temperatures and %.4f of the rest\n",
           n[0] + n[1], (double)same[1] / n[1], (double)same[0] / n[0]);
}

int main(int argc, char** argv) {
    const char* dir = argc > 1 ? argv[1] : ".";
    params_and_derive();
    oracle_and_tip();
    formats(dir);
    peritectic(dir);
    thermo_forms();
    table_log();
    printf(g_fail ? "host checks FAILED\n" : "host checks ok\n");
    return g_fail;
}
