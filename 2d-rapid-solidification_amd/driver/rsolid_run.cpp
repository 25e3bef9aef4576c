// rsolid_run -- command-line driver in place of the reference's main()
// (fun.cu:344-546): the same loop order, cadences and output files, with
// every field update in the HIP library behind include/rsolid.h.
//
//   rsolid_run [key=value ...]
//
// Keys of the reference's compile-time -D flags (run.sh:75-90), read at run
// time: Nx_all Ny_all total_time Vp GG eps1 epk1 path_input if_load
// run_time. Further keys: preset=reference|config, N, tfc, alpha (config
// preset), precision=fp64|mixed|surrogate, seed (config initialiser), at_coef V_D
// k_part (antitrapping/trapping extension), alpha0 beta0 gamma0 (grain
// orientation, degrees), di dj (output stride),
// state=1 (also write the exact state file for a bit-exact restart),
// load_state=FILE (restart from such a file instead of the checkpoints),
// pull_back=N (pull-back cadence in steps instead of the derived one),
// finite_check=N (stop with an error naming the field and cell when psi, c
// or the coarse T holds a NaN or Inf, checked every N steps).
//
// Files (under path_input, as the reference writes them): data/c_N.txt and
// data/phi0_N.txt every step_fields, tip.txt every step_tip,
// data/offset.txt and data/c_history_temp.txt at every pull-back,
// data/check_point_{c,phi,temperature}.bin and parameters.txt every
// step_check_point and at the end.
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <string>
#include <vector>

#include "rsolid.h"

namespace {

std::map<std::string, std::string> g_args;

const char* arg(const char* k, const char* def) {
    const auto it = g_args.find(k);
    return it == g_args.end() ? def : it->second.c_str();
}
double argd(const char* k, double def) {
    const auto it = g_args.find(k);
    return it == g_args.end() ? def : atof(it->second.c_str());
}
long argl(const char* k, long def) {
    const auto it = g_args.find(k);
    return it == g_args.end() ? def : atol(it->second.c_str());
}

void die(rs_ctx* ctx, int rc, const char* what) {
    fprintf(stderr, "rsolid_run: %s failed (%d): %s\n", what, rc, ctx ? rs_last_error(ctx) : "");
    exit(1);
}
#define CHECK(ctx, call)                    \
    do {                                    \
        const int rc_ = (call);             \
        if (rc_ != RS_OK) die(ctx, rc_, #call); \
    } while (0)

void ensure_dir(const std::string& d) {
    if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) {
        fprintf(stderr, "rsolid_run: cannot create %s\n", d.c_str());
        exit(1);
    }
}

// T field (NxT_all x NyT_all) as the reference's checkpoint stores it:
// NyT rows of NxT values (fun.cu:1874-1891)
std::vector<double> pack_T(const std::vector<double>& T, const rs_constants& k) {
    std::vector<double> out((size_t)k.NxT * k.NyT);
    for (int j = 0; j < k.NyT; ++j)
        for (int i = 0; i < k.NxT; ++i) out[(size_t)j * k.NxT + i] = T[(size_t)j * k.NxT_all + i];
    return out;
}

struct Run {
    rs_params p;
    rs_ctx* ctx = nullptr;
    std::string dir, data;
    int di = 1, dj = 1;
    int32_t i_ignore = 0;
    bool state_file = false;
};

void write_fields(Run& r, long step, const rs_constants& k) {
    const int64_t n = (int64_t)r.p.Nx * r.p.Ny;
    std::vector<double> c(n), phi(n);
    CHECK(r.ctx, rs_download(r.ctx, RS_FIELD_C, c.data(), n));
    CHECK(r.ctx, rs_download(r.ctx, RS_FIELD_PHI, phi.data(), n));
    const long index = step / k.step_fields;
    char name[512];
    snprintf(name, sizeof name, "%s/c_%ld.txt", r.data.c_str(), index);
    CHECK(r.ctx, rs_write_field_txt(name, c.data(), r.p.Nx, r.p.Ny, r.di, r.dj));
    snprintf(name, sizeof name, "%s/phi0_%ld.txt", r.data.c_str(), index);
    CHECK(r.ctx, rs_write_field_txt(name, phi.data(), r.p.Nx, r.p.Ny, r.di, r.dj));
}

void write_checkpoint(Run& r) {
    rs_constants k;
    CHECK(r.ctx, rs_get_constants(r.ctx, &k));
    const int64_t n = (int64_t)r.p.Nx * r.p.Ny;
    std::vector<double> c(n), phi(n);
    CHECK(r.ctx, rs_download(r.ctx, RS_FIELD_C, c.data(), n));
    CHECK(r.ctx, rs_download(r.ctx, RS_FIELD_PHI, phi.data(), n));
    CHECK(r.ctx, rs_write_checkpoint((r.data + "/check_point_c.bin").c_str(), c.data(), n));
    CHECK(r.ctx, rs_write_checkpoint((r.data + "/check_point_phi.bin").c_str(), phi.data(), n));
    if (r.p.ifTFC) {
        std::vector<double> T((size_t)k.NxT_all * k.NyT_all);
        CHECK(r.ctx, rs_download(r.ctx, RS_FIELD_T, T.data(), (int64_t)T.size()));
        const std::vector<double> Tp = pack_T(T, k);
        CHECK(r.ctx, rs_write_checkpoint((r.data + "/check_point_temperature.bin").c_str(), Tp.data(),
                                         (int64_t)Tp.size()));
    }
    int32_t it = 0;
    CHECK(r.ctx, rs_tip_column(r.ctx, &it));
    CHECK(r.ctx, rs_write_parameters((r.dir + "/parameters.txt").c_str(), &r.p, &k, it));
    if (r.state_file) CHECK(r.ctx, rs_save_state(r.ctx, (r.data + "/state.bin").c_str()));
}

// restart from the reference's files (if_load, fun.cu:584-616): state from
// parameters.txt, c and phi (psi := phi) and the coarse T from the
// checkpoints
void load_reference_files(Run& r) {
    rs_constants k;
    CHECK(r.ctx, rs_get_constants(r.ctx, &k));
    CHECK(r.ctx, rs_read_parameters((r.dir + "/parameters.txt").c_str(), &r.p, &k));
    const int64_t n = (int64_t)r.p.Nx * r.p.Ny;
    std::vector<double> c(n), phi(n);
    CHECK(r.ctx, rs_read_checkpoint((r.data + "/check_point_c.bin").c_str(), c.data(), n));
    CHECK(r.ctx, rs_read_checkpoint((r.data + "/check_point_phi.bin").c_str(), phi.data(), n));
    CHECK(r.ctx, rs_upload(r.ctx, RS_FIELD_C, c.data(), n));
    CHECK(r.ctx, rs_upload(r.ctx, RS_FIELD_PSI, phi.data(), n));
    if (r.p.ifTFC) {
        std::vector<double> Tp((size_t)k.NxT * k.NyT), T((size_t)k.NxT_all * k.NyT_all, 0.0);
        CHECK(r.ctx, rs_read_checkpoint((r.data + "/check_point_temperature.bin").c_str(), Tp.data(),
                                        (int64_t)Tp.size()));
        for (int j = 0; j < k.NyT; ++j)
            for (int i = 0; i < k.NxT; ++i) T[(size_t)j * k.NxT_all + i] = Tp[(size_t)j * k.NxT + i];
        CHECK(r.ctx, rs_upload(r.ctx, RS_FIELD_T, T.data(), (int64_t)T.size()));
    }
    k.if_begin_tip = 1;
    k.if_begin_field = 1;
    CHECK(r.ctx, rs_set_constants(r.ctx, &k));
}

}  // namespace

int main(int argc, char** argv) {
    for (int q = 1; q < argc; ++q) {
        const char* s = argv[q];
        while (*s == '-') ++s;
        const char* eq = strchr(s, '=');
        if (!eq) {
            fprintf(stderr, "usage: rsolid_run [key=value ...]  (bad argument %s)\n", argv[q]);
            return 2;
        }
        g_args[std::string(s, eq - s)] = eq + 1;
    }
    Run r;
    const std::string preset = arg("preset", "reference");
    int rc;
    if (preset == "config")
        rc = rs_params_config(&r.p, (int32_t)argl("N", 256), (int32_t)argl("tfc", 1), argd("alpha", 100.0));
    else
        rc = rs_params_reference(&r.p, (int32_t)argl("Nx_all", 768), (int32_t)argl("Ny_all", 832));
    if (rc) die(nullptr, rc, "parameter preset");
    // run.sh -D keys (fun.cu macros) at run time
    r.p.total_time = argd("total_time", r.p.total_time);
    r.p.Vp = argd("Vp", r.p.Vp);
    r.p.GG = argd("GG", r.p.GG);
    r.p.eps1 = argd("eps1", r.p.eps1);
    r.p.epk1 = argd("epk1", r.p.epk1);
    r.p.at_coef = argd("at_coef", r.p.at_coef);
    r.p.V_D = argd("V_D", r.p.V_D);
    r.p.k_part = argd("k_part", r.p.k_part);
    r.p.alpha0 = argd("alpha0", r.p.alpha0);  // fun.cu:129-131 grain orientation
    r.p.beta0 = argd("beta0", r.p.beta0);
    r.p.gamma0 = argd("gamma0", r.p.gamma0);
    const std::string prec = arg("precision", "fp64");
    r.p.precision = prec == "mixed"       ? RS_PREC_MIXED
                    : prec == "surrogate" ? RS_PREC_SURROGATE
                                          : RS_PREC_FP64;
    r.di = (int)argl("di", 1);
    r.dj = (int)argl("dj", 1);
    r.state_file = argl("state", 0) != 0;
    r.dir = arg("path_input", ".");
    r.data = r.dir + "/data";
    ensure_dir(r.dir);
    ensure_dir(r.data);
    const double run_time_h = argd("run_time", 1e6);

    CHECK(nullptr, rs_create(&r.p, &r.ctx));
    const std::string load_state = arg("load_state", "");
    if (!load_state.empty()) {
        CHECK(r.ctx, rs_load_state(r.ctx, load_state.c_str()));
    } else if (argl("if_load", 0)) {
        load_reference_files(r);
    } else {
        if (preset == "config")
            CHECK(r.ctx, rs_init_config(r.ctx, (uint64_t)argl("seed", 1)));
        else
            CHECK(r.ctx, rs_init_reference(r.ctx));
        rs_constants k0;
        CHECK(r.ctx, rs_get_constants(r.ctx, &k0));
        CHECK(r.ctx, rs_write_parameters((r.dir + "/parameters.txt").c_str(), &r.p, &k0, 0));
    }
    if (r.p.precision == RS_PREC_SURROGATE) {
        rs_surrogate_opts so;
        rs_surrogate_defaults(&so);
        double rep[RS_SURR_NREPORT];
        CHECK(r.ctx, rs_surrogate_train(r.ctx, &so, rep, RS_SURR_NREPORT));
        printf("surrogate trained in %.2f s, trusted bins %.3f\n", rep[7], rep[1]);
    }
    if (argl("pull_back", 0) > 0) {
        rs_constants kp;
        CHECK(r.ctx, rs_get_constants(r.ctx, &kp));
        kp.step_pull_back = (int32_t)argl("pull_back", 0);
        CHECK(r.ctx, rs_set_constants(r.ctx, &kp));
    }
    if (argl("finite_check", 0) > 0)
        CHECK(r.ctx, rs_set_finite_check(r.ctx, (int32_t)argl("finite_check", 0)));
    char dev[256] = "";
    rs_device_name(dev, sizeof dev);
    printf("rsolid_run on %s, %d x %d, precision %s\n", dev, r.p.Nx, r.p.Ny, prec.c_str());

    const auto t0 = std::chrono::steady_clock::now();
    const double tau0 = (r.p.S * r.p.W0) * (r.p.S * r.p.W0) / (r.p.Gamma * r.p.muk0);
    rs_constants k;
    CHECK(r.ctx, rs_get_constants(r.ctx, &k));
    const long initial = k.step;
    const std::string tip_path = r.dir + "/tip.txt";
    for (;;) {
        CHECK(r.ctx, rs_get_constants(r.ctx, &k));
        const long step = k.step;
        if (step >= k.total_step) break;
        // fun.cu:437-446: fields, tip, checkpoint
        if (step % k.step_fields == 0 || step == initial) {
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (k.if_begin_field == 1) {
                printf("(1)        (2)        (3)        (4)\n");
                printf("step       t[us]      t[s]       t[h]\n");
                k.if_begin_field = 0;
                CHECK(r.ctx, rs_set_constants(r.ctx, &k));
            }
            printf("%-10ld %-10.3f %-10.0f %-10.1f\n", step, step * k.dt * tau0 / 1000, el, el / 3600);
            fflush(stdout);
            write_fields(r, step, k);
        }
        if (step % k.step_tip == 0) {
            rs_tip_row t;
            const int header = k.if_begin_tip;
            rc = rs_tip(r.ctx, &t);
            if (rc != RS_OK && rc != RS_ERR_DOMAIN) die(r.ctx, rc, "rs_tip");
            CHECK(r.ctx, rs_write_tip_row(tip_path.c_str(), &t, header));
        }
        if (step % k.step_check_point == 0 && step != initial) write_checkpoint(r);
        // the pull-back of this step (fun.cu:448-463) runs inside rs_step;
        // its history records need the columns it shifts out
        std::vector<double> cols;
        int32_t i_offset = 0;
        if (step % k.step_pull_back == 0) {
            int32_t it = 0;
            CHECK(r.ctx, rs_tip_column(r.ctx, &it));
            i_offset = it - r.p.i_tip_target;
            if (i_offset > 0) {
                cols.resize((size_t)r.p.Ny * i_offset);
                CHECK(r.ctx, rs_download_columns(r.ctx, RS_FIELD_C, 0, i_offset, cols.data()));
            }
        }
        // advance to the next output or pull-back step
        long next = k.total_step;
        for (long cad : {(long)k.step_fields, (long)k.step_tip, (long)k.step_check_point,
                         (long)k.step_pull_back})
            next = std::min(next, (step / cad + 1) * cad);
        CHECK(r.ctx, rs_step(r.ctx, next - step));
        if (step % k.step_pull_back == 0) {
            rs_constants ka;
            CHECK(r.ctx, rs_get_constants(r.ctx, &ka));
            CHECK(r.ctx, rs_append_offset((r.data + "/offset.txt").c_str(), step, ka.i_offset,
                                          ka.i_offset_history, ka.step_pull_back));
            CHECK(r.ctx, rs_append_history((r.data + "/c_history_temp.txt").c_str(), cols.data(),
                                           std::max(i_offset, 0), r.p.Ny, ka.i_offset_history, r.di,
                                           r.dj, &r.i_ignore));
        }
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (el > run_time_h * 3600.0 - 1000.0) break;  // fun.cu:511-513 (wall time here)
    }
    write_checkpoint(r);  // fun.cu:516-517
    rs_destroy(r.ctx);
    return 0;
}
