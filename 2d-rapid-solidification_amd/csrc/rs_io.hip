// rs_io.hip -- the reference's on-disk formats (host code): %g field text,
// tip.txt rows, offset.txt and c_history_temp.txt records, raw FP64
// checkpoints, parameters.txt, plus an exact binary state file and a
// column-block download for the pull-back history.
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "rs_host.hpp"

using namespace rs;

namespace {

int io_fail(int code) { return code; }

struct File {
    FILE* f = nullptr;
    File(const char* path, const char* mode) : f(fopen(path, mode)) {}
    ~File() {
        if (f) fclose(f);
    }
};

constexpr uint64_t kStateMagic = 0x3253544153525352ULL;  // "RSRSATS2" (rs_params with grain angles)

}  // namespace

extern "C" {

// save_field (fun.cu:1768-1790): rows j = 0, dj, ... of columns i = 0, di, ...
// as "%g\t", one line per row
int rs_write_field_txt(const char* path, const double* data, int32_t Nx, int32_t Ny, int32_t di,
                       int32_t dj) {
    if (!path || !data || Nx < 1 || Ny < 1 || di < 1 || dj < 1) return RS_ERR_ARG;
    File fp(path, "w");
    if (!fp.f) return io_fail(RS_ERR_DOMAIN);
    for (int j = 0; j < Ny; j += dj) {
        for (int i = 0; i < Nx; i += di) fprintf(fp.f, "%g\t", data[(int64_t)j * Nx + i]);
        fprintf(fp.f, "\n");
    }
    return ferror(fp.f) ? RS_ERR_DOMAIN : RS_OK;
}

// save_tip (src/main.cu:2671-2709): appends one row; header first when asked
int rs_write_tip_row(const char* path, const rs_tip_row* t, int32_t header) {
    if (!path || !t) return RS_ERR_ARG;
    File fp(path, "a");
    if (!fp.f) return RS_ERR_DOMAIN;
    if (header) {
        fprintf(fp.f, "(1)        (2)        (3)        (4)        (5)        (6)        (7)        "
                      "(8)        (9)        (10)       (11)\n");
        fprintf(fp.f, "step       t[us]      \xCE\x94tip       Ttip[K]    Vtip[m/s]  Vp[m/s]    "
                      "tip.R[nm]  tip.i      tip.j      tip.x      tip.y\n");
    }
    fprintf(fp.f, "%-10lld %-10g %-10g %-10g %-10g %-10g %-10g %-10d %-10d %-10g %-10g \n",
            (long long)t->step, t->t_us, t->delta, t->T_tip, t->V_tip, t->Vp, t->R, t->i, t->j,
            t->x, t->y);
    return ferror(fp.f) ? RS_ERR_DOMAIN : RS_OK;
}

// offset.txt record of a pull-back (fun.cu:710-716)
int rs_append_offset(const char* path, int64_t step, int32_t i_offset, int32_t i_offset_history,
                     int32_t step_pull_back) {
    if (!path) return RS_ERR_ARG;
    File fp(path, "a");
    if (!fp.f) return RS_ERR_DOMAIN;
    fprintf(fp.f, "%lld\t%d\t%d\t%d\n", (long long)step, i_offset, i_offset_history, step_pull_back);
    return RS_OK;
}

// c_history_temp.txt records of the columns a pull-back shifts out
// (fun.cu:719-740): cols is Ny x ncols row-major (rs_download_columns),
// column q is the field column q; columns with (q + i_offset_history) % di
// != 0 are skipped, and *i_ignore counts columns still to skip. Each written
// column is one line of rows j = 0, dj, ...
int rs_append_history(const char* path, const double* cols, int32_t ncols, int32_t Ny,
                      int32_t i_offset_history, int32_t di, int32_t dj, int32_t* i_ignore) {
    if (!path || (!cols && ncols > 0) || Ny < 1 || di < 1 || dj < 1 || !i_ignore) return RS_ERR_ARG;
    File fp(path, "a");
    if (!fp.f) return RS_ERR_DOMAIN;
    *i_ignore -= ncols;
    for (int q = 0; q < ncols; ++q) {
        if ((q + i_offset_history) % di != 0) continue;
        if (*i_ignore <= 0) {
            for (int j = 0; j < Ny; j += dj) fprintf(fp.f, "%g\t", cols[(int64_t)j * ncols + q]);
            fprintf(fp.f, "\n");
        }
        --*i_ignore;
        if (*i_ignore < 0) *i_ignore = 0;
    }
    return RS_OK;
}

// raw FP64 checkpoints (save_check_point / load_check_point, fun.cu:1848-1916)
int rs_write_checkpoint(const char* path, const double* data, int64_t n) {
    if (!path || !data || n < 0) return RS_ERR_ARG;
    File fp(path, "wb");
    if (!fp.f) return RS_ERR_DOMAIN;
    return fwrite(data, sizeof(double), (size_t)n, fp.f) == (size_t)n ? RS_OK : RS_ERR_DOMAIN;
}

int rs_read_checkpoint(const char* path, double* data, int64_t n) {
    if (!path || !data || n < 0) return RS_ERR_ARG;
    File fp(path, "rb");
    if (!fp.f) return RS_ERR_DOMAIN;
    return fread(data, sizeof(double), (size_t)n, fp.f) == (size_t)n ? RS_OK : RS_ERR_DOMAIN;
}

// parameters.txt (save_parameters, fun.cu:2060-2202): the same lines in the
// same order, so the reference's loader reads it; quantities this build has
// no parameter for (dilute-alloy d0/ke/b/lambda, molar weight, density,
// the second grain's angle, eps2, random seed) are written as 0.
int rs_write_parameters(const char* path, const rs_params* p, const rs_constants* k, int32_t i_tip) {
    if (!path || !p || !k) return RS_ERR_ARG;
    File out(path, "w");
    FILE* fp = out.f;
    if (!fp) return RS_ERR_DOMAIN;
    (void)i_tip;
    const Derived d = derive_constants(*p, k->dt);
    const double tau0 = d.tau0, dt = k->dt;
    const double SW0 = p->S * p->W0;
    fprintf(fp, "----------------------------------------- Dynamically changing variables\n");
    fprintf(fp, "%-13lld step\n", (long long)k->step);
    fprintf(fp, "%-13d i_offset_history, accumulated i offset\n", k->i_offset_history);
    fprintf(fp, "%-13g l_offset_Vp, unit: [W]; offset due to pulling velocity\n", k->l_offset_Vp);
    fprintf(fp, "%-13g delta_initial, initial dimensionless temperature\n", k->delta);
    fprintf(fp, "\n");
    fprintf(fp, "-----------------------------------------\n");
    fprintf(fp, "%-13d Nx_all\n", p->Nx);
    fprintf(fp, "%-13d Ny_all\n", p->Ny);
    fprintf(fp, "%-13g dx => %g nm\n", p->dx, p->dx * SW0);
    fprintf(fp, "%-13g W0 [nm], physical interface thickness\n", p->W0);
    fprintf(fp, "%-13g S, scale factor\n", (double)p->S);
    fprintf(fp, "%-13g W=S*W0, nm, effective interface thickness\n", SW0);
    fprintf(fp, "%-13d i_tip_target => %g um\n", p->i_tip_target, p->i_tip_target * p->dx * SW0 * 1e-3);
    fprintf(fp, "\n");
    fprintf(fp, "%-13g dt => %g ns\n", dt, dt * tau0);
    fprintf(fp, "%-13g tau0 [ns]; characteristic time\n", tau0);
    fprintf(fp, "%-13lld total steps      => %g us\n", (long long)k->total_step, p->total_time * 1e-3);
    fprintf(fp, "%-13d step_fields      => %g us\n", k->step_fields, k->step_fields * tau0 * dt * 1e-3);
    fprintf(fp, "%-13d step_check_point => %g us\n", k->step_check_point,
            k->step_check_point * tau0 * dt * 1e-3);
    fprintf(fp, "%-13d step_tip         => %g ns\n", k->step_tip, k->step_tip * tau0 * dt);
    fprintf(fp, "%-13d step_pull_back   => %g ns\n", k->step_pull_back, k->step_pull_back * tau0 * dt);
    fprintf(fp, "\n");
    fprintf(fp, "%-13g Gamma [K*nm], Gibbs-Thomson coefficient\n", p->Gamma);
    fprintf(fp, "%-13g muk0 [nm/ns/K], interface kinetic coefficient\n", p->muk0);
    fprintf(fp, "%-13g nm^2/ns => %g m^2/s, Dl, diffusivity in the liquid\n", p->Dl, p->Dl * 1e-9);
    fprintf(fp, "%-13g K/nm => %g K/m, G, thermal gradient\n", p->GG, p->GG * 1e9);
    fprintf(fp, "%-13g Vp [m/s], pulling velocity \n", p->Vp);
    fprintf(fp, "\n");
    fprintf(fp, "%-13g Tmelt [K], melting temperature \n", p->Tmelt);
    fprintf(fp, "%-13g cinf, nominal composition \n", p->cinf);
    fprintf(fp, "%-13g deltaT0 [K], |me|*cinf for dilute alloys \n", p->deltaT0);
    fprintf(fp, "%-13g d0 [nm] = Gamma/deltaT0, capillary length, only for dilute alloys\n", 0.0);
    fprintf(fp, "%-13g ke, partition coefficient, only for dilute alloys\n", 0.0);
    fprintf(fp, "%-13g b=ln(ke)/2, only for dilute alloys\n", 0.0);
    fprintf(fp, "%-13g lambda, only for dilute alloys\n", 0.0);
    fprintf(fp, "\n");
    fprintf(fp, "%-13g gamma0 [degrees], orientation angle\n", p->gamma0);
    fprintf(fp, "%-13g gamma1 [degrees], orientation angle\n", 0.0);
    fprintf(fp, "%-13g eps1, anisotropy parameter of the interfacial free energy \n", p->eps1);
    fprintf(fp, "%-13g eps2, anisotropy parameter of the interfacial free energy \n", 0.0);
    fprintf(fp, "%-13g epk1, anisotropy parameter of the kinetic coefficient \n", p->epk1);
    fprintf(fp, "%-13ld random_seed\n", 0L);
    fprintf(fp, "%-13s path_input\n", "./");
    fprintf(fp, "\n");
    fprintf(fp, "%-13g latentHeat [J/mol], latent heat\n", p->latentHeat);
    fprintf(fp, "%-13g cp [J/mol/K], specific heat\n", p->cp);
    fprintf(fp, "%-13g Lcp = latentHeat/cp [K]\n", d.Lcp);
    fprintf(fp, "%-13g molar_weight [g/mol] \n", 0.0);
    fprintf(fp, "%-13g density [g/cm^3]\n", 0.0);
    fprintf(fp, "%-13g num_moles [mol/m^3], number of moles per m^3\n", p->num_moles);
    fprintf(fp, "%-13g latentHeat2 [J/m^3], latent heat\n", p->latentHeat * p->num_moles);
    fprintf(fp, "%-13g surface_tension [J/m^2]\n", p->Gamma * 1e-9 * p->latentHeat * p->num_moles / p->Tmelt);
    fprintf(fp, "%-13g hh = surface_tension/(a1*S*W0*1e-9) [J/m^3]\n", d.hh);
    fprintf(fp, "%-13g h0 = RR*Tmelt/v0 [J/m^3]\n", d.h0);
    fprintf(fp, "\n");
    if (p->ifTFC) {
        const double dxT = p->dx * p->ratio_dxT_dx;
        fprintf(fp, "----------------------------------------- Variables for temperature field calculation (TFC)\n");
        fprintf(fp, "-------------------- Dynamically changing variables\n");
        fprintf(fp, "%-13d i_left_end_initial, The i index of the left border of PF in the T field.\n",
                k->i_left_end_initial);
        fprintf(fp, "%-13d i_left_end, The i index of the left border of PF in the T field.\n", k->i_left_end);
        fprintf(fp, "\n");
        fprintf(fp, "%-13g dxT => %g nm\n", dxT, dxT * SW0);
        fprintf(fp, "%-13d dxT/dx, larger dx for the temperature field\n", p->ratio_dxT_dx);
        fprintf(fp, "%-13d NxT => %g um\n", k->NxT, k->NxT * dxT / 1000);
        fprintf(fp, "%-13d NyT => %g um\n", k->NyT, k->NyT * dxT / 1000);
        fprintf(fp, "%-13d NxT_all => %g um, some lattices not used\n", k->NxT_all, k->NxT_all * dxT / 1000);
        fprintf(fp, "%-13d NyT_all => %g um, some lattices not used\n", k->NyT_all, k->NyT_all * dxT / 1000);
        fprintf(fp, "%-13g sample length [nm] => %g um\n", k->length_sample, k->length_sample / 1000);
        fprintf(fp, "\n");
        fprintf(fp, "%-13g dtT => %g ns\n", k->dtT, k->dtT * tau0);
        fprintf(fp, "%-13d dt/dtT, smaller dt for the temperature field\n", k->ratio_dt_dtT);
        fprintf(fp, "\n");
        fprintf(fp, "%-13g D_T [nm^2/ns], thermal diffusivity\n", p->DT);
        fprintf(fp, "%-13g Tcold [K]\n", k->Tcold);
        fprintf(fp, "%-13g Thot [K]\n", k->Thot);
    }
    return ferror(fp) ? RS_ERR_DOMAIN : RS_OK;
}

// load_parameters (fun.cu:2204-2388): the values the reference's restart
// takes from parameters.txt, keyed by the description that follows each
// value (the first numeric token of the line). k holds the derived
// constants on entry (rs_derive); the restart fields are overwritten.
int rs_read_parameters(const char* path, const rs_params* p, rs_constants* k) {
    if (!path || !p || !k) return RS_ERR_ARG;
    File in(path, "r");
    if (!in.f) return RS_ERR_DOMAIN;
    char line[1024];
    bool have_step = false;
    while (fgets(line, sizeof line, in.f)) {
        char* end = nullptr;
        errno = 0;
        const double v = strtod(line, &end);
        if (end == line || errno) continue;
        while (*end == ' ') ++end;
        std::string d(end);
        auto starts = [&](const char* s) { return d.compare(0, strlen(s), s) == 0; };
        if (starts("step\n") || d == "step") {
            k->initial_step = (int64_t)llround(v);
            have_step = true;
        } else if (starts("i_offset_history")) {
            k->i_offset_history_initial = (int32_t)lround(v);
        } else if (starts("l_offset_Vp")) {
            k->l_offset_Vp_initial = v;
        } else if (starts("delta_initial")) {
            k->delta_initial = v;
        } else if (starts("dt =>")) {
            k->dt = v;
        } else if (starts("i_left_end_initial")) {
            k->i_left_end_initial = (int32_t)lround(v);
        } else if (starts("i_left_end,")) {
            k->i_left_end = (int32_t)lround(v);
        } else if (starts("dtT =>")) {
            k->dtT = v;
        } else if (starts("dt/dtT")) {
            k->ratio_dt_dtT = (int32_t)lround(v);
        } else if (starts("Tcold")) {
            k->Tcold = v;
        } else if (starts("Thot")) {
            k->Thot = v;
        } else if (starts("sample length")) {
            k->length_sample = v;
        }
    }
    if (!have_step) return RS_ERR_DOMAIN;
    k->step = k->initial_step;
    k->i_offset_history = k->i_offset_history_initial;
    k->l_offset_Vp = k->l_offset_Vp_initial;
    k->delta = k->delta_initial;
    k->i_left_end_interp = k->i_left_end;
    (void)p;
    return RS_OK;
}

// Columns [i0, i0 + ncols) of a field into host[Ny x ncols] (row-major): the
// columns a pull-back is about to shift out, without the whole field
int rs_download_columns(rs_ctx* ctx, int32_t field, int32_t i0, int32_t ncols, double* host) {
    if (!ctx || !host || ncols < 0 || i0 < 0 || i0 + ncols > ctx->p.Nx) return RS_ERR_ARG;
    if (ncols == 0) return RS_OK;
    const int Nx = ctx->p.Nx, Ny = ctx->p.Ny;
    hipError_t e = hipSuccess;
    if (field == RS_FIELD_C || (field == RS_FIELD_PSI && !ctx->mixed)) {
        const double* src = (field == RS_FIELD_C ? ctx->c[ctx->cur] : ctx->psi[ctx->cur]) + i0;
        e = hipMemcpy2DAsync(host, sizeof(double) * ncols, src, sizeof(double) * Nx,
                             sizeof(double) * ncols, Ny, hipMemcpyDeviceToHost, ctx->stream);
    } else if (field == RS_FIELD_PSI) {
        std::vector<float> tmp((size_t)Ny * ncols);
        e = hipMemcpy2DAsync(tmp.data(), sizeof(float) * ncols, ctx->psi32[ctx->cur] + i0,
                             sizeof(float) * Nx, sizeof(float) * ncols, Ny, hipMemcpyDeviceToHost,
                             ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        for (size_t q = 0; q < tmp.size(); ++q) host[q] = (double)tmp[q];
        return e == hipSuccess ? RS_OK : RS_ERR_HIP;
    } else {
        return RS_ERR_ARG;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? RS_OK : RS_ERR_HIP;
}

// Exact state file: magic, the parameter block, the run state, the parities
// and the device fields in their stored precision (psi, c, coarse T with its
// spare row, and the FP32 T copy on the mixed paths). rs_load_state restores
// it into a context whose parameters differ at most in the run length and
// output counts: the fields and the evolving run state are taken from the
// file, the cadences from the context's own parameters, so the run
// continues bit for bit.
int rs_save_state(rs_ctx* ctx, const char* path) {
    if (!ctx || !path) return RS_ERR_ARG;
    File out(path, "wb");
    if (!out.f) return RS_ERR_DOMAIN;
    const int64_t n = ctx->n, nT = ctx->nT + ctx->k.NxT_all;
    std::vector<char> buf;
    auto put = [&](const void* q, size_t b) {
        const char* c = static_cast<const char*>(q);
        buf.insert(buf.end(), c, c + b);
    };
    put(&kStateMagic, 8);
    put(&ctx->p, sizeof(rs_params));
    put(&ctx->k, sizeof(rs_constants));
    int32_t par[2] = {ctx->cur, ctx->tcur};
    put(par, sizeof par);
    const size_t ps = ctx->mixed ? sizeof(float) : sizeof(double);
    std::vector<char> field((size_t)std::max<int64_t>(n, nT) * 8);
    auto dev = [&](const void* src, size_t bytes) -> hipError_t {
        hipError_t e = hipMemcpyAsync(field.data(), src, bytes, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) put(field.data(), bytes);
        return e;
    };
    hipError_t e = dev(ctx->mixed ? (const void*)ctx->psi32[ctx->cur] : (const void*)ctx->psi[ctx->cur],
                       ps * n);
    if (e == hipSuccess) e = dev(ctx->c[ctx->cur], sizeof(double) * n);
    if (e == hipSuccess) e = dev(ctx->T[ctx->tcur], sizeof(double) * nT);
    if (e == hipSuccess && ctx->mixed) e = dev(ctx->T32[ctx->tcur], sizeof(float) * nT);
    if (e != hipSuccess) return RS_ERR_HIP;
    return fwrite(buf.data(), 1, buf.size(), out.f) == buf.size() ? RS_OK : RS_ERR_DOMAIN;
}

int rs_load_state(rs_ctx* ctx, const char* path) {
    if (!ctx || !path) return RS_ERR_ARG;
    File in(path, "rb");
    if (!in.f) return RS_ERR_DOMAIN;
    uint64_t magic = 0;
    rs_params p;
    rs_constants k;
    int32_t par[2];
    if (fread(&magic, 8, 1, in.f) != 1 || magic != kStateMagic || fread(&p, sizeof p, 1, in.f) != 1 ||
        fread(&k, sizeof k, 1, in.f) != 1 || fread(par, sizeof par, 1, in.f) != 1)
        return RS_ERR_DOMAIN;
    p.total_time = ctx->p.total_time;
    p.num_fields_output = ctx->p.num_fields_output;
    p.num_tip_output = ctx->p.num_tip_output;
    p.num_check_point_output = ctx->p.num_check_point_output;
    if (memcmp(&p, &ctx->p, sizeof p) != 0) return RS_ERR_ARG;  // another run's state
    const int64_t n = ctx->n, nT = ctx->nT + ctx->k.NxT_all;
    const size_t ps = ctx->mixed ? sizeof(float) : sizeof(double);
    std::vector<char> field((size_t)std::max<int64_t>(n, nT) * 8);
    // both parity buffers receive the state; the run continues from parity 0
    auto dev = [&](void* dst0, void* dst1, size_t bytes) -> int {
        if (fread(field.data(), 1, bytes, in.f) != bytes) return RS_ERR_DOMAIN;
        hipError_t e = hipMemcpyAsync(dst0, field.data(), bytes, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dst1, field.data(), bytes, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        return e == hipSuccess ? RS_OK : RS_ERR_HIP;
    };
    int rc = ctx->mixed ? dev(ctx->psi32[0], ctx->psi32[1], ps * n) : dev(ctx->psi[0], ctx->psi[1], ps * n);
    if (!rc) rc = dev(ctx->c[0], ctx->c[1], sizeof(double) * n);
    if (!rc) rc = dev(ctx->T[0], ctx->T[1], sizeof(double) * nT);
    if (!rc && ctx->mixed) rc = dev(ctx->T32[0], ctx->T32[1], sizeof(float) * nT);
    if (rc) return rc;
    rs_constants kk;
    rs_derive(&ctx->p, &kk);
    kk.step = k.step;
    kk.initial_step = k.initial_step;
    kk.if_begin_tip = k.if_begin_tip;
    kk.if_begin_field = k.if_begin_field;
    kk.i_offset = k.i_offset;
    kk.i_offset_history_initial = k.i_offset_history_initial;
    kk.i_offset_history = k.i_offset_history;
    kk.i_ignore = k.i_ignore;
    kk.x_tip_history = k.x_tip_history;
    kk.delta_initial = k.delta_initial;
    kk.delta = k.delta;
    kk.l_offset_Vp_initial = k.l_offset_Vp_initial;
    kk.l_offset_Vp = k.l_offset_Vp;
    kk.i_left_end_initial = k.i_left_end_initial;
    kk.i_left_end = k.i_left_end;
    kk.i_left_end_interp = k.i_left_end_interp;
    ctx->k = kk;
    ctx->cur = 0;
    ctx->tcur = 0;
    // the loaded composition field: the solute ledger starts over
    rc = solute_ledger_reset(ctx);
    if (rc) return rc;
    return hipStreamSynchronize(ctx->stream) == hipSuccess ? RS_OK : RS_ERR_HIP;
}

}  // extern "C"
