// rs_params.hip -- parameter presets and the derived-parameter setup
// (host code of the library).
#include <math.h>
#include <string.h>

#include "rs_host.hpp"

namespace rs {

// Newton solve of the liquid/bcc common tangent at temperature T
// (equal chemical potentials and equal tangent intercepts). cl, cs carry the
// initial guesses in and the compositions (Ti fraction) out.
int common_tangent(const rs_params& p, double T, double* cl, double* cs) {
    const RK rk = rk_of(p);
    double a = *cl, b = *cs;
    for (int it = 0; it < 200; ++it) {
        const Thermo tl = thermo_fp64(rk, T, a);
        const Thermo ts = thermo_fp64(rk, T, b);
        const double fl = Gl_fp64(rk, T, a), fs = Gs_fp64(rk, T, b);
        const double F1 = tl.muL - ts.muS;
        const double F2 = fl - fs - tl.muL * (a - b);
        const double dl = p.RR * T / (a * (1.0 - a)) - 2.0 * p.L0_liq + p.L1_liq * (-6.0 + 12.0 * a);
        const double ds = p.RR * T / (b * (1.0 - b)) - 2.0 * p.L0_bcc + p.L1_bcc * (-6.0 + 12.0 * b);
        const double J11 = dl, J12 = -ds, J21 = -dl * (a - b), J22 = tl.muL - ts.muS;
        const double det = J11 * J22 - J12 * J21;
        if (det == 0.0 || !isfinite(det)) return -1;
        const double da = (F1 * J22 - J12 * F2) / det;
        const double db = (J11 * F2 - J21 * F1) / det;
        double lam = 1.0;
        while ((a - lam * da <= 1e-9 || a - lam * da >= 1 - 1e-9 || b - lam * db <= 1e-9 ||
                b - lam * db >= 1 - 1e-9) &&
               lam > 1e-6)
            lam *= 0.5;
        a -= lam * da;
        b -= lam * db;
        if (fabs(da) < 1e-14 && fabs(db) < 1e-14) {
            *cl = a;
            *cs = b;
            return 0;
        }
    }
    *cl = a;
    *cs = b;
    return 1;
}

// liquidus temperature of a liquid of Ti fraction c_liq (bisection on the
// common-tangent liquid composition, which falls as T rises)
int liquidus(const rs_params& p, double c_liq, double* TL, double* c_sol) {
    double lo = 1942.0, hi = 2700.0;
    double cl = c_liq, cs = c_liq - 0.02;
    for (int it = 0; it < 80; ++it) {
        const double mid = 0.5 * (lo + hi);
        double a = cl, b = cs;
        if (common_tangent(p, mid, &a, &b) != 0) return -1;
        if (a > c_liq) lo = mid; else hi = mid;
        cl = a;
        cs = b;
    }
    *TL = 0.5 * (lo + hi);
    double a = c_liq, b = cs;
    if (common_tangent(p, *TL, &a, &b) != 0) return -1;
    *c_sol = b;
    return 0;
}

}  // namespace rs

using namespace rs;

extern "C" int rs_params_reference(rs_params* p, int32_t Nx, int32_t Ny) {
    if (!p || Nx < 272 || Ny < 12) return RS_ERR_ARG;
    memset(p, 0, sizeof(*p));
    p->Nx = Nx;
    p->Ny = Ny;
    p->i_tip_target = Nx - 268;          // fun.cu:41
    p->S = 5;                            // fun.cu:63
    p->ratio_dxT_dx = 10;                // fun.cu:109
    p->ifTFC = 1;                        // fun.cu:25
    p->if_Vp_ramp = 1;                   // fun.cu:30
    p->dt_cap = 1;                       // fun.cu:1989-1991
    p->num_fields_output = 120;          // fun.cu:50-52
    p->num_tip_output = 120;
    p->num_check_point_output = 120;
    p->precision = RS_PREC_FP64;
    p->W0 = 0.25;                        // fun.cu:64
    p->Gamma = 199.0;                    // fun.cu:65
    p->muk0 = 0.5795;                    // fun.cu:68
    p->Tmelt = 2750.0;                   // fun.cu:72
    p->deltaT0 = 2750.0 - 2161.59;       // fun.cu:73
    p->Dl = 0.9;                         // fun.cu:74
    p->DT = 1.1e4;                       // fun.cu:111
    p->GG = 0.0236;                      // run.sh:13
    p->Vp = 0.0842;                      // run.sh:12
    p->Vp_ramp_rate = 0.0107587;         // fun.cu:1977
    p->eps1 = 0.012;                     // run.sh:21
    p->epk1 = 0.1;                       // run.sh:23
    p->cinf = 0.632514105;               // fun.cu:62
    p->latentHeat = 22003.0;             // fun.cu:87
    p->num_moles = 90183.49558101478;    // fun.cu:92
    p->v0 = 1.1088503429118772e-05;      // fun.cu:91
    p->cp = 40.0;                        // fun.cu:94
    p->RR = 8.3145;                      // fun.cu:103
    p->dx = 0.6;                         // fun.cu:171-173 (SS == 5)
    p->AA = 12.0;
    p->TCOLD = 2000.0;                   // fun.cu:108
    p->L0_bcc = 12315.25;                // fun.cu:1668
    p->L0_liq = 5144.29;                 // fun.cu:1669
    p->L1_bcc = 0.0;
    p->L1_liq = 0.0;
    p->dt_coef = 0.6;                    // fun.cu:1989
    p->dtT_coef = 0.9;                   // fun.cu:2029
    p->length_sample_um = 50.0;          // fun.cu:2037
    p->lap_w_axis = 2.0 / 3.0;           // fun.cu:1051
    p->lap_w_diag = 1.0 / 3.0;           // fun.cu:1052
    p->q_solid = 0.0;
    p->total_time = 120000.0;            // run.sh:18 (ns)
    p->delta_initial = -1.0;             // fun.cu:2013
    p->phi_interface = 0.0;              // fun.cu:126
    p->c_liq0 = p->cinf;
    p->c_sol0 = 0.515285917;             // fun.cu:62 comment (solid partner of cinf)
    p->at_coef = 0.0;                    // no antitrapping in the reference
    p->V_D = 0.0;
    p->k_part = p->c_sol0 / p->cinf;
    return RS_OK;
}

// Ti-10at%Nb configuration: S = 1, dx = 0.8 W, nondimensional D_l = 1,
// D_s = 1e-4 D_l, 5-point Laplacian, eps4 = 0.02, G and material data from
// src/constants.hpp, V from run.sh, planar front at x0 = 25 cells.
extern "C" int rs_params_config(rs_params* p, int32_t N, int32_t tfc, double alpha_ratio) {
    if (!p || N < 64) return RS_ERR_ARG;
    memset(p, 0, sizeof(*p));
    p->Nx = N;
    p->Ny = N;
    p->i_tip_target = 25;
    p->S = 1;
    p->ratio_dxT_dx = 10;
    p->ifTFC = tfc ? 1 : 0;
    p->if_Vp_ramp = 0;
    p->dt_cap = 1;
    p->num_fields_output = 120;
    p->num_tip_output = 120;
    p->num_check_point_output = 120;
    p->precision = RS_PREC_FP64;
    p->W0 = 0.25;                       // constants.hpp:34
    p->Gamma = 199.0;                   // constants.hpp:35
    p->muk0 = 0.57;                     // constants.hpp:39
    p->Dl = p->Gamma * p->muk0;         // D_l = 1 in units of W^2/tau0
    p->DT = (alpha_ratio > 0 ? alpha_ratio : 100.0) * p->Dl;
    p->GG = 0.01;                       // constants.hpp:10
    p->Vp = 0.0842;                     // run.sh:12
    p->Vp_ramp_rate = 0.0;
    p->eps1 = 0.02;
    p->epk1 = 0.1;                      // run.sh:23
    p->latentHeat = 22003.0;            // constants.hpp:61
    p->num_moles = 93353.43470588748;   // constants.hpp:63
    p->v0 = 1.17e-5;                    // constants.hpp:76
    p->cp = 40.0;                       // constants.hpp:66
    p->RR = 8.3145;                     // constants.hpp:74
    p->dx = 0.8;                        // fun.cu:165-167 (SS == 1)
    p->AA = 1.0;
    p->L0_bcc = 12315.25;
    p->L0_liq = 5144.29;
    p->dt_coef = 0.9;
    p->dtT_coef = 0.9;
    p->lap_w_axis = 1.0;
    p->lap_w_diag = 0.0;
    p->q_solid = 1e-4;
    p->delta_initial = -1.0;
    p->phi_interface = 0.0;
    // Ti-10at%Nb liquid: liquidus and solidus partner from the CALPHAD
    // common tangent; undercooling from src/constants.hpp:45
    p->c_liq0 = 0.90;
    double TL = 0, cs = 0;
    if (liquidus(*p, p->c_liq0, &TL, &cs) != 0) return RS_ERR_DOMAIN;
    p->Tmelt = TL;
    p->c_sol0 = cs;
    p->at_coef = 0.0;                   // antitrapping / trapping extension off
    p->V_D = 0.0;
    p->k_part = cs / p->c_liq0;
    p->cinf = p->c_liq0;
    p->deltaT0 = 2161.59 - 2118.35;
    // coarse T sample twice the window, window centred in it
    const double cell = p->dx * p->S * p->W0 * p->ratio_dxT_dx;
    p->length_sample_um = 2.0 * (N - 2) * p->dx * p->S * p->W0 / 1000.0;
    const int ile_target = (N - 2) / 2;
    p->TCOLD = (p->Tmelt - p->deltaT0) - p->GG * p->dx * p->S * p->W0 * (ile_target + p->i_tip_target);
    (void)cell;
    // total_time: 2000 steps of dt (cadences derive from it)
    rs_constants k;
    p->total_time = 1.0;
    rs_derive(p, &k);
    const double sw = p->S * p->W0;
    const double tau0 = (sw * sw) / (p->Gamma * p->muk0);
    p->total_time = (2000.0 + 0.5) * tau0 * k.dt;
    return RS_OK;
}

// set_parameters (fun.cu:1984-2058)
extern "C" int rs_derive(const rs_params* p, rs_constants* k) {
    if (!p || !k) return RS_ERR_ARG;
    memset(k, 0, sizeof(*k));
    const double sw = p->S * p->W0;
    const double tau0 = (sw * sw) / (p->Gamma * p->muk0);
    const double DD = p->Dl * tau0 / (sw * sw);
    double dt = p->dt_coef * p->dx * p->dx / (4.0 * DD);
    if (p->dt_cap && dt > p->dt_coef * p->dx * p->dx / 4.0) dt = p->dt_coef * p->dx * p->dx / 4.0;
    k->dt = dt;
    k->x_tip_history = p->i_tip_target * p->dx;
    k->if_begin_tip = 1;
    k->if_begin_field = 1;
    k->total_step = (int64_t)(p->total_time / (tau0 * dt));
    k->step_fields = (int32_t)(p->total_time / p->num_fields_output / (tau0 * dt));
    k->step_tip = (int32_t)(p->total_time / p->num_tip_output / (tau0 * dt));
    k->step_check_point = (int32_t)(p->total_time / p->num_check_point_output / (tau0 * dt));
    const double vpb = p->if_Vp_ramp ? 1.0 : p->Vp;   // fun.cu:2007-2012
    k->step_pull_back = (int32_t)(p->S * p->W0 * p->dx / (vpb * dt * tau0));
    if (k->step_pull_back < 1) k->step_pull_back = 1;
    if (k->step_fields < 1) k->step_fields = 1;
    if (k->step_tip < 1) k->step_tip = 1;
    if (k->step_check_point < 1) k->step_check_point = 1;
    k->delta_initial = p->delta_initial;
    k->delta = k->delta_initial;
    const int r = p->ratio_dxT_dx;
    const double swr = p->S * p->W0 * r;
    const double DDT = p->DT * tau0 / (swr * swr);
    const double dtT = p->dtT_coef * p->dx * p->dx / (4.0 * DDT);
    k->ratio_dt_dtT = 1;
    if (dtT < k->dt) k->ratio_dt_dtT = (int32_t)ceil(k->dt / dtT);
    k->dtT = k->dt / k->ratio_dt_dtT;
    const double cell = p->dx * p->S * p->W0 * r;
    k->length_sample = ceil(p->length_sample_um * 1000 / cell) * cell;
    k->NxT = (int32_t)(k->length_sample / cell + 2);
    k->NxT_all = (int32_t)(ceil(1.0 * k->NxT / 64) * 64);
    k->NyT = ((p->Ny - 2) / r) + 2;
    k->NyT_all = (int32_t)(ceil(1.0 * k->NyT / 4) * 4);
    k->Tcold = p->TCOLD;
    k->Thot = k->Tcold + p->GG * k->length_sample;
    k->i_left_end_initial =
        (int32_t)((p->Tmelt - p->deltaT0 - k->Tcold) / (p->GG * p->dx * p->S * p->W0) - p->i_tip_target);
    k->i_left_end = k->i_left_end_initial;
    k->i_left_end_interp = k->i_left_end;
    return RS_OK;
}
