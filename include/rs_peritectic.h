/*
 * rs_peritectic.h -- C ABI of the 3D three-phase peritectic Ti-Ag stepper on
 * one GPU (one z-slab of the reference's multi-GPU model).
 *
 * The reference (multi/peritectic_multiGPU/TiAg.cu) advances two phase
 * fields phi1, phi2 (phi3 = 1 - phi1 - phi2) and a composition c on an
 * Nx x Ny x Nz grid split into z-slabs of devNz planes, one per GPU
 * (TiAg.cu:19-35). Its interface is the set of kernels main() launches each
 * step (TiAg.cu:288-313): computeMu + boundaryMu{X,Y,Z} (TiAg.cu:446-483),
 * then diffuse + boundary{X,Y,Z} (TiAg.cu:486-579), then a pointer swap;
 * initial() sets the two-nuclei start (TiAg.cu:411-444) and
 * save/load_break_point, data2File2D write the files (TiAg.cu:630-760).
 * With one GPU the slab is the whole grid (Nz = devNz, TiAg.cu:25-27) and
 * its z faces are the model's no-flux boundaries.
 *
 * This header replaces that surface: one call per step runs computeMu,
 * diffuse and all six boundary kernels fused (mu is never stored; it is a
 * pointwise function of phi1, phi2, c). Conventions as in rsolid.h: int
 * status codes (RS_OK, RS_ERR_*), no exceptions, host arrays never
 * retained, layout pos(i,j,k) = Nx*Ny*k + Nx*j + i with ghost cells
 * (TiAg.cu:66).
 */
#ifndef RS_PERITECTIC_H
#define RS_PERITECTIC_H

#include <stdint.h>

#include "rsolid.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Model constants: the TiAg.cu macros (TiAg.cu:19-62) as run-time values. */
typedef struct tp_params {
    int32_t Nx, Ny, Nz;        /* grid with ghosts (TiAg.cu:23-25); one GPU:
                                  Nz = devNz (TiAg.cu:27)                   */
    int32_t zcut_from_top;     /* initial(): zcut = Nz - 100 (TiAg.cu:418)   */
    double DT;                 /* time step (TiAg.cu:19)                     */
    double DXin, DYin, DZin;   /* inverse grid spacings (TiAg.cu:20-22)      */
    double a1, a2, a3, b;      /* multi-well coefficients (TiAg.cu:36-39)    */
    double lambda;             /* coupling (TiAg.cu:41)                      */
    double alpha;              /* solute mobility (TiAg.cu:42)               */
    double Da, Db;             /* relative mobility in phases 2, 3 (:43-44)  */
    double A1, A2;             /* TiAg.cu:52-53                              */
    double B11, B21, B12, B22; /* TiAg.cu:54-57                              */
    double T;                  /* reduced temperature (TiAg.cu:50)           */
    /* initial(): spherical caps of base radius R and contact angle theta
     * sitting on the plane zcut, centred at (cap_x[q], cap_y[q])
     * (TiAg.cu:411-444); ncaps <= 4 */
    double R, theta;
    int32_t ncaps, reserved;
    double cap_x[4], cap_y[4];
} tp_params;

typedef struct tp_ctx tp_ctx;

enum {
    TP_FIELD_PHI1 = 0,
    TP_FIELD_PHI2 = 1,
    TP_FIELD_C = 2,
    TP_FIELD_MU = 3 /* download only: computeMu + boundaryMu of the current
                       fields (TiAg.cu:446-483)                             */
};

/* sizeof(tp_params), for binding-side layout checks (no counterpart in the
 * reference, which has no ABI). */
int tp_abi_info(int32_t* size_of_params);

/* The TiAg.cu constants on an Nx x Ny x Nz grid (the reference runs
 * 512 x 512 x 506 over 4 GPUs, 512 x 512 x 128 planes per GPU). */
int tp_params_reference(tp_params* p, int32_t Nx, int32_t Ny, int32_t Nz);
/* Derived compositions of the model (TiAg.cu:58-62): out[0] = ca,
 * out[1] = cb, out[2] = cla, out[3] = clb, out[4] = B1, out[5] = B2. */
int tp_derived(const tp_params* p, double* out, int32_t n);

int tp_create(const tp_params* p, tp_ctx** out);
int tp_destroy(tp_ctx* ctx);
const char* tp_last_error(const tp_ctx* ctx); /* never NULL */

/* initial() (TiAg.cu:411-444): phase 3 below zcut, phase 2 in the caps,
 * phase 1 elsewhere, c at the matching composition; step := 0. */
int tp_init_reference(tp_ctx* ctx);
/* Nx*Ny*Nz values; TP_FIELD_MU cannot be uploaded (the stepper derives it). */
int tp_upload(tp_ctx* ctx, int32_t field, const double* host, int64_t count);
int tp_download(tp_ctx* ctx, int32_t field, double* host, int64_t count);
/* nsteps passes of the loop body TiAg.cu:288-345 on one GPU. */
int tp_step(tp_ctx* ctx, int64_t nsteps);
int tp_get_step(tp_ctx* ctx, int64_t* step);
int tp_set_step(tp_ctx* ctx, int64_t step);
/* The reference's progress figures (TiAg.cu:229-262): out[0] = mean of c
 * over the interior cells, out[1] = largest |c| over the interior cells of
 * plane k_check (its stop criterion is out[1] > 0.1 at plane 20). FP64,
 * deterministic reductions. */
int tp_report(tp_ctx* ctx, int32_t k_check, double* out, int32_t n);
int tp_sync(tp_ctx* ctx);
/* HIP-event timing of the fused step kernel over every step since the
 * last call with reset != 0: out[0] average ms, out[1] launches. */
int tp_timing(tp_ctx* ctx, int32_t reset, double* out, int32_t n);

/* ---- files (host only) --------------------------------------------------
 * data2File2D (TiAg.cu:630-665): "%.4f\t" per value, one line per z plane;
 * plane 0 = xz cross-section at y = index, 1 = yz at x = index. */
int tp_write_cross_section(const char* path, const double* data, int32_t Nx, int32_t Ny,
                           int32_t Nz, int32_t plane, int32_t index);
/* save_break_point / load_break_point (TiAg.cu:667-760): the 10-line VTK
 * structured-points header, then Nx*Ny*Nz native FP64 values. */
int tp_write_break_point(const char* path, const double* data, int32_t Nx, int32_t Ny, int32_t Nz);
int tp_read_break_point(const char* path, double* data, int32_t Nx, int32_t Ny, int32_t Nz);

#ifdef __cplusplus
}
#endif
#endif /* RS_PERITECTIC_H */
