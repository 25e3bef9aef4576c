// rs_noalloc.hip -- rs_solute_mass_noalloc: the solute-mass reduction of
// rs_solute_mass without a context and without any allocation, for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops). The per-row partial sums live in a file-scope device array,
// `scratch`; each GPU holds its own copy, looked up once per device.
#include <mutex>
#include <string>

#include <hip/hip_runtime.h>

#include "rs_host.hpp"

namespace {

constexpr int kMaxRows = 65536;  // Ny limit: 512 KiB of partial sums per GPU
constexpr int kMaxDevices = 64;

// per-row partial sums of every call on this GPU (the last call's contents
// stay; each call writes the rows it reads, so nothing needs clearing)
static __device__ __align__(256) double scratch[kMaxRows];

std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// device address of this GPU's scratch array (host code never passes the
// array's name, a host-side stand-in, to a kernel)
int scratch_ptr(double** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return RS_ERR_HIP;
    if (dev < 0 || dev >= kMaxDevices) return RS_ERR_HIP;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RS_ERR_HIP;
        g_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_ptr[dev];
    return RS_OK;
}

}  // namespace

extern "C" int rs_solute_mass_noalloc(const double* c, int32_t Nx, int32_t Ny, double* mass,
                                      void* stream) {
    if (!c || !mass) {
        rs::set_thread_error("rs_solute_mass_noalloc: null field or output pointer");
        return RS_ERR_ARG;
    }
    if (Nx < 3 || Ny < 3) {
        rs::set_thread_error("rs_solute_mass_noalloc: the grid needs Nx >= 3 and Ny >= 3");
        return RS_ERR_ARG;
    }
    if (Ny > kMaxRows) {
        rs::set_thread_error("rs_solute_mass_noalloc: Ny = " + std::to_string(Ny) +
                             " exceeds the scratch array's " + std::to_string(kMaxRows) +
                             " rows (use rs_solute_mass)");
        return RS_ERR_ARG;
    }
    double* rows = nullptr;
    const int rc = scratch_ptr(&rows);
    if (rc) {
        rs::set_thread_error("rs_solute_mass_noalloc: no device address for the scratch array");
        return rc;
    }
    const hipError_t e = rs::launch_mass(c, Nx, Ny, rows, mass, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        rs::set_thread_error(std::string("rs_solute_mass_noalloc: ") + hipGetErrorString(e));
        return RS_ERR_HIP;
    }
    return RS_OK;
}
