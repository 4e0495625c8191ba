// wenofd_noalloc.hip — calculateDeltaT of a caller-owned device state without
// any allocation (wenofd_delta_t_noalloc, include/wenofd_amd.h).
//
// The partial minima of the CFL reduction live in a file-scope device array
// (`scratch`, kNoallocParts doubles + kNoallocParts ints, 12 KiB per GPU for
// the life of the process) instead of buffers of a solver handle. Its address
// on each device is taken once, with that device current, by
// hipGetSymbolAddress; kernels receive that address, never the host-side
// symbol. Every call on a GPU shares that GPU's array: calls that overlap on
// one GPU race (serialise them on one stream). The kernels are the
// calculateDeltaT kernels of the solver handles (k_dt_partial, k_dt_final),
// which overwrite every partial slot, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "wenofd_amd.h"
#include "wenofd_host.hpp"
#include "wenofd_kernels.hpp"

constexpr int kNoallocParts = 1024;       // partial minima: one per reduction block
constexpr int kNoallocMaxPoints = 65536;  // grid points a side (the solver's own limit)
constexpr int kMaxDevices = 64;

// kNoallocParts partial minima (double), then their NaN flags (int).
static __device__ __align__(256) unsigned char scratch[kNoallocParts * (sizeof(double) + sizeof(int))];

namespace wenofd {
namespace {

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device (null: not taken yet)

// The device address of `scratch` on device dev (current), taken on first use.
int scratch_on(int dev, void** out) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_last_error(WENOFD_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return 0;
}

// dt = min(dt, t_left) on the device (src/WENOFD.cpp:434-435).
__global__ void k_clip_dt(double* out, double t_left) {
    if (threadIdx.x == 0 && t_left < out[0]) out[0] = t_left;
}

}  // namespace
}  // namespace wenofd

using namespace wenofd;

extern "C" int wenofd_delta_t_noalloc(const double* U, int nx, int ny, int pitch, double gamma, double dx, double dy,
                                      double cfl, double t_left, int fp_contract, double* out, void* stream) {
    if (!U || !out) return set_last_error(WENOFD_E_ARG, "null argument");
    if (nx < 5 || ny < 5 || nx > kNoallocMaxPoints || ny > kNoallocMaxPoints)
        return set_last_error(WENOFD_E_ARG, "nx and ny must be in [5, " + std::to_string(kNoallocMaxPoints) + "]");
    if (pitch < nx + 8) return set_last_error(WENOFD_E_ARG, "pitch must be at least nx + 8");
    if (!(dx > 0) || !(dy > 0) || !(gamma > 1) || !(cfl > 0))
        return set_last_error(WENOFD_E_ARG, "dx, dy and cfl must be positive and gamma above 1");
    if (fp_contract != kFpExact && fp_contract != kFpFused) return set_last_error(WENOFD_E_ARG, "fp_contract must be 0 or 1");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_last_error(WENOFD_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return set_last_error(WENOFD_E_HIP, "device index beyond the scratch table");
    void* sc = nullptr;
    int rc = scratch_on(dev, &sc);
    if (rc) return rc;
    double* part = static_cast<double*>(sc);
    int* part_nan = reinterpret_cast<int*>(part + kNoallocParts);
    Geom g{};
    g.nx = nx;
    g.ny = ny;
    g.pitch = pitch;
    g.plane = (long)pitch * (ny + 8);
    g.dx = dx;
    g.dy = dy;
    g.rdx = 1.0 / dx;
    g.rdy = 1.0 / dy;
    g.gamma = gamma;
    hipStream_t s = static_cast<hipStream_t>(stream);
    e = WENOFD_FP(fp_contract, launch_dt(U, g, cfl, part, part_nan, kNoallocParts, out, s));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_clip_dt, dim3(1), dim3(64), 0, s, out, t_left);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return set_last_error(WENOFD_E_HIP, std::string("delta_t_noalloc: ") + hipGetErrorString(e));
    return 0;
}
