// axs3d_scratch.hip -- the device-resident cardinal scratch of the
// allocation-free stiffness entry point (axs3d_compute_stiffness_static).
//
// The cardinal path keeps the strain and stress rows of a batch of 3D
// elements in a scratch buffer between its gradient, stress and
// quadrature kernels.  axs3d_compute_stiffness uses the buffer the context
// allocated at axs3d_finalize; the static entry point uses this array
// instead, a file-scope device array of STATIC_SCRATCH_BYTES that every
// GPU the library runs on holds until the process ends (each GPU has its
// own copy at its own address).  The array's name is a host-side
// stand-in, not a device address: host code only ever passes the address
// that hipGetSymbolAddress returned for the current device.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "axs3d_domain.hpp"

namespace axs3d {

static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of this device's copy of the array; looked up once per
// device (the caller has made `device` current) and cached.
void *static_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices)
        throw Error("Domain::Domain || Device index out of range for the "
                    "static scratch.");
    if (g_scratch_ptr[device] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            throw Error(std::string("Domain::Domain || hipGetSymbolAddress "
                                    "of the static scratch failed: ") +
                        hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

}  // namespace axs3d
