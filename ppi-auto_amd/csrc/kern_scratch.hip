// kern_scratch.hip — the device work area of ppi_engine_evaluate_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc) evaluate through an entry point
// whose programs, fold programs, partials, bad-row records and results live in
// one statically allocated array per device.  Each GPU the library runs on holds
// kScratchBytes for the life of the process.  The array's host-side name is only a
// handle: its device address comes from hipGetSymbolAddress, taken once per device
// when an engine is created on it (scratch_setup) and used for every launch and
// copy afterwards (scratch_region).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine_internal.h"

namespace {

constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[ppi::kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

namespace ppi {

ppi_status scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device index beyond the scratch table (" + std::to_string(kMaxDevices) + " devices)");
        return PPI_E_ARG;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[device]) return PPI_OK;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) {
        set_error("scratch_setup: device " + std::to_string(device) + " is not current");
        return PPI_E_DEVICE;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return PPI_E_DEVICE;
    }
    g_ptr[device] = p;
    return PPI_OK;
}

void* scratch_region(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    return g_ptr[device];
}

}  // namespace ppi
