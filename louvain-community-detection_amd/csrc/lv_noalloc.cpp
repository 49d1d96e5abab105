// lv_noalloc.cpp -- lv_run_noalloc: the whole Louvain run with no device
// allocation, for callers that cannot allocate.
//
// Every device buffer the run needs is carved from `scratch`, a file-scope
// device array of kScratchBytes.  Each GPU the library runs on holds its own
// copy of the array until the process ends.  Its device address on a GPU is
// taken once, with that GPU current, by hipGetSymbolAddress, and only that
// address is handed to kernels, copies and fills: in host code `scratch`
// names a host-side stand-in, not device memory.
//
// Calls on one GPU share its array and run one at a time (a mutex per device
// in this process; each call synchronises its stream before it returns).  A
// caller that bypasses this library to touch the array would race with it.
// The array keeps the previous call's contents; the run clears everything it
// needs empty (tables, assignment) before it starts.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/louvain_hip.h"
#include "lv_host.h"

namespace {

constexpr size_t kScratchBytes = 128ull << 20;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex device_lock[kMaxDevices];
void *scratch_ptr[kMaxDevices];  // device address of `scratch` per GPU, once known
alignas(64) unsigned char host_counters[kMaxDevices][lv_host::kHostCountersBytes];

// The device address of `scratch` on the current device (`device`); the
// caller holds device_lock[device].
int scratch_address(int device, char **out) {
    if (!scratch_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return LV_ERR_HIP;
        }
        scratch_ptr[device] = p;
    }
    *out = static_cast<char *>(scratch_ptr[device]);
    return LV_OK;
}

}  // namespace

extern "C" int lv_run_noalloc(int device, uint32_t n, const uint32_t *row, const uint32_t *col,
                              const float *w, float min_gain, void *stream,
                              uint32_t *assignment, lv_result *res) {
    if (!row || !assignment || !res) return LV_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) {
        (void)hipGetLastError();
        return LV_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= ndev) return LV_ERR_NO_DEVICE;
    if (device >= kMaxDevices) return LV_ERR_ARG;
    std::lock_guard<std::mutex> hold(device_lock[device]);
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return LV_ERR_HIP;
    }
    char *base = nullptr;
    const int rc = scratch_address(device, &base);
    if (rc) return rc;
    return lv_host::run_fixed(device, base, kScratchBytes, host_counters[device], n, row, col, w,
                              min_gain, stream, assignment, res);
}
