// Entry points that allocate no device memory: their buffers come from a
// fixed array in the library's code object, one copy per GPU, whose device
// address is looked up once per device (hipGetSymbolAddress) and then passed
// to the kernels like any other buffer.  For callers that cannot call
// hipMalloc (graph capture, tight serving loops, harnesses that forbid it).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/qgpusim.h"
#include "qsim_internal.h"

namespace {

constexpr size_t kScratchBytes = QSIM_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

// Device-side scratch.  In host code this name is only a handle for the
// symbol lookup below, never a device address.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceScratch {
    std::mutex mu;         // one call at a time per device: the calls share the array
    void *ptr = nullptr;   // the array's address on this device (null until looked up)
};
DeviceScratch g_dev[kMaxDevices];

// The array's address on `device` (looked up on first use, with that device
// current, and kept).  Caller holds g_dev[device].mu.
int scratch_on(int device, void **out)
{
    DeviceScratch &d = g_dev[device];
    if (!d.ptr) {
        int prev = 0;
        hipError_t e = hipGetDevice(&prev);
        if (e != hipSuccess) return qsim::set_error(std::string("HIP error: ") + hipGetErrorString(e));
        if (prev != device && (e = hipSetDevice(device)) != hipSuccess)
            return qsim::set_error(std::string("HIP error: ") + hipGetErrorString(e));
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != device) hipSetDevice(prev);
        if (e != hipSuccess || !p)
            return qsim::set_error(std::string("sample_noalloc: scratch address lookup failed: ") +
                                   hipGetErrorString(e));
        d.ptr = p;
    }
    *out = d.ptr;
    return 0;
}

}  // namespace

extern "C" int qsim_sample_noalloc(qsim_state *st, uint64_t shots, uint64_t seed, uint64_t *indices_out)
{
    if (!st) return qsim::set_error("null state");
    const int device = qsim::state_device(st);
    if (device < 0 || device >= kMaxDevices) return qsim::set_error("sample_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_dev[device].mu);
    void *p = nullptr;
    if (scratch_on(device, &p)) return -1;
    return qsim::sample_impl(st, shots, seed, indices_out, p, kScratchBytes);
}
