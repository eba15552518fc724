// The fixed device scratch of gt_lp_matvec_batch_noalloc: the L_p product's transposed row spectra (T1) and their
// products with the kernel (T2) for a whole batch, in one file-scope device array instead of the stack's workspace,
// so that the entry point allocates nothing. Every call on a device shares that device's copy of the array; it keeps
// the last call's contents (no pass reads an entry before the same call wrote it, so nothing is cleared).
#include <mutex>

#include "stack.hpp"

static __device__ __align__(256) unsigned char scratch[GT_LP_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;

std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};               // the array's device address on each device
hipError_t g_status[MAX_DEVICES] = {};      // what the look-up returned there

}  // namespace

namespace gt {

// The array's address on `device`, looked up once per device by the library's init call (gt_ctx_create) with that
// device current. A failed look-up is kept and reported by noalloc_scratch.
void noalloc_scratch_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_status[device] = e;
    if (e == hipSuccess) g_ptr[device] = p;
}

// The device address of the scratch on `device` (GT_ERR_DEVICE when its look-up failed or never ran)
void* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0 || device >= MAX_DEVICES || !g_ptr[device]) {
        const hipError_t e = device >= 0 && device < MAX_DEVICES ? g_status[device] : hipErrorInvalidDevice;
        fail(GT_ERR_DEVICE, std::string("no-allocation scratch unavailable on this device: ") +
                                (e == hipSuccess ? "not set up" : hipGetErrorString(e)));
    }
    return g_ptr[device];
}

}  // namespace gt
