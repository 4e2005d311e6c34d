// The no-allocation entry point, tsg_batch_run_noalloc: a variant of
// tsg_batch_run for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Every device buffer of the run (sweep partials, per-chunk running sums,
// first-tree counts, combine and finals buffers, the batch's packed host
// arrays, the results) is carved from `scratch`, a file-scope device array
// of TSG_SCRATCH_BYTES that each GPU the library runs on holds until the
// process ends.  In host code `scratch` names a host-side stand-in, not
// device memory: its device address on each GPU is looked up once, with
// that GPU current, by hipGetSymbolAddress, and only that address is passed
// on.  Every call on a GPU shares that GPU's array and finds the previous
// call's contents there: the kernels write every word they later read
// (the first-tree counts are cleared by a fill kernel), so nothing needs to
// start at zero.  Calls that overlap on one GPU (from several host threads)
// race on the array; the caller serialises them.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>

#include "tsg_common.h"

struct tsg_batch;
extern "C" int tsg_batch_run_fixed(tsg_batch *b, void *scratch, size_t cap);

// 64 MiB: enough for a batch of ~50 outputs over 1000 windows on a
// tree sequence of 1e5 samples cut into ~17000 chunks
#define TSG_SCRATCH_BYTES ((size_t) 64 << 20)
#define TSG_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[TSG_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mutex;
void *g_scratch_ptr[TSG_MAX_DEVICES];  // device address of scratch, per GPU

// The address of this GPU's copy of scratch (looked up on the first call on
// the device, with the device current).
int scratch_of(int dev, void **out) {
    if (dev < 0 || dev >= TSG_MAX_DEVICES) return TSG_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || ptr == nullptr) {
            (void) hipGetLastError();
            return TSG_ERR_DEVICE;
        }
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" {

int tsg_batch_run_noalloc(tsg_batch *batch) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void) hipGetLastError();
        return TSG_ERR_NO_DEVICE;
    }
    void *ptr = nullptr;
    const int ret = scratch_of(dev, &ptr);
    if (ret != 0) return ret;
    return tsg_batch_run_fixed(batch, ptr, TSG_SCRATCH_BYTES);
}

}  // extern "C"
