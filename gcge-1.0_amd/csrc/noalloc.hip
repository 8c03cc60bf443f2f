// noalloc.hip -- the column dot products <X_j, Y_j> without any allocation,
// for callers that cannot allocate device or pinned memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The kernels' partial sums go to a file-scope device array (scratch below)
// of a fixed size instead of the library's growing scratch buffer, and the
// results are copied on the backend stream straight into the caller's host
// array.  Every GPU the library runs on holds its own copy of the array
// until the process ends; its address is looked up once per device when
// the library's context is set up (noalloc_setup, called by ensure_ctx)
// and only that address is handed to kernels.  The array keeps the last
// call's contents; every slot a call reads is written by the same call
// first, so nothing needs clearing.  Calls that overlap on one GPU (from
// several host threads) race on the array.
#include <hip/hip_runtime.h>

#include "hip_common.h"

namespace gh {
bool dots_with(int prec, int64_t n, const void *x, int64_t ldx, const void *y, int64_t ldy, int k, double *partial,
               double *out);
int dots_blocks(int64_t n);
}  // namespace gh

// columns one call may take, and the partial-sum slots they need: k_dots
// writes k * dots_blocks(n) <= 128 * 256 partials, k_reduce_rows k sums
constexpr int kNoAllocMaxCols = 128;
constexpr size_t kNoAllocScratch = (size_t)kNoAllocMaxCols * 256 + kNoAllocMaxCols;  // doubles (263 KiB)

static __device__ __align__(256) double scratch[kNoAllocScratch];

namespace {
constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {nullptr};
}  // namespace

namespace gh {

// Once per device, with that device current: the device address of this
// GPU's copy of the array (the array's name is a host-side stand-in).
bool noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return true;  // the entry point then reports the missing set-up
    if (g_scratch_ptr[device]) return true;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (e != hipSuccess) {
        latch((int)e, "noalloc set-up: hipGetSymbolAddress: %s", hipGetErrorString(e));
        return false;
    }
    g_scratch_ptr[device] = static_cast<double *>(p);
    return true;
}

}  // namespace gh

using namespace gh;

// Variant of the MultiVecInnerProdDiag slot (out[j] = <X_{start[0]+j},
// Y_{start[1]+j}>, j < end[0] - start[0], FP64 accumulation, host out): the
// same kernels and summation order, so the same bits, with no allocation.
extern "C" GCGE_INT GCGE_HIP_MultiVecInnerProdDiagNoAlloc(void **x, void **y, GCGE_DOUBLE *out, GCGE_INT *start,
                                                          GCGE_INT *end)
{
    MV *mx = as_mv(x), *my = as_mv(y);
    if (!mx || !my || !out || !start || !end) {
        latch(200, "MultiVecInnerProdDiagNoAlloc: bad arguments");
        return 200;
    }
    const int k = end[0] - start[0];
    if (k < 0 || start[0] < 0 || end[0] > mx->ncols || start[1] < 0 || start[1] + k > my->ncols) {
        latch(201, "MultiVecInnerProdDiagNoAlloc: column range [%d,%d) / [%d,%d) outside %d / %d columns", start[0],
              end[0], start[1], start[1] + k, mx->ncols, my->ncols);
        return 201;
    }
    if (k > kNoAllocMaxCols) {
        latch(202, "MultiVecInnerProdDiagNoAlloc: %d columns exceed the limit of %d per call", k, kNoAllocMaxCols);
        return 202;
    }
    if (mx->n != my->n || mx->prec != my->prec) {
        latch(203, "MultiVecInnerProdDiagNoAlloc: multivectors differ in rows or precision");
        return 203;
    }
    if (k == 0) return 0;
    if (failed()) return GCGE_HIP_GetLastError(nullptr, 0);
    int dev = -1;
    if (!ctx().init || hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices || !g_scratch_ptr[dev]) {
        latch(204, "MultiVecInnerProdDiagNoAlloc: the library context is not set up on this device");
        return 204;
    }
    if ((size_t)k * dots_blocks(mx->n) + k > kNoAllocScratch) {
        latch(205, "MultiVecInnerProdDiagNoAlloc: partial sums exceed the scratch array");
        return 205;
    }
    const size_t es = mx->prec == GCGE_HIP_FP64 ? sizeof(double) : sizeof(float);
    const char *xb = static_cast<const char *>(mx->data) + (size_t)start[0] * mx->ld * es;
    const char *yb = static_cast<const char *>(my->data) + (size_t)start[1] * my->ld * es;
    if (!dots_with(mx->prec, mx->n, xb, mx->ld, yb, my->ld, k, g_scratch_ptr[dev], out))
        return GCGE_HIP_GetLastError(nullptr, 0);
    return 0;
}
