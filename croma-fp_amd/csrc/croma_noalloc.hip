// croma_noalloc.hip -- the batched Chang-Cooper solve for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// It runs the kernel of croma_hip_solve_cc_batch (k_solve_cc_batch: x <-
// dt Q + x, then the Thomas sweep of fp_shared_core.c:226-257 per system) on
// caller-owned device arrays.  The kernel's only scratch -- the modified
// upper diagonal c' of every system and the pivot-failure word -- lives in
// the file-scope device array `scratch` below: one copy per GPU, held until
// the process ends, whose address is looked up once per device with
// hipGetSymbolAddress.  Nothing here calls hipMalloc, hipFree or an
// allocating library routine.
//
// Layout of `scratch`: bytes [0, 256) hold the pivot-failure word (cleared by
// a fill kernel at the start of every call), c' starts at byte 256.  Calls
// that overlap on one GPU (two host threads, or two streams) share the array
// and race.
#include <hip/hip_runtime.h>

#include <pthread.h>
#include <stddef.h>

#include "croma_kernels.h"

#include "../../include/croma_hip.h"

extern "C" void ch_set_error(const char *fmt, ...);

#define NOALLOC_HEAD_DOUBLES 32 /* 256 bytes: the pivot-failure word */
#define NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) double scratch[NOALLOC_HEAD_DOUBLES + CROMA_HIP_NOALLOC_MAX_CELLS];

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static double *g_scratch[NOALLOC_MAX_DEVICES]; /* device address of `scratch`, per device */

/* Address of this device's copy of `scratch` (looked up on the first call
 * made with that device current). */
static double *scratch_for_device(int dev)
{
    double *p;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return 0;
    pthread_mutex_lock(&g_lock);
    p = g_scratch[dev];
    if (!p) {
        void *addr = 0;
        const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
        if (e == hipSuccess && addr) {
            p = (double *)addr;
            g_scratch[dev] = p;
        } else {
            ch_set_error("solve_cc_batch_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        }
    }
    pthread_mutex_unlock(&g_lock);
    return p;
}

extern "C" int croma_hip_solve_cc_batch_noalloc(int nsys, int nrow, double dt, const double *a, const double *b,
                                                const double *c, const double *source, const double *state_init,
                                                double *state_out, void *stream)
{
    int dev = -1, err = 0;
    double *base;
    hipError_t e;
    if (nsys <= 0 || nrow <= 0 || !a || !b || !c || !source || !state_init || !state_out) {
        ch_set_error("solve_cc_batch_noalloc: invalid argument");
        return CROMA_HIP_EINVAL;
    }
    if ((size_t)nsys * (size_t)nrow > (size_t)CROMA_HIP_NOALLOC_MAX_CELLS) {
        ch_set_error("solve_cc_batch_noalloc: nsys * nrow = %zu exceeds the scratch limit of %d cells",
                     (size_t)nsys * (size_t)nrow, CROMA_HIP_NOALLOC_MAX_CELLS);
        return CROMA_HIP_EINVAL;
    }
    if (hipGetDevice(&dev) != hipSuccess) {
        ch_set_error("solve_cc_batch_noalloc: no current device");
        return CROMA_HIP_EDEVICE;
    }
    base = scratch_for_device(dev);
    if (!base) return CROMA_HIP_EDEVICE;
    int *d_err = (int *)base;
    double *d_cp = base + NOALLOC_HEAD_DOUBLES;
    if (ck_fill_i32(d_err, 0, 1, stream) ||
        ck_solve_cc_batch(nsys, nrow, dt, a, b, c, source, state_init, state_out, d_cp, d_err, stream)) {
        ch_set_error("solve_cc_batch_noalloc: %s", ck_last_error());
        return CROMA_HIP_EDEVICE;
    }
    e = hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) {
        ch_set_error("solve_cc_batch_noalloc: %s", hipGetErrorString(e));
        return CROMA_HIP_EDEVICE;
    }
    return err ? CROMA_HIP_EPIVOT : 0;
}
