// croma_nsub.hip -- the step before the solve, on the device: substep counts
// of every tracer and snapshot of a window, the cost order of the queue and
// the segment rows of the split secondary contraction.
//
// Reference: tracer_fp_prepare_nsub_loss_endpoints / compute_nsubsteps_core
// (tracer_fp_nsub.c:94-282) for nsub and n_on; the bucketing by cost
// (tracer_fp_bucket.c:208-262) becomes one counting sort over 256 buckets per
// octave of the task cost (descending), and the coefficient segments of
// tracer_fp_setup.c:73-95 give each tracer its rows of the split GEMM.
//
// Three launches:
//   k_nsub_count  one thread per tracer: nsub, n_on for every slot, the cost
//                 bucket (histogram), the tracer's segment rows (block scan)
//   k_nsub_scan   one block: exclusive scans of the block row totals and of
//                 the histogram in descending bucket order
//   k_nsub_place  one thread per tracer: segment offsets and the task slot
// Both per-tracer kernels count in a block-local LDS histogram and touch the
// global one once per non-empty bucket of the block: most tracers of a
// workload share a few buckets, and one global atomic per tracer on them
// serialises.  Within one cost bucket the task order follows the atomics;
// results do not depend on it (each tracer's update is independent of its
// queue position).
#include <hip/hip_runtime.h>

#include "croma_device.h"

#define NS_BLOCK 256

__global__ __launch_bounds__(NS_BLOCK) void k_nsub_count(CkGrid g, const CkSlot *__restrict__ slots, int count,
                                                         int ntr, int protons, int split, int *__restrict__ key,
                                                         int *__restrict__ segloc, int *__restrict__ blk,
                                                         int *__restrict__ hist,
                                                         unsigned long long *__restrict__ cells)
{
    __shared__ int s_wsum[NS_BLOCK / 64];
    __shared__ int s_hist[CK_NSUB_HIST];
    const int i = blockIdx.x * NS_BLOCK + threadIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int b = threadIdx.x; b < CK_NSUB_HIST; b += NS_BLOCK) s_hist[b] = 0;
    __syncthreads();
    int rows = 0;
    unsigned long long cs = 0;
    if (i < ntr) {
        double cost = 0.0;
        for (int q = 0; q < count; ++q) {
            const CkSlot &S = slots[q];
            int ns, non;
            d_nsub_one(g, S.prm, S.dt_snap, S.ric_lo, S.ric_hi, S.bgc[CK_BG_NGAS][i], S.bgc[CK_BG_KBT][i],
                       S.bgc[CK_BG_B][i], S.bgc[CK_BG_DIVV][i], S.bgc[CK_BG_LTURB][i], S.bgc[CK_BG_DV][i], ns, non);
            const_cast<int *>(S.nsub)[i] = ns;
            const_cast<int *>(S.n_on)[i] = non;
            const int nseg = d_choose_segments(ns, S.allow_interp, S.prm.coeff_interp_min_steps,
                                               S.prm.coeff_interp_max_segments);
            cost += (double)ns * (protons ? 2 : 1) + 3.0 * nseg;
            rows += nseg;
            cs += (unsigned long long)ns;
        }
        const int k = d_cost_bucket(cost);
        key[i] = k;
        atomicAdd(&s_hist[k], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < CK_NSUB_HIST; b += NS_BLOCK) {
        const int c = s_hist[b];
        if (c) atomicAdd(&hist[b], c);
    }
    // sum nsub: one atomic per wave
    for (int o = 32; o > 0; o >>= 1) cs += __shfl_xor(cs, o, 64);
    if (lane == 0 && cs) atomicAdd(cells, cs);
    if (!split) return;
    // block-exclusive prefix of the segment rows (wave scan, then wave totals)
    int incl = rows;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) s_wsum[wv] = incl;
    __syncthreads();
    int wbase = 0, btot = 0;
    for (int w = 0; w < NS_BLOCK / 64; ++w) {
        if (w < wv) wbase += s_wsum[w];
        btot += s_wsum[w];
    }
    if (i < ntr) segloc[i] = wbase + incl - rows;
    if (threadIdx.x == 0) blk[blockIdx.x] = btot;
}

// Exclusive scan of v[0..len) in place by one block (chunks of blockDim.x);
// desc scans from the last entry down.  Returns the total through *total.
__device__ void ns_block_scan(int *v, int len, bool desc, int *total)
{
    __shared__ int s_w[16];
    __shared__ int s_carry;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, nw = blockDim.x >> 6;
    if (t == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < len; base += blockDim.x) {
        const int j = base + t;
        const int idx = desc ? len - 1 - j : j;
        const int x = j < len ? v[idx] : 0;
        int incl = x;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        if (lane == 63) s_w[wv] = incl;
        __syncthreads();
        int wbase = 0, ctot = 0;
        for (int w = 0; w < nw; ++w) {
            if (w < wv) wbase += s_w[w];
            ctot += s_w[w];
        }
        const int carry = s_carry;
        if (j < len) v[idx] = carry + wbase + incl - x;
        __syncthreads();
        if (t == 0) s_carry = carry + ctot;
        __syncthreads();
    }
    if (t == 0 && total) *total = s_carry;
}

__global__ __launch_bounds__(1024) void k_nsub_scan(int nblk, int *__restrict__ blk, int *__restrict__ hist,
                                                   int split)
{
    if (split) ns_block_scan(blk, nblk, false, blk + nblk);
    ns_block_scan(hist, CK_NSUB_HIST, true, nullptr);
}

__global__ __launch_bounds__(NS_BLOCK) void k_nsub_place(const CkSlot *__restrict__ slots, int count, int ntr,
                                                         int split, const int *__restrict__ key,
                                                         const int *__restrict__ segloc, const int *__restrict__ blk,
                                                         int *__restrict__ hist, CkTask *__restrict__ tasks)
{
    __shared__ int s_hist[CK_NSUB_HIST];
    const int i = blockIdx.x * NS_BLOCK + threadIdx.x;
    for (int b = threadIdx.x; b < CK_NSUB_HIST; b += NS_BLOCK) s_hist[b] = 0;
    __syncthreads();
    // rank of the tracer among the block's tracers of its bucket
    const int k = (i < ntr) ? key[i] : 0;
    const int local = (i < ntr) ? atomicAdd(&s_hist[k], 1) : 0;
    __syncthreads();
    // the block's range in each non-empty bucket
    for (int b = threadIdx.x; b < CK_NSUB_HIST; b += NS_BLOCK) {
        const int c = s_hist[b];
        if (c) s_hist[b] = atomicAdd(&hist[b], c);
    }
    __syncthreads();
    if (i >= ntr) return;
    if (split) {
        int off = blk[blockIdx.x] + segloc[i];
        for (int q = 0; q < count; ++q) {
            const CkSlot &S = slots[q];
            const_cast<int *>(S.seg_off)[i] = off;
            off += d_choose_segments(S.nsub[i], S.allow_interp, S.prm.coeff_interp_min_steps,
                                     S.prm.coeff_interp_max_segments);
        }
    }
    const int pos = s_hist[k] + local;
    CkTask t;
    t.slot = 0;
    t.tracer = i;
    t.nslot = count;
    tasks[pos] = t;
}

extern "C" int ck_nsub_order(CkGrid g, const CkSlot *d_slot, int count, int ntr, int protons, int split,
                             CkTask *tasks, int *key, int *segloc, int *blk, int *hist, unsigned long long *cells,
                             void *stream)
{
    if (ntr <= 0 || count <= 0) return -1;
    hipStream_t st = (hipStream_t)stream;
    const int nblk = (ntr + NS_BLOCK - 1) / NS_BLOCK;
    if (ck_fill_i32(hist, 0, CK_NSUB_HIST, stream)) return -1;
    hipLaunchKernelGGL(k_nsub_count, dim3(nblk), dim3(NS_BLOCK), 0, st, g, d_slot, count, ntr, protons, split, key,
                       segloc, blk, hist, cells);
    hipLaunchKernelGGL(k_nsub_scan, dim3(1), dim3(1024), 0, st, nblk, blk, hist, split);
    hipLaunchKernelGGL(k_nsub_place, dim3(nblk), dim3(NS_BLOCK), 0, st, d_slot, count, ntr, split, key, segloc, blk,
                       hist, tasks);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ __launch_bounds__(NS_BLOCK) void k_nsub_arrays(CkGrid g, CkParams prm, int ntr, double dt_snap,
                                                          double ric_lo, double ric_hi,
                                                          const double *__restrict__ n_gas,
                                                          const double *__restrict__ kbt,
                                                          const double *__restrict__ b_field,
                                                          const double *__restrict__ divv,
                                                          const double *__restrict__ l_turb,
                                                          const double *__restrict__ dv, int *__restrict__ nsub,
                                                          int *__restrict__ n_on)
{
    const int i = blockIdx.x * NS_BLOCK + threadIdx.x;
    if (i >= ntr) return;
    int ns, non;
    d_nsub_one(g, prm, dt_snap, ric_lo, ric_hi, n_gas[i], kbt[i], b_field[i], divv[i], l_turb[i], dv[i], ns, non);
    nsub[i] = ns;
    n_on[i] = non;
}

extern "C" int ck_nsub_arrays(CkGrid g, CkParams prm, int ntr, double dt_snap, double ric_lo, double ric_hi,
                              const double *n_gas, const double *kbt, const double *b_field, const double *divv,
                              const double *l_turb, const double *dv, int *nsub, int *n_on, void *stream)
{
    if (ntr <= 0) return 0;
    hipLaunchKernelGGL(k_nsub_arrays, dim3((ntr + NS_BLOCK - 1) / NS_BLOCK), dim3(NS_BLOCK), 0,
                       (hipStream_t)stream, g, prm, ntr, dt_snap, ric_lo, ric_hi, n_gas, kbt, b_field, divv, l_turb,
                       dv, nsub, n_on);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
