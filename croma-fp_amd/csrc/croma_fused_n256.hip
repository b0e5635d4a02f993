// croma_fused_n256.hip -- instantiations of the fused persistent kernel for
// n = 256 (one translation unit per grid size so they compile in parallel).
#include "croma_device.h"

#include <stdio.h>

template <int N, int R, int WPB, bool PROT>
static int launch_fused_t(int nblocks, CkGrid grid, const CkSlot *slots, const CkTask *tasks, int ntasks,
                          int *counter, int *err, hipStream_t st)
{
    hipLaunchKernelGGL((k_fused<N, R, WPB, PROT>), dim3(nblocks), dim3(WPB * 64), 0, st, grid, slots, tasks,
                       ntasks, counter, err);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int N, int R, bool PROT>
static int launch_fused_w(int wpb, int nblocks, CkGrid grid, const CkSlot *slots, const CkTask *tasks,
                          int ntasks, int *counter, int *err, hipStream_t st)
{
    switch (wpb) {
    case 1: return launch_fused_t<N, R, 1, PROT>(nblocks, grid, slots, tasks, ntasks, counter, err, st);
    case 2: return launch_fused_t<N, R, 2, PROT>(nblocks, grid, slots, tasks, ntasks, counter, err, st);
    case 4: return launch_fused_t<N, R, 4, PROT>(nblocks, grid, slots, tasks, ntasks, counter, err, st);
    default: return -1;
    }
}

extern "C" int ck_launch_fused_n256(int r, int wpb, int nblocks, int protons, CkGrid grid, const CkSlot *slots,
                                   const CkTask *tasks, int ntasks, int *counter, int *err, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (r == 4)
        return protons ? launch_fused_w<256, 4, true>(wpb, nblocks, grid, slots, tasks, ntasks, counter, err, st)
                       : launch_fused_w<256, 4, false>(wpb, nblocks, grid, slots, tasks, ntasks, counter, err, st);
    if (r == 8)
        return protons ? launch_fused_w<256, 8, true>(wpb, nblocks, grid, slots, tasks, ntasks, counter, err, st)
                       : launch_fused_w<256, 8, false>(wpb, nblocks, grid, slots, tasks, ntasks, counter, err, st);
    if (r == 16)
        return protons ? launch_fused_w<256, 16, true>(wpb, nblocks, grid, slots, tasks, ntasks, counter, err, st)
                       : launch_fused_w<256, 16, false>(wpb, nblocks, grid, slots, tasks, ntasks, counter, err, st);
    return -1;
}
