// croma_fused_n512.hip -- fused persistent kernel instances for n = 512
// (rows per lane: 8 16; 1 or 2 waves per SIMD; all launch modes).
#include "croma_fused_launch.h"

extern "C" int ck_launch_fused_n512(int r, int occ, int nblocks, int mode, CkGrid grid, const CkSlot *slots,
                                   const CkTask *tasks, int ntasks, int *counter, int *err, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    switch (r) {
    case 8: return ck_launch_r<512, 8>(occ, mode, nblocks, grid, slots, tasks, ntasks, counter, err, st);
    case 16: return ck_launch_r<512, 16>(occ, mode, nblocks, grid, slots, tasks, ntasks, counter, err, st);
    default: return -1;
    }
}
