// croma_fused_launch.h -- host-side dispatch of the fused kernel's template
// instances for one grid size (included by croma_fused_n*.hip, one
// translation unit per n so the instances compile in parallel).
#pragma once
#include "croma_device.h"

#define CK_FUSED_WPB 4

template <int N, int R, int MODE, int OCC, typename T>
static int ck_launch_inst(int nblocks, const CkGrid &grid, const CkSlot *slots, const CkTask *tasks, int ntasks,
                          int *counter, int *err, hipStream_t st)
{
    if constexpr (sizeof(T) != sizeof(double) && !fused_use_lg(N, R, CK_FUSED_WPB, MODE, OCC, (int)sizeof(T))) {
        return -1;  // no FP32-coefficient instance: its float tables do not fit the LDS
    } else {
        hipLaunchKernelGGL((k_fused<N, R, CK_FUSED_WPB, MODE, OCC, T>), dim3(nblocks), dim3(CK_FUSED_WPB * 64), 0,
                           st, grid, slots, tasks, ntasks, counter, err);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
}

template <int N, int R>
static int ck_launch_r(int occ, int mode, int nblocks, const CkGrid &grid, const CkSlot *slots,
                       const CkTask *tasks, int ntasks, int *counter, int *err, hipStream_t st)
{
#define CK_MODE_CASE(M, O)                                                                                             \
    case M: return ck_launch_inst<N, R, M, O, double>(nblocks, grid, slots, tasks, ntasks, counter, err, st);            \
    case M + CK_MODE_F32: return ck_launch_inst<N, R, M, O, float>(nblocks, grid, slots, tasks, ntasks, counter, err, st);
#define CK_OCC_CASE(O)                                                                                                 \
    case O:                                                                                                            \
        switch (mode) {                                                                                                \
            CK_MODE_CASE(0, O)                                                                                         \
            CK_MODE_CASE(1, O)                                                                                         \
            CK_MODE_CASE(2, O)                                                                                         \
            CK_MODE_CASE(3, O)                                                                                         \
            CK_MODE_CASE(CK_LAUNCH_NOSRC, O)                                                                           \
        default: return -1;                                                                                            \
        }
    switch (occ) {
        CK_OCC_CASE(1)
        CK_OCC_CASE(2)
    default: return -1;
    }
#undef CK_OCC_CASE
#undef CK_MODE_CASE
}
