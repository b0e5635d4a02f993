// Secondary-electron contraction of the split two-species launches: the
// proton pass records each segment's starting proton spectrum (a row of
// rec), this kernel turns every row into the segment's secondary-electron
// injection row, and the electron pass reads those rows:
//   inje[r][je] = sum_{jp >= np_min[je]} fqe_t[jp][je] rec[r][jp],  r < rows
// (the product of prepare_secondary_sources_batch, fp_shared_core.c:1508-1589,
// over the weights of prepare_secondary_kernel_flat, :1453-1506).
//
// FP64 MFMA (v_mfma_f64_16x16x4_f64), no LDS: a block takes 32 rows and its
// four waves each own a quarter of the n output columns as 16-column tiles,
// every weight fragment serving both 16-row tiles and every row fragment all
// of the wave's column tiles.  The rows come straight from memory into the A
// operand (the four waves of a block read the same rows, the later ones from
// the cache); the weights (n x n doubles, L2-resident) are the B operand.
// The wave's k loop starts at the first proton bin with a nonzero weight in
// any of its columns (np_min), so the zero triangle of the weights costs
// nothing; the weights below a column's own start are zero and add exact
// zeros, so no branch separates the products.  The row count can come from
// device memory (the device-side segment count of the window), so the
// launch shape depends only on the capacity and a captured graph replays it
// unchanged; blocks loop over row tiles and stop at the count.
#include <hip/hip_runtime.h>

#include "croma_kernels.h"

namespace {

typedef double sec_d4 __attribute__((ext_vector_type(4)));

template <int N>
__global__ __launch_bounds__(256) void k_sec_contract(const double *__restrict__ fqe_t, const int *__restrict__ np_min,
                                                      const double *__restrict__ rec, double *__restrict__ inje,
                                                      long rows_cap, const int *__restrict__ d_rows)
{
    constexpr int MT = 2;        // 16-row MFMA tiles per block (32 rows)
    constexpr int NT = N / 16;   // 16-column tiles
    constexpr int TW = NT / 4;   // column tiles per wave
    static_assert(NT % 4 == 0 && N % 16 == 0, "tile shape");
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i16 = lane & 15, q = lane >> 4;
    const long rows = d_rows ? min((long)*d_rows, rows_cap) : rows_cap;
    // the wave's k loop starts at the first nonzero weight of any of its
    // columns, rounded down to a multiple of 16
    int kmin = N;
#pragma unroll
    for (int i = 0; i < TW; ++i) {
        int v = np_min[(w * TW + i) * 16 + i16];
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) v = min(v, __shfl_xor(v, o, 64));
        // wave-uniform (every 16-lane block reduced the same columns)
        kmin = min(kmin, __builtin_amdgcn_readfirstlane(max(v, 0) & ~15));
    }
    // Within a group of 16 k the k order is permuted alike for both
    // operands: product u of the group takes k = k0 + 4 q + u in lane block
    // q, so a lane's four A values are one contiguous 32-byte load and its B
    // values four rows of the weights.
    const double *bcol = fqe_t + (size_t)(4 * q) * N + w * TW * 16 + i16;
    for (long r0 = (long)blockIdx.x * (16 * MT); r0 < rows; r0 += (long)gridDim.x * (16 * MT)) {
        const double *arow[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            // rows past the capacity read the last row (in bounds; their
            // products are not stored)
            const long r = min(r0 + 16 * m + i16, rows_cap - 1);
            arow[m] = rec + (size_t)r * N + 4 * q;
        }
        sec_d4 acc[MT][TW];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int i = 0; i < TW; ++i) acc[m][i] = sec_d4{0.0, 0.0, 0.0, 0.0};
        for (int k0 = kmin; k0 < N; k0 += 16) {
            double4 a[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m) a[m] = *reinterpret_cast<const double4 *>(arow[m] + k0);
            double bb[4][TW];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int i = 0; i < TW; ++i) bb[u][i] = bcol[(size_t)(k0 + u) * N + 16 * i];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int i = 0; i < TW; ++i)
#pragma unroll
                    for (int m = 0; m < MT; ++m) {
                        const double av = u == 0 ? a[m].x : u == 1 ? a[m].y : u == 2 ? a[m].z : a[m].w;
                        acc[m][i] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bb[u][i], acc[m][i], 0, 0, 0);
                    }
        }
        // D layout of the FP64 16x16 MFMA: register v of a lane holds row
        // 4 v + q of the tile, column i16 (one row per 16-lane block per
        // register, unlike the 4 consecutive rows of the FP32 16x16 forms)
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int i = 0; i < TW; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const long r = r0 + 16 * m + 4 * v + q;
                    if (r < rows) inje[(size_t)r * N + (w * TW + i) * 16 + i16] = acc[m][i][v];
                }
    }
}

}  // namespace

extern "C" int ck_sec_contract(int n, const double *fqe_t, const int *np_min, const double *rec, double *inje,
                               long rows_cap, const int *d_rows, int nblocks, void *stream)
{
    if (rows_cap <= 0) return 0;
    if (nblocks <= 0) return -1;
    const long tiles = (rows_cap + 31) / 32;
    const dim3 grid((unsigned)(tiles < nblocks ? tiles : nblocks));
    hipStream_t st = (hipStream_t)stream;
    switch (n) {
    case 64: hipLaunchKernelGGL(k_sec_contract<64>, grid, dim3(256), 0, st, fqe_t, np_min, rec, inje, rows_cap, d_rows); break;
    case 128: hipLaunchKernelGGL(k_sec_contract<128>, grid, dim3(256), 0, st, fqe_t, np_min, rec, inje, rows_cap, d_rows); break;
    case 256: hipLaunchKernelGGL(k_sec_contract<256>, grid, dim3(256), 0, st, fqe_t, np_min, rec, inje, rows_cap, d_rows); break;
    case 512: hipLaunchKernelGGL(k_sec_contract<512>, grid, dim3(256), 0, st, fqe_t, np_min, rec, inje, rows_cap, d_rows); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
