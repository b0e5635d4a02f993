// croma_emission.hip -- emission contractions on the device.
//
// IC, pp gamma and pp neutrino emissivities are one contraction each:
//   out[i][k] = row_scale[i] * sum_j spectra[i][j] * kmat[j][k]
// (prepare_ic_emission_cell_major fp_shared_core.c:2131-2157 with the dp
// weights folded into kmat; prepare_{gamma,neutrino}_emission_cell_major
// :2166-2243 with the trapezoid weights, beta and E_k folded into kmat and
// n_gas as the row scale).  Tiles of 64 tracers x 64 bins per block:
//   mode 0  FP32 on the vector ALUs
//   mode 1  split-BF16 x3 on MFMA (v_mfma_f32_16x16x32_bf16), FP32 accumulate
//   mode 2  one BF16 pass on MFMA, FP32 accumulate
//   mode 3  FP64 on the vector ALUs
// The reduced-precision modes scale every tracer's spectrum by a power of two
// (its largest entry) and every kernel column likewise before rounding, and
// undo both on output, so spectra spanning many decades keep their relative
// precision.  Synchrotron (prepare_synch_emission_cell_major :2070-2129)
// interpolates its kernel per tracer in log B and runs on the vector ALUs.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "croma_kernels.h"

#define EM_TR 64   // tracers per block
#define EM_TB 64   // bins per block
#define EM_KC 128  // K chunk staged in LDS (reduced-precision modes)
#define EM_KC64 64 // K chunk (FP64 mode)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t em_bf16_bits(float f)
{
    // round to nearest even (inputs are finite)
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float em_bf16_val(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Largest binary exponent of row i of spectra (frexp convention), 0 for an
// all-zero row.  4 threads per row.
__device__ __forceinline__ int em_row_shift(const double *__restrict__ row, int K, int sub)
{
    int e = -100000;
    for (int j = sub; j < K; j += 4) {
        const double v = fabs(row[j]);
        if (v > 0.0 && isfinite(v)) {
            const int ev = __builtin_amdgcn_frexp_exp(v);
            e = ev > e ? ev : e;
        }
    }
    e = max(e, __shfl_xor(e, 1, 4));
    e = max(e, __shfl_xor(e, 2, 4));
    return e == -100000 ? 0 : e;
}

// kmat_t: [nbins_pad][K] as bf16 hi and lo halves (column-scaled on the host
// side of this file), col_shift [nbins_pad].
template <int SPLIT>
__global__ __launch_bounds__(256) void k_emis_mfma(int ntr, int K, int nbins, int ld_out,
                                                   const double *__restrict__ spectra,
                                                   const uint16_t *__restrict__ kt_hi,
                                                   const uint16_t *__restrict__ kt_lo,
                                                   const int *__restrict__ col_shift,
                                                   const double *__restrict__ row_scale,
                                                   double *__restrict__ out)
{
    __shared__ __attribute__((aligned(16))) uint16_t a_hi[EM_TR][EM_KC + 8];
    __shared__ __attribute__((aligned(16))) uint16_t a_lo[EM_TR][EM_KC + 8];
    __shared__ __attribute__((aligned(16))) uint16_t b_hi[EM_TB][EM_KC + 8];
    __shared__ __attribute__((aligned(16))) uint16_t b_lo[EM_TB][EM_KC + 8];
    __shared__ int s_shift[EM_TR];
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int i0 = blockIdx.x * EM_TR;
    const int c0 = blockIdx.y * EM_TB;
    {
        const int r = t >> 2, sub = t & 3;
        int sh = 0;
        if (i0 + r < ntr) sh = em_row_shift(spectra + (size_t)(i0 + r) * K, K, sub);
        if (sub == 0) s_shift[r] = sh;
    }
    __syncthreads();
    floatx4 acc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += EM_KC) {
        const int kc = min(EM_KC, K - k0);
        // A chunk: rows i0.., columns k0..k0+kc, coalesced along each row
        for (int e = t; e < EM_TR * EM_KC; e += 256) {
            const int r = e / EM_KC, kk = e % EM_KC;
            float v = 0.f;
            if (i0 + r < ntr && kk < kc) v = (float)ldexp(spectra[(size_t)(i0 + r) * K + k0 + kk], -s_shift[r]);
            const uint16_t h = em_bf16_bits(v);
            a_hi[r][kk] = h;
            a_lo[r][kk] = SPLIT ? em_bf16_bits(v - em_bf16_val(h)) : 0;
        }
        for (int e = t; e < EM_TB * EM_KC; e += 256) {
            const int cc = e / EM_KC, kk = e % EM_KC;
            const size_t g = (size_t)(c0 + cc) * K + k0 + kk;
            b_hi[cc][kk] = (kk < kc) ? kt_hi[g] : 0;
            b_lo[cc][kk] = (SPLIT && kk < kc) ? kt_lo[g] : 0;
        }
        __syncthreads();
        for (int kk = 0; kk < kc; kk += 32) {
            const int ar = 16 * w + (lane & 15), ak = kk + (lane >> 4) * 8;
            const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(&a_hi[ar][ak]);
            const bf16x8 al = *reinterpret_cast<const bf16x8 *>(&a_lo[ar][ak]);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int bc = 16 * c + (lane & 15);
                const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(&b_hi[bc][ak]);
                acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[c], 0, 0, 0);
                if (SPLIT) {
                    const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(&b_lo[bc][ak]);
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[c], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D layout of the 16x16 MFMA: lane holds rows 4*(lane/16)+q, column lane%16
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int col = c0 + 16 * c + (lane & 15);
        if (col >= nbins) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = 16 * w + 4 * (lane >> 4) + q;
            if (i0 + r >= ntr) continue;
            double v = ldexp((double)acc[c][q], s_shift[r] + col_shift[col]);
            if (row_scale) v *= row_scale[i0 + r];
            out[(size_t)(i0 + r) * ld_out + col] = v;
        }
    }
}

// Streaming MFMA contraction for K = 32 KS <= 256 (reduced-precision modes).
// The kernel columns c0..c0+63 of the block stay in LDS for the block's
// life; each wave streams tiles of 16 tracers straight from HBM into the
// registers of the MFMA A operand (no LDS staging of spectra), scales each
// row by its largest power of two, rounds to bf16 hi (and lo) and runs
// v_mfma_f32_16x16x32_bf16 over 4 column tiles.
//
// Lane (r = lane % 16, q = lane / 16) loads, for k-step ks and m = 0..3, the
// double pair at k = 32 ks + 8 m + 2 q (+1), so the four q-lanes of a row
// read 64 contiguous bytes per load.  Its operand slot s = 2 m + (0|1) thus
// holds k(q, s) = 32 ks + 8 (s / 2) + 2 q + (s & 1); the B operand in LDS is
// stored with the same k order (a contraction does not depend on the order
// of k, only on A and B agreeing).
template <int SPLIT, int KS>
__global__ __launch_bounds__(256) void k_emis_stream(int ntr, int nbins, int ld_out,
                                                     const double *__restrict__ spectra,
                                                     const uint16_t *__restrict__ kt_hi,
                                                     const uint16_t *__restrict__ kt_lo,
                                                     const int *__restrict__ col_shift,
                                                     const double *__restrict__ row_scale,
                                                     double *__restrict__ out)
{
    constexpr int K = 32 * KS;
    // [ks][q][column][slot]: one 16-byte operand per (ks, q, column)
    __shared__ __attribute__((aligned(16))) uint16_t b_hi[KS][4][EM_TB][8];
    __shared__ __attribute__((aligned(16))) uint16_t b_lo[SPLIT ? KS : 1][4][EM_TB][8];
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int c0 = blockIdx.y * EM_TB;
    for (int e = t; e < KS * 4 * EM_TB * 8; e += 256) {
        const int sl = e & 7, c = (e >> 3) % EM_TB, q = (e / (8 * EM_TB)) & 3, ks = e / (32 * EM_TB);
        const int k = 32 * ks + 8 * (sl >> 1) + 2 * q + (sl & 1);
        const size_t g = (size_t)(c0 + c) * K + k;
        b_hi[ks][q][c][sl] = kt_hi[g];
        if (SPLIT) b_lo[ks][q][c][sl] = kt_lo[g];
    }
    __syncthreads();
    const int r = lane & 15, q = lane >> 4;
    const int ntiles = (ntr + 15) / 16;
    // the row streams in chunks of KC x 32 bins, the next chunk's loads in
    // flight while the current one converts and multiplies (registers for two
    // chunks instead of the whole row: more waves per SIMD)
    constexpr int KC = KS >= 4 ? 4 : (KS >= 2 ? 2 : 1);
    constexpr int NCH = KS / KC;
    constexpr int NONE = -100000;
    for (int tile = blockIdx.x * 4 + w; tile < ntiles; tile += gridDim.x * 4) {
        const int i = tile * 16 + r;
        const bool live = i < ntr;
        const double *row = spectra + (size_t)(live ? i : 0) * K + 2 * q;
        floatx4 acc[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
        // each row's sum is kept scaled by 2^-s_row, s_row the largest binary
        // exponent (frexp convention) of the row's entries seen so far; the
        // four lanes of row r are r, r + 16, r + 32, r + 48
        int s_row = NONE;
        double2 nxt[KC][4];
#pragma unroll
        for (int kk = 0; kk < KC; ++kk)
#pragma unroll
            for (int m = 0; m < 4; ++m)
                nxt[kk][m] = live ? *reinterpret_cast<const double2 *>(row + 32 * kk + 8 * m) : make_double2(0.0, 0.0);
#pragma unroll 1
        for (int ch = 0; ch < NCH; ++ch) {
            double2 a[KC][4];
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                for (int m = 0; m < 4; ++m) a[kk][m] = nxt[kk][m];
            if (ch + 1 < NCH) {
#pragma unroll
                for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                    for (int m = 0; m < 4; ++m)
                        nxt[kk][m] = live ? *reinterpret_cast<const double2 *>(row + 32 * (KC * (ch + 1) + kk) + 8 * m)
                                          : make_double2(0.0, 0.0);
            }
            int e = NONE;
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const double v0 = fabs(a[kk][m].x), v1 = fabs(a[kk][m].y);
                    if (v0 > 0.0 && isfinite(v0)) e = max(e, __builtin_amdgcn_frexp_exp(v0));
                    if (v1 > 0.0 && isfinite(v1)) e = max(e, __builtin_amdgcn_frexp_exp(v1));
                }
            e = max(e, __shfl_xor(e, 16, 64));
            e = max(e, __shfl_xor(e, 32, 64));
            const int s_new = max(s_row, e);
            if (ch > 0) {
                // rows whose scale rose: the partial sums move to the new scale
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int rr = 4 * q + j;
                    const int so = __shfl(s_row, rr, 64), sn = __shfl(s_new, rr, 64);
                    if (so != sn && so != NONE) {
                        const float f = ldexpf(1.0f, max(so - sn, -160));
#pragma unroll
                        for (int c = 0; c < 4; ++c) acc[c][j] *= f;
                    }
                }
            }
            s_row = s_new;
            const int sh = s_row == NONE ? 0 : s_row;
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) {
                const int ks = KC * ch + kk;
                bf16x8 ah, al;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const float f0 = (float)ldexp(a[kk][m].x, -sh), f1 = (float)ldexp(a[kk][m].y, -sh);
                    const uint16_t h0 = em_bf16_bits(f0), h1 = em_bf16_bits(f1);
                    ah[2 * m] = __builtin_bit_cast(__bf16, h0);
                    ah[2 * m + 1] = __builtin_bit_cast(__bf16, h1);
                    if (SPLIT) {
                        al[2 * m] = __builtin_bit_cast(__bf16, em_bf16_bits(f0 - em_bf16_val(h0)));
                        al[2 * m + 1] = __builtin_bit_cast(__bf16, em_bf16_bits(f1 - em_bf16_val(h1)));
                    }
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(&b_hi[ks][q][16 * c + r][0]);
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[c], 0, 0, 0);
                    if (SPLIT) {
                        const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(&b_lo[ks][q][16 * c + r][0]);
                        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[c], 0, 0, 0);
                        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[c], 0, 0, 0);
                    }
                }
            }
        }
        const int sh = s_row == NONE ? 0 : s_row;
        // D layout: lane holds rows 4 (lane / 16) + j of the tile, column lane % 16
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int rr = 4 * q + j;
            const int ii = tile * 16 + rr;
            const int shr = __shfl(sh, rr, 64);
            if (ii >= ntr) continue;
            const double rs = row_scale ? row_scale[ii] : 1.0;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int col = c0 + 16 * c + r;
                if (col >= nbins) continue;
                double v = ldexp((double)acc[c][j], shr + col_shift[col]);
                if (row_scale) v *= rs;
                out[(size_t)ii * ld_out + col] = v;
            }
        }
    }
}

// FP32 vector-ALU contraction: 4 x 4 outputs per thread.
__global__ __launch_bounds__(256) void k_emis_f32(int ntr, int K, int nbins, int ld_out,
                                                  const double *__restrict__ spectra,
                                                  const float *__restrict__ kt, const int *__restrict__ col_shift,
                                                  const double *__restrict__ row_scale, double *__restrict__ out)
{
    __shared__ float a_s[EM_KC][EM_TR + 4];
    __shared__ float b_s[EM_KC][EM_TB + 4];
    __shared__ int s_shift[EM_TR];
    const int t = threadIdx.x;
    const int i0 = blockIdx.x * EM_TR;
    const int c0 = blockIdx.y * EM_TB;
    {
        const int r = t >> 2, sub = t & 3;
        int sh = 0;
        if (i0 + r < ntr) sh = em_row_shift(spectra + (size_t)(i0 + r) * K, K, sub);
        if (sub == 0) s_shift[r] = sh;
    }
    __syncthreads();
    const int tr = t >> 4, tc = t & 15;
    float acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
    for (int k0 = 0; k0 < K; k0 += EM_KC) {
        const int kc = min(EM_KC, K - k0);
        for (int e = t; e < EM_TR * EM_KC; e += 256) {
            const int r = e / EM_KC, kk = e % EM_KC;
            float v = 0.f;
            if (i0 + r < ntr && kk < kc) v = (float)ldexp(spectra[(size_t)(i0 + r) * K + k0 + kk], -s_shift[r]);
            a_s[kk][r] = v;
        }
        for (int e = t; e < EM_TB * EM_KC; e += 256) {
            const int cc = e / EM_KC, kk = e % EM_KC;
            b_s[kk][cc] = (kk < kc) ? kt[(size_t)(c0 + cc) * K + k0 + kk] : 0.f;
        }
        __syncthreads();
        for (int kk = 0; kk < kc; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = a_s[kk][4 * tr + q];
                b[q] = b_s[kk][tc + 16 * q];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] = fmaf(a[p], b[q], acc[p][q]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int r = 4 * tr + p;
        if (i0 + r >= ntr) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int col = c0 + tc + 16 * q;
            if (col >= nbins) continue;
            double v = ldexp((double)acc[p][q], s_shift[r] + col_shift[col]);
            if (row_scale) v *= row_scale[i0 + r];
            out[(size_t)(i0 + r) * ld_out + col] = v;
        }
    }
}

// FP64 vector-ALU contraction (the reference precision), 4 x 4 per thread.
__global__ __launch_bounds__(256) void k_emis_f64(int ntr, int K, int nbins, int ld_out,
                                                  const double *__restrict__ spectra,
                                                  const double *__restrict__ kt64,
                                                  const double *__restrict__ row_scale, double *__restrict__ out)
{
    __shared__ double a_s[EM_KC64][EM_TR + 2];
    __shared__ double b_s[EM_KC64][EM_TB + 2];
    const int t = threadIdx.x;
    const int i0 = blockIdx.x * EM_TR;
    const int c0 = blockIdx.y * EM_TB;
    const int tr = t >> 4, tc = t & 15;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int k0 = 0; k0 < K; k0 += EM_KC64) {
        const int kc = min(EM_KC64, K - k0);
        for (int e = t; e < EM_TR * EM_KC64; e += 256) {
            const int r = e / EM_KC64, kk = e % EM_KC64;
            a_s[kk][r] = (i0 + r < ntr && kk < kc) ? spectra[(size_t)(i0 + r) * K + k0 + kk] : 0.0;
        }
        for (int e = t; e < EM_TB * EM_KC64; e += 256) {
            const int cc = e / EM_KC64, kk = e % EM_KC64;
            b_s[kk][cc] = (kk < kc) ? kt64[(size_t)(c0 + cc) * K + k0 + kk] : 0.0;
        }
        __syncthreads();
        for (int kk = 0; kk < kc; ++kk) {
            double a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = a_s[kk][4 * tr + q];
                b[q] = b_s[kk][tc + 16 * q];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] = fma(a[p], b[q], acc[p][q]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int r = 4 * tr + p;
        if (i0 + r >= ntr) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int col = c0 + tc + 16 * q;
            if (col >= nbins) continue;
            double v = acc[p][q];
            if (row_scale) v *= row_scale[i0 + r];
            out[(size_t)(i0 + r) * ld_out + col] = v;
        }
    }
}

extern "C" int ck_emission_contract(int ntr, int K, int nbins, int nbins_pad, const double *spectra,
                                    const uint16_t *kt_hi, const uint16_t *kt_lo, const float *kt32,
                                    const double *kt64, const int *col_shift, const double *row_scale,
                                    double *out, int mode, void *stream)
{
    if (ntr <= 0 || K <= 0 || nbins <= 0 || nbins_pad < nbins || nbins_pad % EM_TB) return -1;
    const dim3 grid((ntr + EM_TR - 1) / EM_TR, nbins_pad / EM_TB);
    hipStream_t st = (hipStream_t)stream;
    switch (mode) {
    case 0:
        hipLaunchKernelGGL(k_emis_f32, grid, dim3(256), 0, st, ntr, K, nbins, nbins, spectra, kt32, col_shift,
                           row_scale, out);
        break;
    case 1:
    case 2:
        if (K == 64 || K == 128 || K == 256) {
            // persistent blocks: two per CU (the B tile of a split K = 256
            // run takes 64 KB of LDS)
            static int ncu = 0;
            if (!ncu && (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess ||
                         ncu <= 0))
                ncu = 256;
            const int tiles = (ntr + 15) / 16;
            int gx = 2 * ncu;
            if (gx > (tiles + 3) / 4) gx = (tiles + 3) / 4;
            const dim3 gs(gx, nbins_pad / EM_TB);
#define EM_STREAM(SP, KS_)                                                                                             \
    hipLaunchKernelGGL((k_emis_stream<SP, KS_>), gs, dim3(256), 0, st, ntr, nbins, nbins, spectra, kt_hi, kt_lo,       \
                       col_shift, row_scale, out)
            if (mode == 1) {
                if (K == 64) EM_STREAM(1, 2);
                else if (K == 128) EM_STREAM(1, 4);
                else EM_STREAM(1, 8);
            } else {
                if (K == 64) EM_STREAM(0, 2);
                else if (K == 128) EM_STREAM(0, 4);
                else EM_STREAM(0, 8);
            }
#undef EM_STREAM
        } else if (mode == 1) {
            hipLaunchKernelGGL(k_emis_mfma<1>, grid, dim3(256), 0, st, ntr, K, nbins, nbins, spectra, kt_hi, kt_lo,
                               col_shift, row_scale, out);
        } else {
            hipLaunchKernelGGL(k_emis_mfma<0>, grid, dim3(256), 0, st, ntr, K, nbins, nbins, spectra, kt_hi, kt_lo,
                               col_shift, row_scale, out);
        }
        break;
    case 3:
        hipLaunchKernelGGL(k_emis_f64, grid, dim3(256), 0, st, ntr, K, nbins, nbins, spectra, kt64, row_scale,
                           out);
        break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Synchrotron pitch-kernel table (prepare_synch_pitch_kernel_table,
// fp_shared_core.c:2008-2052): table[ib][nf][je] = sum_k F(x_k) w_k with
// log x_k = logy[nf][je][k] - logb_ib and F interpolated log-linearly in the
// log F(x) table (0 below its first node, its last value past the end).  One
// thread per (ib, nf, je); the k sum runs in the reference's order.
__global__ __launch_bounds__(256) void k_synch_table(int n, int nfreq, int ntheta, int nlogb, double logb_min,
                                                     double inv_dlogb, int nx_tab, double xmin,
                                                     const double *__restrict__ logx_tab,
                                                     const double *__restrict__ logfx_tab,
                                                     const double *__restrict__ fx_tab,
                                                     const double *__restrict__ logy,
                                                     const double *__restrict__ pitch_weight,
                                                     double *__restrict__ table)
{
#pragma clang fp contract(off)
    const size_t o = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= (size_t)nlogb * nfreq * n) return;
    const int je = (int)(o % n);
    const int nf = (int)((o / n) % nfreq);
    const int ib = (int)(o / ((size_t)n * nfreq));
    const double logb = logb_min + (double)ib / inv_dlogb;
    const double inv_dlogx = (nx_tab > 1) ? 1.0 / (logx_tab[1] - logx_tab[0]) : 0.0;
    const double xlast = logx_tab[nx_tab - 1];
    const double *ly = logy + ((size_t)nf * n + je) * ntheta;
    double int_th = 0.0;
    for (int k = 0; k < ntheta; ++k) {
        double Fx = 0.0;
        const double logx = ly[k] - logb;
        if (logx <= xlast) {
            const int ix = (int)floor((logx - xmin) * inv_dlogx);
            if (ix >= 0) {
                if (ix >= nx_tab - 1) {
                    Fx = fx_tab[nx_tab - 1];
                } else {
                    const double frac = (logx - logx_tab[ix]) * inv_dlogx;
                    Fx = pow(10.0, logfx_tab[ix] + frac * (logfx_tab[ix + 1] - logfx_tab[ix]));
                }
            }
        }
        int_th += Fx * pitch_weight[k];
    }
    table[o] = int_th;
}

extern "C" int ck_synch_table(int n, int nfreq, int ntheta, int nlogb, double logb_min, double inv_dlogb,
                              int nx_tab, double xmin, const double *logx_tab, const double *logfx_tab,
                              const double *fx_tab, const double *logy, const double *pitch_weight, double *table,
                              void *stream)
{
    if (n <= 0 || nfreq <= 0 || ntheta <= 0 || nlogb <= 0 || nx_tab < 1) return -1;
    const size_t tot = (size_t)nlogb * nfreq * n;
    hipLaunchKernelGGL(k_synch_table, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n,
                       nfreq, ntheta, nlogb, logb_min, inv_dlogb, nx_tab, xmin, logx_tab, logfx_tab, fx_tab, logy,
                       pitch_weight, table);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Logarithms of the synchrotron kernel table, transposed to [ib][je][nf]
// (frequency fastest, so the lanes of a wave read one coalesced row); NaN
// marks entries that are not positive (linear interpolation there).
__global__ __launch_bounds__(256) void k_synch_logt(int nlogb, int nfreq, int n, const double *__restrict__ table,
                                                    double *__restrict__ logt)
{
    const size_t o = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= (size_t)nlogb * nfreq * n) return;
    const int nf = (int)(o % nfreq);
    const int je = (int)((o / nfreq) % n);
    const size_t ib = o / ((size_t)nfreq * n);
    const double v = table[(ib * nfreq + nf) * n + je];
    logt[o] = v > 0.0 ? log(v) : __builtin_nan("");
}

// The same table as float log2 (FP32 mode), layout [ib][je][nf]: -inf for a
// zero entry, NaN for a negative one.
__global__ __launch_bounds__(256) void k_synch_log2t32(int nlogb, int nfreq, int n, const double *__restrict__ table,
                                                       float *__restrict__ lt)
{
    const size_t o = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= (size_t)nlogb * nfreq * n) return;
    const int nf = (int)(o % nfreq);
    const int je = (int)((o / nfreq) % n);
    const size_t ib = o / ((size_t)nfreq * n);
    const double v = table[(ib * nfreq + nf) * n + je];
    lt[o] = v > 0.0 ? (float)log2(v) : (v == 0.0 ? -__builtin_inff() : __builtin_nanf(""));
}

// Synchrotron: one wave per tracer, lanes over frequencies.  The kernel is
// interpolated in log B between the table planes ib, ib+1 as the reference
// does (log-linear when both entries are positive, else linear); the sum
// runs over je = 2.. up to the first bin whose successor is not positive.
// With logt (FP64), the logarithms come from the table of k_synch_logt.
// Table plane ib of a tracer's field (the same arithmetic as the
// interpolation weight below).
__device__ __forceinline__ int em_synch_plane(double bval, double logb_min, double inv_dlogb, int nlogb)
{
    if (!(bval > 0.0) || !isfinite(bval)) bval = 1.0e-30;
    const double u = (log10(bval / 1.0e-6) - logb_min) * inv_dlogb;
    int ib = (int)floor(u);
    if (ib < 0) ib = 0;
    if (ib > nlogb - 2) ib = nlogb - 2;
    return ib;
}

// Tracers grouped by table plane (counting sort): waves that run one after
// the other read the same two planes, which then stay in L2.
// Up to EM_LDS_HIST planes the per-tracer kernels count in a block-local LDS
// histogram and add to the global one once per non-empty plane of the block
// (one global atomic per tracer serialises on the planes many tracers share).
#define EM_LDS_HIST 2048
__global__ __launch_bounds__(256) void k_synch_plane_count(int ntr, const double *__restrict__ b_field,
                                                           double logb_min, double inv_dlogb, int nlogb,
                                                           int *__restrict__ key, int *__restrict__ hist)
{
    __shared__ int s_hist[EM_LDS_HIST];
    const bool local = nlogb <= EM_LDS_HIST;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (local) {
        for (int b = threadIdx.x; b < nlogb; b += 256) s_hist[b] = 0;
        __syncthreads();
    }
    if (i < ntr) {
        const int ib = em_synch_plane(b_field[i], logb_min, inv_dlogb, nlogb);
        key[i] = ib;
        atomicAdd(local ? &s_hist[ib] : &hist[ib], 1);
    }
    if (local) {
        __syncthreads();
        for (int b = threadIdx.x; b < nlogb; b += 256) {
            const int c = s_hist[b];
            if (c) atomicAdd(&hist[b], c);
        }
    }
}

__global__ __launch_bounds__(256) void k_synch_plane_scan(int nh, int *__restrict__ hist)
{
    __shared__ int s_part[256];
    __shared__ int s_carry;
    const int t = threadIdx.x;
    if (t == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < nh; base += 256) {
        const int x = base + t < nh ? hist[base + t] : 0;
        s_part[t] = x;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int y = t >= o ? s_part[t - o] : 0;
            __syncthreads();
            s_part[t] += y;
            __syncthreads();
        }
        const int carry = s_carry;
        if (base + t < nh) hist[base + t] = carry + s_part[t] - x;
        __syncthreads();
        if (t == 0) s_carry = carry + s_part[255];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_synch_plane_place(int ntr, int nlogb, const int *__restrict__ key,
                                                           int *__restrict__ hist, int *__restrict__ perm)
{
    __shared__ int s_hist[EM_LDS_HIST];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (nlogb > EM_LDS_HIST) {
        if (i < ntr) perm[atomicAdd(&hist[key[i]], 1)] = i;
        return;
    }
    for (int b = threadIdx.x; b < nlogb; b += 256) s_hist[b] = 0;
    __syncthreads();
    const int k = i < ntr ? key[i] : 0;
    const int rank = i < ntr ? atomicAdd(&s_hist[k], 1) : 0;
    __syncthreads();
    for (int b = threadIdx.x; b < nlogb; b += 256) {
        const int c = s_hist[b];
        if (c) s_hist[b] = atomicAdd(&hist[b], c);
    }
    __syncthreads();
    if (i < ntr) perm[s_hist[k] + rank] = i;
}

template <typename T>
__global__ __launch_bounds__(256) void k_emis_synch(int ntr, int n, int nfreq, int nlogb, double logb_min,
                                                    double inv_dlogb, const double *__restrict__ table,
                                                    const double *__restrict__ logt, const int *__restrict__ perm,
                                                    const double *__restrict__ cre, const double *__restrict__ dp,
                                                    const double *__restrict__ b_field, double *__restrict__ eps)
{
    __shared__ T s_w[4][512];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int slot = blockIdx.x * 4 + w;
    if (slot >= ntr) return;
    const int it = perm ? perm[slot] : slot;
    const double *cl = cre + (size_t)it * n;
    // end of the summation: first je >= 2 with cl[je + 1] <= 0 (or n - 1)
    int jend = n - 1;
    for (int base = 2; base < n - 1; base += 64) {
        const int je = base + lane;
        const bool stop = je < n - 1 && !(cl[je + 1] > 0.0);
        const unsigned long long m = __ballot(stop);
        if (m) {
            jend = base + __ffsll((long long)m) - 1;
            break;
        }
    }
    for (int je = lane; je < n; je += 64) s_w[w][je] = (T)(cl[je] * dp[je]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double bval = b_field[it];
    if (!(bval > 0.0) || !isfinite(bval)) bval = 1.0e-30;
    const double logb = log10(bval / 1.0e-6);
    const double u = (logb - logb_min) * inv_dlogb;
    int ib = (int)floor(u);
    if (ib < 0) ib = 0;
    if (ib > nlogb - 2) ib = nlogb - 2;
    double tt = u - (double)ib;
    if (tt < 0.0) tt = 0.0;
    if (tt > 1.0) tt = 1.0;
    const T t = (T)tt;
    const double A = sqrt(3.0) * pow(4.8e-10, 3.0) * b_field[it] / (2.0 * 9.11e-28 * 2.999e10 * 2.999e10) /
                     (4.0 * 3.14159265358979323846);
    const bool inner = logt && tt > 0.0 && tt < 1.0;
    for (int nf = lane; nf < nfreq; nf += 64) {
        const double *r0 = table + ((size_t)ib * nfreq + nf) * n;
        const double *r1 = table + ((size_t)(ib + 1) * nfreq + nf) * n;
        T integral = 0;
        if (inner) {
            const double *l0p = logt + (size_t)ib * n * nfreq + nf;
            const double *l1p = l0p + (size_t)n * nfreq;
            // loads of 8 bins are issued before their exps (the sum keeps its order)
            for (int j0 = 2; j0 < jend; j0 += 8) {
                double l0[8], l1[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int je = j0 + u < jend ? j0 + u : j0;
                    l0[u] = l0p[(size_t)je * nfreq];
                    l1[u] = l1p[(size_t)je * nfreq];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int je = j0 + u;
                    if (je >= jend) break;
                    T K;
                    if (!__builtin_isnan(l0[u]) && !__builtin_isnan(l1[u])) {
                        K = exp(((T)1 - t) * (T)l0[u] + t * (T)l1[u]);
                    } else {
                        const T k0 = (T)r0[je], k1 = (T)r1[je];
                        K = ((T)1 - t) * k0 + t * k1;
                    }
                    integral += s_w[w][je] * K;
                }
            }
        } else {
            for (int je = 2; je < jend; ++je) {
                const T k0 = (T)r0[je], k1 = (T)r1[je];
                T K;
                if (tt <= 0.0) K = k0;
                else if (tt >= 1.0) K = k1;
                else if (k0 > (T)0 && k1 > (T)0) K = exp(((T)1 - t) * log(k0) + t * log(k1));
                else K = ((T)1 - t) * k0 + t * k1;
                integral += s_w[w][je] * K;
            }
        }
        eps[(size_t)it * nfreq + nf] = A * (double)integral;
    }
}

// FP32 mode of the same sum (emissivities within 1e-4).  Every term
// w K = sign(w) 2^(log2|w| + log2 K) is accumulated relative to the running
// largest exponent of the lane's sum (rescaled once per block of 8 bins), so
// no term leaves float's range however far the spectrum and the kernel span:
// log2|w| per bin in LDS (w = N dp), log2 K interpolated on the float log2
// table, the hardware exp2, and the FP64 result A acc 2^E.
__global__ __launch_bounds__(256) void k_emis_synch32(int ntr, int n, int nfreq, int nlogb, double logb_min,
                                                      double inv_dlogb, const double *__restrict__ table,
                                                      const float *__restrict__ lt, const int *__restrict__ perm,
                                                      const double *__restrict__ cre, const double *__restrict__ dp,
                                                      const double *__restrict__ b_field, double *__restrict__ eps,
                                                      const int *__restrict__ count)
{
    __shared__ float s_lw[4][512];
    __shared__ float s_sg[4][512];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int slot = blockIdx.x * 4 + w;
    // count (device): the number of tracers listed in perm, when the list
    // was made on the device
    if (slot >= (count ? *count : ntr)) return;
    const int it = perm ? perm[slot] : slot;
    const double *cl = cre + (size_t)it * n;
    int jend = n - 1;
    for (int base = 2; base < n - 1; base += 64) {
        const int je = base + lane;
        const bool stop = je < n - 1 && !(cl[je + 1] > 0.0);
        const unsigned long long m = __ballot(stop);
        if (m) {
            jend = base + __ffsll((long long)m) - 1;
            break;
        }
    }
    for (int je = lane; je < n; je += 64) {
        const double v = cl[je] * dp[je];
        s_lw[w][je] = v != 0.0 ? (float)log2(fabs(v)) : -__builtin_inff();
        s_sg[w][je] = v > 0.0 ? 1.0f : (v < 0.0 ? -1.0f : 0.0f);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double bval = b_field[it];
    if (!(bval > 0.0) || !isfinite(bval)) bval = 1.0e-30;
    const double u = (log10(bval / 1.0e-6) - logb_min) * inv_dlogb;
    int ib = (int)floor(u);
    if (ib < 0) ib = 0;
    if (ib > nlogb - 2) ib = nlogb - 2;
    double tt = u - (double)ib;
    if (tt < 0.0) tt = 0.0;
    if (tt > 1.0) tt = 1.0;
    const float t = (float)tt;
    const double A = sqrt(3.0) * pow(4.8e-10, 3.0) * b_field[it] / (2.0 * 9.11e-28 * 2.999e10 * 2.999e10) /
                     (4.0 * 3.14159265358979323846);
    const float ninf = -__builtin_inff();
    // one plane's entry zero: the linear interpolation is t K1 or (1 - t) K0
    const float l2t = tt > 0.0 ? (float)log2(tt) : ninf;
    const float l2u = tt < 1.0 ? (float)log2(1.0 - tt) : ninf;
    for (int nf = lane; nf < nfreq; nf += 64) {
        const float *l0p = lt + (size_t)ib * n * nfreq + nf;
        const float *l1p = l0p + (size_t)n * nfreq;
        const double *r0 = table + ((size_t)ib * nfreq + nf) * n;
        const double *r1 = table + ((size_t)(ib + 1) * nfreq + nf) * n;
        float E = ninf, acc = 0.0f;
        for (int j0 = 2; j0 < jend; j0 += 8) {
            float a[8], b[8], x[8], sg[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int je = j0 + q < jend ? j0 + q : j0;
                a[q] = l0p[(size_t)je * nfreq];
                b[q] = l1p[(size_t)je * nfreq];
            }
            float mx = ninf;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int je = j0 + q;
                x[q] = ninf;
                sg[q] = 0.0f;
                if (je < jend) {
                    // log-linear where both entries are positive, else linear
                    const bool pa = a[q] > ninf, pb = b[q] > ninf;
                    float lk = (pa && pb) ? fmaf(t, b[q] - a[q], a[q]) : (pa ? l2u + a[q] : (pb ? l2t + b[q] : ninf));
                    float sk = 1.0f;
                    if (__builtin_isnan(a[q]) || __builtin_isnan(b[q])) {
                        // a negative entry: the interpolation in FP64
                        const double K = (1.0 - tt) * r0[je] + tt * r1[je];
                        lk = K != 0.0 ? (float)log2(fabs(K)) : ninf;
                        sk = K < 0.0 ? -1.0f : 1.0f;
                    }
                    x[q] = s_lw[w][je] + lk;
                    sg[q] = s_sg[w][je] * sk;
                }
                mx = fmaxf(mx, x[q]);
            }
            if (mx > E) {
                acc *= __builtin_amdgcn_exp2f(E - mx);
                E = mx;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (x[q] > ninf) acc = fmaf(sg[q], __builtin_amdgcn_exp2f(x[q] - E), acc);
        }
        eps[(size_t)it * nfreq + nf] = (E > ninf) ? A * (double)acc * exp2((double)E) : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Synchrotron on MFMA.  Between table planes p and p + 1 the reference
// interpolates the kernel log-linearly in log B where both entries are
// positive: K(t) = K0^(1-t) K1^t = sqrt(K0 K1) e^(b s), s = 2 t - 1,
// b = ln(K1 / K0) / 2, else linearly: K = (K0 + K1) / 2 + s (K1 - K0) / 2.
// With the Chebyshev expansion e^(b s) = I_0(b) + 2 sum_m I_m(b) T_m(s)
// (I_m: modified Bessel functions) truncated after T_7, every tracer of plane
// interval p has
//   eps_f = A(B) sum_m T_m(s) sum_j N_j [dp_j C_m(p, j, f)],
// a GEMM of the plane's spectra against 8 columns per frequency followed by
// an 8-term sum.  The truncation error is 2 I_8(b) / I_0(b) < 3e-6 for
// |b| < 1.4; b = (d ln K / d ln x) (Delta ln B) / 2 with d ln K / d ln x ~ x
// in the exponential tail, so with table planes no coarser than 0.02 in
// log10 B it holds for every entry with K above e^-57 of the kernel's peak
// (the entries that reach emissivities within 25 decades of a tracer's
// peak).  Entries with |b| > 30 (ratios beyond e^60 between adjacent planes,
// far out in the tail) use the linear form.
// ---------------------------------------------------------------------------
#define EM_SY_M 8 // Chebyshev terms per (plane interval, bin, frequency)
// Both operands are scaled so that their largest entry is 2^59 (not 1): a
// product of the largest entries is then 2^118 and a sum of 256 of them stays
// below FP32's 2^128, while products down to 2^-126 (244 binary orders, 73
// decades below the largest) stay normal.  With unit maxima the products of
// a spectrum's cutoff bins (N_j ~ 1e-38 of the row's largest) with the
// kernel's entries would leave FP32's range although they carry frequencies
// within 25 decades of the tracer's peak emissivity.
#define EM_SY_OFF 59

// I_0(b) .. I_7(b) by their power series (|b| <= 30).
__device__ __forceinline__ void em_bessel_i8(double b, double (&I)[EM_SY_M])
{
    const double h = 0.5 * b, h2 = h * h;
    double pm = 1.0; // h^m / m!
#pragma unroll
    for (int m = 0; m < EM_SY_M; ++m) {
        if (m > 0) pm *= h / (double)m;
        double term = pm, sum = pm;
        for (int k = 1; k < 96; ++k) {
            term *= h2 / ((double)k * (double)(k + m));
            sum += term;
            if (fabs(term) <= 1.0e-17 * fabs(sum)) break;
        }
        I[m] = sum;
    }
}

// Chebyshev coefficients of the interpolated kernel at one entry, times dp.
__device__ __forceinline__ void em_synch_cheb(double K0, double K1, double w, double (&c)[EM_SY_M])
{
#pragma unroll
    for (int m = 0; m < EM_SY_M; ++m) c[m] = 0.0;
    if (K0 > 0.0 && K1 > 0.0) {
        const double b = 0.5 * log(K1 / K0);
        if (fabs(b) <= 30.0) {
            double I[EM_SY_M];
            em_bessel_i8(b, I);
            const double base = sqrt(K0) * sqrt(K1) * w;
            c[0] = base * I[0];
#pragma unroll
            for (int m = 1; m < EM_SY_M; ++m) c[m] = 2.0 * base * I[m];
            return;
        }
    }
    c[0] = 0.5 * (K0 + K1) * w;
    c[1] = 0.5 * (K1 - K0) * w;
}

// Pass 0: the largest binary exponent of every column (p, f * 8 + m) over its
// bins (atomic max); pass 1: the entries scaled by their column's power of
// two, split into bf16 hi / lo.  One thread per (p, f, j); bins 0 and 1 are
// zero (the sum starts at bin 2).  kt [p][f * 8 + m][j], col_shift [p][f * 8 + m].
__global__ __launch_bounds__(256) void k_synch_cheb(int pass, int n, int nfreq, int nlogb,
                                                    const double *__restrict__ table, const double *__restrict__ dp,
                                                    uint16_t *__restrict__ kt_hi, uint16_t *__restrict__ kt_lo,
                                                    int *__restrict__ col_shift)
{
    const size_t o = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t tot = (size_t)(nlogb - 1) * nfreq * n;
    if (o >= tot) return;
    const int j = (int)(o % n);
    const int f = (int)((o / n) % nfreq);
    const int pl = (int)(o / ((size_t)n * nfreq));
    double c[EM_SY_M];
    if (j < 2) {
#pragma unroll
        for (int m = 0; m < EM_SY_M; ++m) c[m] = 0.0;
    } else {
        em_synch_cheb(table[((size_t)pl * nfreq + f) * n + j], table[((size_t)(pl + 1) * nfreq + f) * n + j], dp[j],
                      c);
    }
    int *cs = col_shift + ((size_t)pl * nfreq + f) * EM_SY_M;
    if (pass == 0) {
#pragma unroll
        for (int m = 0; m < EM_SY_M; ++m) {
            const double v = fabs(c[m]);
            if (v > 0.0 && isfinite(v)) atomicMax(&cs[m], __builtin_amdgcn_frexp_exp(v));
        }
        return;
    }
#pragma unroll
    for (int m = 0; m < EM_SY_M; ++m) {
        const int sh = cs[m] < -2000 ? 0 : cs[m] - EM_SY_OFF;
        const float v = (float)ldexp(isfinite(c[m]) ? c[m] : 0.0, -sh);
        const uint16_t h = em_bf16_bits(v);
        const size_t g = (((size_t)pl * nfreq + f) * EM_SY_M + m) * n + j;
        kt_hi[g] = h;
        kt_lo[g] = em_bf16_bits(v - em_bf16_val(h));
    }
}

// Per tracer: the end of its sum (first bin whose successor is not positive,
// as k_emis_synch32), s = 2 t - 1 of its field in its plane interval, the
// prefactor A(B) and the largest binary exponent of the bins before the end
// (the contraction's row scale; EM_SY_NONE for an empty row).  One wave per
// tracer.
#define EM_SY_NONE (-100000)
__global__ __launch_bounds__(256) void k_synch_prep(int ntr, int n, int nlogb, double logb_min, double inv_dlogb,
                                                    const double *__restrict__ cre,
                                                    const double *__restrict__ b_field, int *__restrict__ jend_out,
                                                    double *__restrict__ s_out, double *__restrict__ pref_out,
                                                    int *__restrict__ srow_out)
{
    const int lane = threadIdx.x & 63;
    const int it = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (it >= ntr) return;
    const double *cl = cre + (size_t)it * n;
    // one pass over the row: the end bin and the exponents of the bins
    // before it (bins 0 and 1 always come before it)
    int e = EM_SY_NONE;
    if (lane < 2) {
        const double v = fabs(cl[lane]);
        if (v > 0.0 && isfinite(v)) e = __builtin_amdgcn_frexp_exp(v);
    }
    int jend = n - 1;
    for (int base = 2; base < n - 1; base += 64) {
        const int je = base + lane;
        const bool in = je < n - 1;
        const double v = in ? fabs(cl[je]) : 0.0;
        const bool stop = in && !(cl[je + 1] > 0.0);
        const unsigned long long m = __ballot(stop);
        const int lim = m ? base + __ffsll((long long)m) - 1 : n - 1;
        if (je < lim && v > 0.0 && isfinite(v)) e = max(e, __builtin_amdgcn_frexp_exp(v));
        if (m) {
            jend = lim;
            break;
        }
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) e = max(e, __shfl_xor(e, o, 64));
    if (lane == 0) {
        double bval = b_field[it];
        if (!(bval > 0.0) || !isfinite(bval)) bval = 1.0e-30;
        const double u = (log10(bval / 1.0e-6) - logb_min) * inv_dlogb;
        int ib = (int)floor(u);
        if (ib < 0) ib = 0;
        if (ib > nlogb - 2) ib = nlogb - 2;
        double tt = u - (double)ib;
        if (tt < 0.0) tt = 0.0;
        if (tt > 1.0) tt = 1.0;
        srow_out[it] = e;
        jend_out[it] = jend;
        s_out[it] = 2.0 * tt - 1.0;
        pref_out[it] = sqrt(3.0) * pow(4.8e-10, 3.0) * b_field[it] / (2.0 * 9.11e-28 * 2.999e10 * 2.999e10) /
                       (4.0 * 3.14159265358979323846);
    }
}

// The plane-grouped contraction: block (plane interval p, column block cb)
// keeps the 64 columns cb * 64 .. (8 frequencies x 8 Chebyshev terms) of
// plane p in LDS and streams the 16-tracer tiles of p's tracers (perm order,
// p's range [hist[p - 1], hist[p])) as k_emis_stream does, the rows masked
// past each tracer's end bin.  Blocks are numbered so that the column
// blocks of one plane share an XCD (blockIdx % 8), whose L2 then serves the
// plane's spectra to all of them.  Output: eps[i][f] = A_i sum_m T_m(s_i) G.
template <int KS>
__global__ __launch_bounds__(256) void k_synch_mfma(int nfreq, int nlogb, int ncb, const double *__restrict__ cre,
                                                    const int *__restrict__ perm, const int *__restrict__ plane_end,
                                                    const int *__restrict__ jend, const double *__restrict__ sv,
                                                    const double *__restrict__ pref,
                                                    const uint16_t *__restrict__ kt_hi,
                                                    const uint16_t *__restrict__ kt_lo,
                                                    const int *__restrict__ col_shift, double *__restrict__ eps)
{
    constexpr int K = 32 * KS;
    __shared__ __attribute__((aligned(16))) uint16_t b_hi[KS][4][EM_TB][8];
    __shared__ __attribute__((aligned(16))) uint16_t b_lo[KS][4][EM_TB][8];
    const int xcd = blockIdx.x & 7, loc = blockIdx.x >> 3;
    const int cb = loc % ncb;
    const int pl = (loc / ncb) * 8 + xcd;
    if (pl >= nlogb - 1) return;
    const int start = pl ? plane_end[pl - 1] : 0, cnt = plane_end[pl] - start;
    if (cnt <= 0) return;
    const int ncol = nfreq * EM_SY_M;
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const size_t pbase = (size_t)pl * ncol;
    for (int e = t; e < KS * 4 * EM_TB * 8; e += 256) {
        const int sl = e & 7, c = (e >> 3) % EM_TB, q = (e / (8 * EM_TB)) & 3, ks = e / (32 * EM_TB);
        const int k = 32 * ks + 8 * (sl >> 1) + 2 * q + (sl & 1);
        const int col = cb * EM_TB + c;
        const size_t g = (pbase + (col < ncol ? col : 0)) * K + k;
        b_hi[ks][q][c][sl] = col < ncol ? kt_hi[g] : (uint16_t)0;
        b_lo[ks][q][c][sl] = col < ncol ? kt_lo[g] : (uint16_t)0;
    }
    __syncthreads();
    const int r = lane & 15, q = lane >> 4;
    const int ntiles = (cnt + 15) / 16;
    constexpr int KC = KS >= 4 ? 4 : (KS >= 2 ? 2 : 1);
    constexpr int NCH = KS / KC;
    constexpr int NONE = -100000;
    for (int tile = w; tile < ntiles; tile += 4) {
        const bool live = tile * 16 + r < cnt;
        const int it = live ? perm[start + tile * 16 + r] : 0;
        const int je = live ? jend[it] : 0;
        const double *row = cre + (size_t)it * K + 2 * q;
        floatx4 acc[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = floatx4{0.f, 0.f, 0.f, 0.f};
        int s_row = NONE;
        double2 nxt[KC][4];
#pragma unroll
        for (int kk = 0; kk < KC; ++kk)
#pragma unroll
            for (int m = 0; m < 4; ++m)
                nxt[kk][m] = live ? *reinterpret_cast<const double2 *>(row + 32 * kk + 8 * m) : make_double2(0.0, 0.0);
#pragma unroll 1
        for (int ch = 0; ch < NCH; ++ch) {
            double2 a[KC][4];
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    // bins past the tracer's end bin do not enter its sum
                    const int k = 32 * (KC * ch + kk) + 8 * m + 2 * q;
                    a[kk][m].x = k < je ? nxt[kk][m].x : 0.0;
                    a[kk][m].y = k + 1 < je ? nxt[kk][m].y : 0.0;
                }
            if (ch + 1 < NCH) {
#pragma unroll
                for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                    for (int m = 0; m < 4; ++m)
                        nxt[kk][m] = live ? *reinterpret_cast<const double2 *>(row + 32 * (KC * (ch + 1) + kk) + 8 * m)
                                          : make_double2(0.0, 0.0);
            }
            int e = NONE;
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const double v0 = fabs(a[kk][m].x), v1 = fabs(a[kk][m].y);
                    if (v0 > 0.0 && isfinite(v0)) e = max(e, __builtin_amdgcn_frexp_exp(v0));
                    if (v1 > 0.0 && isfinite(v1)) e = max(e, __builtin_amdgcn_frexp_exp(v1));
                }
            e = max(e, __shfl_xor(e, 16, 64));
            e = max(e, __shfl_xor(e, 32, 64));
            const int s_new = max(s_row, e);
            if (ch > 0) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int rr = 4 * q + j;
                    const int so = __shfl(s_row, rr, 64), sn = __shfl(s_new, rr, 64);
                    if (so != sn && so != NONE) {
                        const float f = ldexpf(1.0f, max(so - sn, -160));
#pragma unroll
                        for (int c = 0; c < 4; ++c) acc[c][j] *= f;
                    }
                }
            }
            s_row = s_new;
            const int sh = s_row == NONE ? 0 : s_row - EM_SY_OFF;
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) {
                const int ks = KC * ch + kk;
                bf16x8 ah, al;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const float f0 = (float)ldexp(a[kk][m].x, -sh), f1 = (float)ldexp(a[kk][m].y, -sh);
                    const uint16_t h0 = em_bf16_bits(f0), h1 = em_bf16_bits(f1);
                    ah[2 * m] = __builtin_bit_cast(__bf16, h0);
                    ah[2 * m + 1] = __builtin_bit_cast(__bf16, h1);
                    al[2 * m] = __builtin_bit_cast(__bf16, em_bf16_bits(f0 - em_bf16_val(h0)));
                    al[2 * m + 1] = __builtin_bit_cast(__bf16, em_bf16_bits(f1 - em_bf16_val(h1)));
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(&b_hi[ks][q][16 * c + r][0]);
                    const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(&b_lo[ks][q][16 * c + r][0]);
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[c], 0, 0, 0);
                }
            }
        }
        const int sh = s_row == NONE ? 0 : s_row - EM_SY_OFF;
        // D layout: lane holds rows 4 (lane / 16) + j, column lane % 16; a
        // column is (frequency cb * 8 + (16 c + r) / 8, Chebyshev term r % 8)
        const int mt = r & 7;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int rr = 4 * q + j;
            const int shr = __shfl(sh, rr, 64);
            const bool lv = tile * 16 + rr < cnt;
            const int itr = lv ? perm[start + tile * 16 + rr] : 0;
            const double sr = lv ? sv[itr] : 0.0;
            // T_mt(s) by the recurrence
            double t0 = 1.0, t1 = sr, tm = (mt == 0) ? 1.0 : sr;
            for (int m = 2; m <= mt; ++m) {
                tm = 2.0 * sr * t1 - t0;
                t0 = t1;
                t1 = tm;
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int col = cb * EM_TB + 16 * c + r;
                double v = 0.0;
                if (col < ncol) v = ldexp((double)acc[c][j], shr + col_shift[pbase + col] - EM_SY_OFF) * tm;
                v += __shfl_xor(v, 1, 64);
                v += __shfl_xor(v, 2, 64);
                v += __shfl_xor(v, 4, 64);
                const int f = col / EM_SY_M;
                if (lv && mt == 0 && f < nfreq) eps[(size_t)itr * nfreq + f] = pref[itr] * v;
            }
        }
    }
}

// Range check of the MFMA contraction.  An operand entry below 2^-126 after
// scaling (or a product below it) is lost, so each column sum of tracer i
// misses at most 256 x 2^-67 x 2^(srow_i + cs - 118) of its true value (cs:
// the column's binary exponent), and eps_f at most
// |A| 8 2^(srow_i + max_m cs(f, m) - 177).  Where that bound exceeds 2^-17
// (1e-5) of an emissivity within 25 decades of the tracer's largest one (or
// the row has content and every emissivity came out zero), the tracer is
// listed for the FP32 log-domain kernel, which has no range limit.
__global__ __launch_bounds__(256) void k_synch_check(int ntr, int nfreq, int nlogb, double logb_min,
                                                     double inv_dlogb, const double *__restrict__ b_field,
                                                     const int *__restrict__ srow, const double *__restrict__ pref,
                                                     const int *__restrict__ col_shift,
                                                     const double *__restrict__ eps, int *__restrict__ fix)
{
    const int lane = threadIdx.x & 63;
    const int it = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (it >= ntr) return;
    const int sr = srow[it];
    if (sr == EM_SY_NONE) return;
    const int pl = em_synch_plane(b_field[it], logb_min, inv_dlogb, nlogb);
    const double *e = eps + (size_t)it * nfreq;
    const double ap = fabs(pref[it]);
    double mx = 0.0;
    for (int f = lane; f < nfreq; f += 64) mx = fmax(mx, fabs(e[f]));
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    bool bad = false;
    for (int f = lane; f < nfreq; f += 64) {
        const int *cs = col_shift + ((size_t)pl * nfreq + f) * EM_SY_M;
        int cm = EM_SY_NONE;
#pragma unroll
        for (int m = 0; m < EM_SY_M; ++m)
            if (cs[m] > -2000) cm = max(cm, cs[m]);
        if (cm == EM_SY_NONE) continue;
        const double lim = ldexp(ap, sr + cm - 157);
        if (lim > 0.0 && lim >= 1.0e-25 * mx && fabs(e[f]) < lim) bad = true;
    }
    if (__ballot(bad) && lane == 0) fix[1 + atomicAdd(fix, 1)] = it;
}

extern "C" int ck_synch_mfma_prepare(int n, int nfreq, int nlogb, const double *table, const double *dp,
                                     uint16_t *kt_hi, uint16_t *kt_lo, int *col_shift, float *lt, void *stream)
{
    if (!(n == 64 || n == 128 || n == 256) || nfreq <= 0 || nlogb < 2) return -1;
    hipStream_t st = (hipStream_t)stream;
    const size_t tot = (size_t)(nlogb - 1) * nfreq * n;
    const size_t ncs = (size_t)(nlogb - 1) * nfreq * EM_SY_M;
    const size_t ttot = (size_t)nlogb * nfreq * n;
    if (ck_fill_i32(col_shift, -100000, ncs, stream)) return -1;
    hipLaunchKernelGGL(k_synch_cheb, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, 0, n, nfreq, nlogb, table,
                       dp, kt_hi, kt_lo, col_shift);
    hipLaunchKernelGGL(k_synch_cheb, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, 1, n, nfreq, nlogb, table,
                       dp, kt_hi, kt_lo, col_shift);
    // float log2 table of the FP32 log-domain kernel (the range fix-up)
    hipLaunchKernelGGL(k_synch_log2t32, dim3((unsigned)((ttot + 255) / 256)), dim3(256), 0, st, nlogb, nfreq, n,
                       table, lt);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int ck_synch_mfma_run(int ntr, int n, int nfreq, int nlogb, double logb_min, double inv_dlogb,
                                 const double *table, const uint16_t *kt_hi, const uint16_t *kt_lo,
                                 const int *col_shift, const float *lt, int *work, int *jend, double *sv,
                                 double *pref, int *srow, int *fix, const double *cre, const double *dp,
                                 const double *b_field, double *eps, void *stream)
{
    if (ntr <= 0 || !(n == 64 || n == 128 || n == 256) || nfreq <= 0 || nlogb < 2) return -1;
    hipStream_t st = (hipStream_t)stream;
    const dim3 g1((ntr + 255) / 256), gw((unsigned)((ntr + 3) / 4));
    int *key = work, *hist = work + ntr, *perm = work + ntr + nlogb;
    if (ck_fill_i32(hist, 0, (size_t)nlogb, stream) || ck_fill_i32(fix, 0, 1, stream)) return -1;
    hipLaunchKernelGGL(k_synch_plane_count, g1, dim3(256), 0, st, ntr, b_field, logb_min, inv_dlogb, nlogb, key, hist);
    hipLaunchKernelGGL(k_synch_plane_scan, dim3(1), dim3(256), 0, st, nlogb, hist);
    hipLaunchKernelGGL(k_synch_plane_place, g1, dim3(256), 0, st, ntr, nlogb, key, hist, perm);
    hipLaunchKernelGGL(k_synch_prep, gw, dim3(256), 0, st, ntr, n, nlogb, logb_min, inv_dlogb, cre, b_field, jend,
                       sv, pref, srow);
    // after the placement hist[p] is the end of plane interval p in perm
    const int ncb = (nfreq * EM_SY_M + EM_TB - 1) / EM_TB;
    const int npg = (nlogb - 1 + 7) / 8;
    const dim3 gm((unsigned)(npg * ncb * 8));
#define EM_SY(KS_)                                                                                                     \
    hipLaunchKernelGGL((k_synch_mfma<KS_>), gm, dim3(256), 0, st, nfreq, nlogb, ncb, cre, perm, hist, jend, sv, pref,  \
                       kt_hi, kt_lo, col_shift, eps)
    if (n == 64) EM_SY(2);
    else if (n == 128) EM_SY(4);
    else EM_SY(8);
#undef EM_SY
    // tracers beyond the contraction's range: the FP32 log-domain kernel
    hipLaunchKernelGGL(k_synch_check, gw, dim3(256), 0, st, ntr, nfreq, nlogb, logb_min, inv_dlogb, b_field, srow,
                       pref, col_shift, eps, fix);
    hipLaunchKernelGGL(k_emis_synch32, gw, dim3(256), 0, st, ntr, n, nfreq, nlogb, logb_min, inv_dlogb, table, lt,
                       (const int *)(fix + 1), cre, dp, b_field, eps, (const int *)fix);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int ck_emission_synch(int ntr, int n, int nfreq, int nlogb, double logb_min, double inv_dlogb,
                                 const double *table, double *logt, int *work, const double *cre, const double *dp,
                                 const double *b_field, double *eps, int fp64, void *stream)
{
    if (ntr <= 0 || n <= 3 || n > 512 || nfreq <= 0 || nlogb < 2) return -1;
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((ntr + 3) / 4), g1((ntr + 255) / 256);
    int *perm = nullptr;
    if (work) {
        int *key = work, *hist = work + ntr;
        perm = work + ntr + nlogb;
        if (ck_fill_i32(hist, 0, (size_t)nlogb, stream)) return -1;
        hipLaunchKernelGGL(k_synch_plane_count, g1, dim3(256), 0, st, ntr, b_field, logb_min, inv_dlogb, nlogb, key,
                           hist);
        hipLaunchKernelGGL(k_synch_plane_scan, dim3(1), dim3(256), 0, st, nlogb, hist);
        hipLaunchKernelGGL(k_synch_plane_place, g1, dim3(256), 0, st, ntr, nlogb, key, hist, perm);
    }
    const size_t tot = (size_t)nlogb * nfreq * n;
    if (fp64 && logt)
        hipLaunchKernelGGL(k_synch_logt, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, nlogb, nfreq, n,
                           table, logt);
    if (!fp64) {
        if (!logt) return -1;
        float *lt = (float *)logt;
        hipLaunchKernelGGL(k_synch_log2t32, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, nlogb, nfreq, n,
                           table, lt);
        hipLaunchKernelGGL(k_emis_synch32, grid, dim3(256), 0, st, ntr, n, nfreq, nlogb, logb_min, inv_dlogb, table,
                           (const float *)lt, (const int *)perm, cre, dp, b_field, eps, nullptr);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    hipLaunchKernelGGL(k_emis_synch<double>, grid, dim3(256), 0, st, ntr, n, nfreq, nlogb, logb_min, inv_dlogb,
                       table, (const double *)logt, (const int *)perm, cre, dp, b_field, eps);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
