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

// Lane permutations inside 16-lane rows by DPP (every lane has a source):
// EM_XOR1 / EM_XOR2 swap lanes l ^ 1 / l ^ 2 (quad_perm), EM_HMIRROR maps
// lane i of an 8-lane half row to 7 - i (pairs its two quads), EM_MIRROR
// lane i of the row to 15 - i (pairs its two halves).
#define EM_XOR1 0xB1
#define EM_XOR2 0x4E
#define EM_HMIRROR 0x141
#define EM_MIRROR 0x140
template <int CTRL>
__device__ __forceinline__ double em_dpp64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

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
                    // (hardware round-to-nearest-even float -> bf16)
                    const float f0 = (float)ldexp(a[kk][m].x, -sh), f1 = (float)ldexp(a[kk][m].y, -sh);
                    const __bf16 h0 = (__bf16)f0, h1 = (__bf16)f1;
                    ah[2 * m] = h0;
                    ah[2 * m + 1] = h1;
                    if (SPLIT) {
                        al[2 * m] = (__bf16)(f0 - (float)h0);
                        al[2 * m + 1] = (__bf16)(f1 - (float)h1);
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
// A block takes a chunk of up to EM_SORT_J x 256 consecutive tracers, so that
// its flush of the histogram (one pass over the planes) serves many tracers.
#define EM_LDS_HIST 2048
#define EM_SORT_J 8
__global__ __launch_bounds__(256) void k_synch_plane_count(int ntr, int chunk, const double *__restrict__ b_field,
                                                           double logb_min, double inv_dlogb, int nlogb,
                                                           int *__restrict__ key, int *__restrict__ hist)
{
    __shared__ int s_hist[EM_LDS_HIST];
    const bool local = nlogb <= EM_LDS_HIST;
    const int i0 = blockIdx.x * chunk, i1 = min(ntr, i0 + chunk);
    if (local) {
        for (int b = threadIdx.x; b < nlogb; b += 256) s_hist[b] = 0;
        __syncthreads();
    }
    for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) {
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

// Tracers per block of the two per-tracer sort kernels: about two blocks per
// CU, at most EM_SORT_J x 256.
static int em_sort_chunk(int ntr)
{
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
            ncu = 256;
    }
    int j = (ntr + 512 * ncu - 1) / (512 * ncu);
    if (j < 1) j = 1;
    if (j > EM_SORT_J) j = EM_SORT_J;
    return 256 * j;
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

__global__ __launch_bounds__(256) void k_synch_plane_place(int ntr, int chunk, int nlogb, const int *__restrict__ key,
                                                           int *__restrict__ hist, int *__restrict__ perm)
{
    __shared__ int s_hist[EM_LDS_HIST];
    const int i0 = blockIdx.x * chunk, i1 = min(ntr, i0 + chunk);
    if (nlogb > EM_LDS_HIST) {
        for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) perm[atomicAdd(&hist[key[i]], 1)] = i;
        return;
    }
    for (int b = threadIdx.x; b < nlogb; b += 256) s_hist[b] = 0;
    __syncthreads();
    // the thread's tracers i0 + threadIdx.x + 256 u: plane and rank in the
    // block's part of the plane
    int k[EM_SORT_J], rank[EM_SORT_J];
#pragma unroll
    for (int u = 0; u < EM_SORT_J; ++u) {
        const int i = i0 + (int)threadIdx.x + 256 * u;
        const bool in = 256 * u < chunk && i < i1;
        k[u] = in ? key[i] : 0;
        rank[u] = in ? atomicAdd(&s_hist[k[u]], 1) : 0;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nlogb; b += 256) {
        const int c = s_hist[b];
        if (c) s_hist[b] = atomicAdd(&hist[b], c);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < EM_SORT_J; ++u) {
        const int i = i0 + (int)threadIdx.x + 256 * u;
        if (256 * u < chunk && i < i1) perm[s_hist[k[u]] + rank[u]] = i;
    }
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
// One tracer of the FP32 log-domain sum (wave w of the block, LDS rows s_lw /
// s_sg [w]).
__device__ __forceinline__ void em_synch32_one(int it, int lane, int w, int n, int nfreq, int nlogb, double logb_min,
                                               double inv_dlogb, const double *__restrict__ table,
                                               const float *__restrict__ lt, const double *__restrict__ cre,
                                               const double *__restrict__ dp, const double *__restrict__ b_field,
                                               double *__restrict__ eps, float (*s_lw)[512], float (*s_sg)[512])
{
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

template <bool LOOP>
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
    // count (device): the number of tracers listed in perm, when the list
    // was made on the device; with LOOP the waves of a smaller grid loop
    // over the slots, else one slot per wave
    const int nslot = count ? *count : ntr;
    if constexpr (!LOOP) {
        const int slot = blockIdx.x * 4 + w;
        if (slot >= nslot) return;
        em_synch32_one(perm ? perm[slot] : slot, lane, w, n, nfreq, nlogb, logb_min, inv_dlogb, table, lt, cre, dp,
                       b_field, eps, s_lw, s_sg);
    } else {
        for (int slot = blockIdx.x * 4 + w; slot < nslot; slot += (int)gridDim.x * 4) {
            em_synch32_one(perm ? perm[slot] : slot, lane, w, n, nfreq, nlogb, logb_min, inv_dlogb, table, lt, cre,
                           dp, b_field, eps, s_lw, s_sg);
            // the next slot rewrites this wave's LDS rows
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
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

// Pass 0: the largest binary exponent of every frequency's 8 columns
// (p, f, m = 0..7) over their bins (atomic max); pass 1: the entries scaled
// by that power of two (one scale per frequency, so the Chebyshev sum of the
// columns needs no per-column rescaling), split into bf16 hi / lo.  One
// thread per (p, f, j); bins 0 and 1 are zero (the sum starts at bin 2).
// col_shift [p][f].  kt holds the columns in blocks of 64 in the order the
// contraction keeps them in LDS: frequencies in groups of 16, each group as
// two blocks (terms 0..3, then 4..7), block column 16 c + r = frequency
// 16 g + r, term 4 h + c; inside a block kt [ks][q][column][sl] for bin
// j = 32 ks + 8 (sl / 2) + 2 q + sl % 2 (the operand k order of
// v_mfma_f32_16x16x32_bf16 permuted alike for both operands), so a block
// copies one column block with 16-byte loads.  Frequencies past nfreq stay
// zero.
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
    int *cs = col_shift + (size_t)pl * nfreq + f;
    if (pass == 0) {
        int e = -100000;
#pragma unroll
        for (int m = 0; m < EM_SY_M; ++m) {
            const double v = fabs(c[m]);
            if (v > 0.0 && isfinite(v)) e = max(e, __builtin_amdgcn_frexp_exp(v));
        }
        if (e > -100000) atomicMax(cs, e);
        return;
    }
    const int sh = *cs < -2000 ? 0 : *cs - EM_SY_OFF;
    const int nblk = 2 * ((nfreq + 15) / 16);
    const int kk = j & 31;
    const size_t kofs = ((size_t)(j >> 5) * 4 + ((kk & 7) >> 1)) * (EM_TB * 8) + 2 * (kk >> 3) + (kk & 1);
#pragma unroll
    for (int m = 0; m < EM_SY_M; ++m) {
        const float v = (float)ldexp(isfinite(c[m]) ? c[m] : 0.0, -sh);
        const uint16_t h = em_bf16_bits(v);
        const int blk = 2 * (f / 16) + m / 4, col = 16 * (m % 4) + f % 16;
        const size_t g = ((size_t)pl * nblk + blk) * (size_t)(n / 32) * 4 * EM_TB * 8 + kofs + (size_t)col * 8;
        kt_hi[g] = h;
        kt_lo[g] = em_bf16_bits(v - em_bf16_val(h));
    }
}

#define EM_SY_NONE (-100000)

// Work items of the fused contraction: plane interval p belongs to XCD p % 8
// (blockIdx % 8 of the persistent grid) and holds ceil(count_p / tpb) items;
// ioff[p] = the items of the XCD's earlier planes, itot[x] = the XCD's
// total.  One wave per XCD.
__global__ __launch_bounds__(512) void k_synch_items(int nlogb, int tpb, const int *__restrict__ plane_end,
                                                     int *__restrict__ ioff, int *__restrict__ itot)
{
    const int lane = threadIdx.x & 63, x = threadIdx.x >> 6;
    const int np = nlogb - 1;
    const int npx = np > x ? (np - x + 7) / 8 : 0;
    int carry = 0;
    for (int base = 0; base < npx; base += 64) {
        const int i = base + lane;
        const int p = x + 8 * i;
        const int c = i < npx ? plane_end[p] - (p ? plane_end[p - 1] : 0) : 0;
        const int items = (c + tpb - 1) / tpb;
        int incl = items;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        if (i < npx) ioff[p] = carry + incl - items;
        carry += __shfl(incl, 63, 64);
    }
    if (lane == 0) itot[x] = carry;
}

// The plane-grouped contraction, one pass over the spectra.  A persistent
// block (XCD x = blockIdx % 8) of NW waves takes work items of x's planes:
// 16 TPW NW tracers of one plane interval p, TPW 16-tracer tiles per wave.
// Per item:
//  1. each wave reads its tiles' rows, finds each row's end bin (the first
//     bin from 3 on that is not positive ends the sum at the bin before: the
//     sum runs over bins 2 .. jend - 1, as k_emis_synch32) and its largest
//     binary exponent over the bins before the end, and keeps the rows in
//     registers as split BF16 operands (scaled to a largest entry of 2^59,
//     masked past the end);
//  2. the block copies plane p's column blocks one after the other into
//     LDS; every wave runs its tiles against each with
//     v_mfma_f32_16x16x32_bf16 (split x3).  Two blocks make a group of 16
//     frequencies: lane (r, q) then holds all 8 Chebyshev terms of frequency
//     16 g + r for rows 4 q + j, and writes
//     eps[i][f] = A_i 2^(scales) sum_m T_m(s_i) G[i][f, m] with no exchange
//     between lanes;
//  3. the range check, on the tracer's own emissivities: an operand entry
//     below 2^-126 after scaling (or a product below it) is lost, so each
//     column sum of tracer i misses at most 256 x 2^-67 x 2^(srow_i + cs - 118)
//     of its true value (cs: the frequency's binary exponent) and eps_f at
//     most lim_f = |A| 8 2^(srow_i + cs_f - 177); where lim_f exceeds 2^-17
//     (1e-5) of an emissivity within 25 decades of the tracer's largest one
//     (or the row has content and every emissivity came out zero), the tracer
//     is listed for the FP32 log-domain kernel, which has no range limit.
// Planes are spread over XCDs by p % 8, so a plane's column blocks (n x 512
// x 4 bytes) are read from HBM once and then served by that XCD's L2.
template <int KS, int TPW, int NW>
__global__ __launch_bounds__(64 * NW, 2) void k_synch_fused(int nfreq, int nlogb, double logb_min, double inv_dlogb,
                                                            const double *__restrict__ cre,
                                                            const double *__restrict__ b_field,
                                                            const int *__restrict__ perm,
                                                            const int *__restrict__ plane_end,
                                                            const int *__restrict__ ioff,
                                                            const int *__restrict__ itot,
                                                            const uint16_t *__restrict__ kt_hi,
                                                            const uint16_t *__restrict__ kt_lo,
                                                            const int *__restrict__ col_shift,
                                                            double *__restrict__ eps, int *__restrict__ fix)
{
    constexpr int K = 32 * KS;
    constexpr int CH = KS * 4 * EM_TB * 8; // bf16 per half of one column block
    constexpr int NONE = EM_SY_NONE;
    constexpr int NT = 64 * NW;            // threads
    constexpr int TPB = 16 * TPW * NW;     // tracers per item
    constexpr int HC = CH / 2; // bf16 per half of one half block (32 columns)
    static_assert(HC / 8 >= NT && (HC / 8) % NT == 0, "every thread issues whole 16-byte loads of a half block");
    // two half-block buffers: one is filled while the other multiplies
    __shared__ __attribute__((aligned(16))) uint16_t b_hi[2][HC];
    __shared__ __attribute__((aligned(16))) uint16_t b_lo[2][HC];
    __shared__ __attribute__((aligned(16))) double s_t[TPB][EM_SY_M]; // T_m(s) of each tracer
    __shared__ double s_a[TPB];
    __shared__ int s_it[TPB], s_sr[TPB];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int x = blockIdx.x & 7, nbx = gridDim.x >> 3;
    const int np = nlogb - 1;
    const int npx = np > x ? (np - x + 7) / 8 : 0;
    const int nitems = itot[x];
    const int ngrp = (nfreq + 15) / 16;
    for (int item = blockIdx.x >> 3; item < nitems; item += nbx) {
        // the item's plane: the last of x's planes whose first item is <= item
        int lo = 0, hi = npx - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ioff[x + 8 * mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        const int p = x + 8 * lo;
        const int pend = plane_end[p];
        const int t0 = (p ? plane_end[p - 1] : 0) + (item - ioff[p]) * TPB;
        // 1. rows -> split BF16 operands
        bf16x8 ah[TPW][KS], al[TPW][KS];
#pragma unroll
        for (int tt = 0; tt < TPW; ++tt) {
            const int loc = 16 * (TPW * w + tt) + r;
            // past the plane's end a lane reads the plane's last tracer (in
            // bounds) and masks the whole row
            const bool live = t0 + loc < pend;
            const int it = perm[min(t0 + loc, pend - 1)];
            // lane (r, q) holds bins 32 ks + 8 m + 2 q (+1) of row r, all read
            // at once (one memory round trip per tile)
            const double *row = cre + (size_t)it * K + 2 * q;
            double2 v[KS][4];
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
#pragma unroll
                for (int m = 0; m < 4; ++m) v[ks][m] = *reinterpret_cast<const double2 *>(row + 32 * ks + 8 * m);
            // bins as compile-time offsets o from the lane's first bin 2 q
            // (k = o + 2 q), so no per-bin index occupies a register
            const int q2 = 2 * q;
            // the end of the sum: the first non-positive bin from 3 on
            // (offsets 0, 1 are bins 2 q, 2 q + 1, from 3 on only for q >= 1)
            int kfo = K;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int o = 32 * ks + 8 * m;
                    if ((o >= 3 || o + q2 >= 3) && !(v[ks][m].x > 0.0)) kfo = min(kfo, o);
                    if ((o + 1 >= 3 || o + 1 + q2 >= 3) && !(v[ks][m].y > 0.0)) kfo = min(kfo, o + 1);
                }
            int kf = kfo + q2;
            if (!live) kf = 0;
            kf = min(kf, __shfl_xor(kf, 16, 64));
            kf = min(kf, __shfl_xor(kf, 32, 64));
            const int jend = kf - 1;
            // a bin of the lane enters the sum if its offset is below jr
            const int jr = jend - q2;
            // the largest exponent of the bins before the end
            int e = NONE;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int o = 32 * ks + 8 * m;
                    const double v0 = fabs(v[ks][m].x), v1 = fabs(v[ks][m].y);
                    if (o < jr && v0 > 0.0 && isfinite(v0)) e = max(e, __builtin_amdgcn_frexp_exp(v0));
                    if (o + 1 < jr && v1 > 0.0 && isfinite(v1)) e = max(e, __builtin_amdgcn_frexp_exp(v1));
                }
            e = max(e, __shfl_xor(e, 16, 64));
            e = max(e, __shfl_xor(e, 32, 64));
            // the split BF16 operands, masked past the end
            const int sh = e == NONE ? 0 : e - EM_SY_OFF;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                bf16x8 h8, l8;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    // (the float -> bf16 conversions round to nearest even in
                    // hardware, v_cvt_pk_bf16_f32)
                    const int o = 32 * ks + 8 * m;
                    const float f0 = o < jr ? (float)ldexp(v[ks][m].x, -sh) : 0.0f;
                    const float f1 = o + 1 < jr ? (float)ldexp(v[ks][m].y, -sh) : 0.0f;
                    const __bf16 h0 = (__bf16)f0, h1 = (__bf16)f1;
                    h8[2 * m] = h0;
                    h8[2 * m + 1] = h1;
                    l8[2 * m] = (__bf16)(f0 - (float)h0);
                    l8[2 * m + 1] = (__bf16)(f1 - (float)h1);
                }
                ah[tt][ks] = h8;
                al[tt][ks] = l8;
            }
            if (q == 0) {
                // per tracer: s = 2 t - 1 of its field in the plane
                // interval, T_0 .. T_7 of s, the prefactor A(B)
                double sv = 0.0, A = 0.0;
                if (live) {
                    double bval = b_field[it];
                    if (!(bval > 0.0) || !isfinite(bval)) bval = 1.0e-30;
                    const double u = (log10(bval / 1.0e-6) - logb_min) * inv_dlogb;
                    int ib = (int)floor(u);
                    if (ib < 0) ib = 0;
                    if (ib > nlogb - 2) ib = nlogb - 2;
                    double tt2 = u - (double)ib;
                    if (tt2 < 0.0) tt2 = 0.0;
                    if (tt2 > 1.0) tt2 = 1.0;
                    sv = 2.0 * tt2 - 1.0;
                    A = sqrt(3.0) * pow(4.8e-10, 3.0) * b_field[it] / (2.0 * 9.11e-28 * 2.999e10 * 2.999e10) /
                        (4.0 * 3.14159265358979323846);
                }
                double c0 = 1.0, c1 = sv;
                s_t[loc][0] = 1.0;
                s_t[loc][1] = sv;
#pragma unroll
                for (int m = 2; m < EM_SY_M; ++m) {
                    const double cm = 2.0 * sv * c1 - c0;
                    c0 = c1;
                    c1 = cm;
                    s_t[loc][m] = cm;
                }
                s_a[loc] = A;
                s_it[loc] = live ? it : -1;
                s_sr[loc] = e;
            }
            // keep the second tile's loads after the first tile's conversion
            // (the raw rows of both tiles at once would not fit the
            // registers)
            __builtin_amdgcn_sched_barrier(0);
        }
        // the range check's running largest emissivity / largest violated
        // bound of the lane's output rows (rows 4 q + j of each tile, the
        // lane's frequencies)
        double rmx[TPW][4], rlm[TPW][4];
#pragma unroll
        for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
            for (int j = 0; j < 4; ++j) rmx[tt][j] = rlm[tt][j] = 0.0;
        // 2. the plane's column blocks, two per group of 16 frequencies, as
        // half blocks of 32 columns (hh = 4 grp + 2 hf + hc: group, block of
        // the group, half of the block) through the two LDS buffers.  A half
        // block goes straight from global memory into LDS (global_load_lds,
        // 16 bytes per lane: a wave's lanes fill 1 KB from the wave's LDS
        // base): unit U = (ks * 4 + q) * 32 + col of the half (its 32
        // columns of one (ks, q) are 512 contiguous bytes of the block).
        const uint4 *gh = reinterpret_cast<const uint4 *>(kt_hi + (size_t)p * 2 * ngrp * CH);
        const uint4 *gl = reinterpret_cast<const uint4 *>(kt_lo + (size_t)p * 2 * ngrp * CH);
        const int nhh = 4 * ngrp;
        auto issue = [&](int hh) {
            const size_t base = (size_t)(hh >> 1) * (CH / 8) + 32 * (hh & 1);
#pragma unroll
            for (int u = 0; u < HC / 8 / NT; ++u) {
                const int U = t + NT * u;
                const size_t gi = base + (size_t)(U >> 5) * 64 + (U & 31);
                __builtin_amdgcn_global_load_lds(
                    gh + gi, (__attribute__((address_space(3))) void *)(reinterpret_cast<uint4 *>(b_hi[hh & 1]) + NT * u + 64 * w),
                    16, 0, 0);
                __builtin_amdgcn_global_load_lds(
                    gl + gi, (__attribute__((address_space(3))) void *)(reinterpret_cast<uint4 *>(b_lo[hh & 1]) + NT * u + 64 * w),
                    16, 0, 0);
            }
        };
        __syncthreads(); // every wave is done with the previous item's buffers
        issue(0);
        for (int grp = 0; grp < ngrp; ++grp) {
            floatx4 acc[TPW][8]; // [tile][term]
#pragma unroll
            for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
                for (int c = 0; c < 8; ++c) acc[tt][c] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int hh = 4 * grp + s4;
                // half block hh has landed and every wave is done with hh - 1,
                // whose buffer the next half block then fills
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
                if (hh + 1 < nhh) issue(hh + 1);
                const uint16_t *bhh = b_hi[hh & 1], *bll = b_lo[hh & 1];
                // the two column tiles of the half: the B fragments of step
                // ks + 1 are read while step ks multiplies, the two tiles'
                // products interleaved (two independent accumulator chains);
                // the scheduling barriers keep the compiler from hoisting
                // further reads, which would not fit the registers beside
                // the operands
                bf16x8 bh[2][2], bl[2][2]; // [buffer][tile]
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    bh[0][c] = *reinterpret_cast<const bf16x8 *>(&bhh[(q * 32 + 16 * c + r) * 8]);
                    bl[0][c] = *reinterpret_cast<const bf16x8 *>(&bll[(q * 32 + 16 * c + r) * 8]);
                }
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    if (ks + 1 < KS) {
#pragma unroll
                        for (int c = 0; c < 2; ++c) {
                            const int o = (((ks + 1) * 4 + q) * 32 + 16 * c + r) * 8;
                            bh[(ks + 1) & 1][c] = *reinterpret_cast<const bf16x8 *>(&bhh[o]);
                            bl[(ks + 1) & 1][c] = *reinterpret_cast<const bf16x8 *>(&bll[o]);
                        }
                    }
#pragma unroll
                    for (int tt = 0; tt < TPW; ++tt) {
                        floatx4 &a0 = acc[tt][2 * s4], &a1 = acc[tt][2 * s4 + 1];
                        a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[tt][ks], bh[ks & 1][0], a0, 0, 0, 0);
                        a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[tt][ks], bh[ks & 1][1], a1, 0, 0, 0);
                        a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[tt][ks], bl[ks & 1][0], a0, 0, 0, 0);
                        a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[tt][ks], bl[ks & 1][1], a1, 0, 0, 0);
                        a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[tt][ks], bh[ks & 1][0], a0, 0, 0, 0);
                        a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[tt][ks], bh[ks & 1][1], a1, 0, 0, 0);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            // D layout: lane (r, q) holds rows 4 q + j, column 16 c + r of
            // each half block: frequency 16 grp + r, term 2 s4 + c
            const int f = 16 * grp + r;
            const int cs = f < nfreq ? col_shift[(size_t)p * nfreq + f] : NONE;
#pragma unroll
            for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    // per-tracer values of output row 4 q + j (written in
                    // phase 1 by this wave; the barriers since order them)
                    const int loc = 16 * (TPW * w + tt) + 4 * q + j;
                    const double ra = s_a[loc];
                    const int rit = s_it[loc], rsr = s_sr[loc];
                    const double4 ta = *reinterpret_cast<const double4 *>(&s_t[loc][0]);
                    const double4 tb = *reinterpret_cast<const double4 *>(&s_t[loc][4]);
                    double v = ta.x * (double)acc[tt][0][j];
                    v = fma(ta.y, (double)acc[tt][1][j], v);
                    v = fma(ta.z, (double)acc[tt][2][j], v);
                    v = fma(ta.w, (double)acc[tt][3][j], v);
                    v = fma(tb.x, (double)acc[tt][4][j], v);
                    v = fma(tb.y, (double)acc[tt][5][j], v);
                    v = fma(tb.z, (double)acc[tt][6][j], v);
                    v = fma(tb.w, (double)acc[tt][7][j], v);
                    const int shr = rsr == NONE ? 0 : rsr - EM_SY_OFF;
                    const bool ok = f < nfreq && cs > -2000;
                    const double ev = ok ? ra * ldexp(v, shr + cs - EM_SY_OFF) : 0.0;
                    if (f < nfreq && rit >= 0) eps[(size_t)rit * nfreq + f] = ev;
                    const double lim = (ok && rsr != NONE) ? ldexp(fabs(ra), rsr + cs - 157) : 0.0;
                    rmx[tt][j] = fmax(rmx[tt][j], fabs(ev));
                    if (lim > 0.0 && fabs(ev) < lim) rlm[tt][j] = fmax(rlm[tt][j], lim);
                }
        }
        // 3. range check of the wave's rows: the largest emissivity and the
        // largest violated bound over the row's 16 lanes
#pragma unroll
        for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                double mx = rmx[tt][j], lm = rlm[tt][j];
                mx = fmax(mx, em_dpp64<EM_XOR1>(mx));
                mx = fmax(mx, em_dpp64<EM_XOR2>(mx));
                mx = fmax(mx, em_dpp64<EM_HMIRROR>(mx));
                mx = fmax(mx, em_dpp64<EM_MIRROR>(mx));
                lm = fmax(lm, em_dpp64<EM_XOR1>(lm));
                lm = fmax(lm, em_dpp64<EM_XOR2>(lm));
                lm = fmax(lm, em_dpp64<EM_HMIRROR>(lm));
                lm = fmax(lm, em_dpp64<EM_MIRROR>(lm));
                const int rit = s_it[16 * (TPW * w + tt) + 4 * q + j];
                if (r == 0 && rit >= 0 && lm > 0.0 && lm >= 1.0e-25 * mx) fix[1 + atomicAdd(fix, 1)] = rit;
            }
    }
}

extern "C" int ck_synch_mfma_prepare(int n, int nfreq, int nlogb, const double *table, const double *dp,
                                     uint16_t *kt_hi, uint16_t *kt_lo, int *col_shift, float *lt, void *stream)
{
    if (!(n == 64 || n == 128 || n == 256) || nfreq <= 0 || nlogb < 2) return -1;
    hipStream_t st = (hipStream_t)stream;
    const size_t tot = (size_t)(nlogb - 1) * nfreq * n;
    const size_t ncs = (size_t)(nlogb - 1) * nfreq;
    const size_t ttot = (size_t)nlogb * nfreq * n;
    // kt: (nlogb - 1) x 2 ceil(nfreq / 16) column blocks of 64 x n bf16; the
    // frequencies past nfreq stay zero
    const size_t nkt = (size_t)(nlogb - 1) * 2 * ((nfreq + 15) / 16) * EM_TB * n;
    if (ck_fill_i32(col_shift, -100000, ncs, stream) || ck_fill_i32((int *)kt_hi, 0, nkt / 2, stream) ||
        ck_fill_i32((int *)kt_lo, 0, nkt / 2, stream))
        return -1;
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
                                 const int *col_shift, const float *lt, int *work, int *fix, const double *cre,
                                 const double *dp, const double *b_field, double *eps, void *stream)
{
    if (ntr <= 0 || !(n == 64 || n == 128 || n == 256) || nfreq <= 0 || nlogb < 2) return -1;
    hipStream_t st = (hipStream_t)stream;
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
            ncu = 256;
    }
    const int chunk = em_sort_chunk(ntr);
    const dim3 g1((ntr + chunk - 1) / chunk);
    int *key = work, *hist = work + ntr, *perm = work + ntr + nlogb;
    int *ioff = perm + ntr, *itot = ioff + nlogb;
    if (ck_fill_i32(hist, 0, (size_t)nlogb, stream) || ck_fill_i32(fix, 0, 1, stream)) return -1;
    hipLaunchKernelGGL(k_synch_plane_count, g1, dim3(256), 0, st, ntr, chunk, b_field, logb_min, inv_dlogb, nlogb, key,
                       hist);
    hipLaunchKernelGGL(k_synch_plane_scan, dim3(1), dim3(256), 0, st, nlogb, hist);
    hipLaunchKernelGGL(k_synch_plane_place, g1, dim3(256), 0, st, ntr, chunk, nlogb, key, hist, perm);
    // after the placement hist[p] is the end of plane interval p in perm
    const int nbx = ncu / 8 > 0 ? ncu / 8 : 1; // CUs per XCD
    // persistent: two blocks of four waves per CU (the same number per XCD),
    // one 16-tracer tile per wave (the accumulators of all 8 terms of 16
    // frequencies for two tiles would not fit the registers beside their
    // operands)
    hipLaunchKernelGGL(k_synch_items, dim3(1), dim3(512), 0, st, nlogb, 64, hist, ioff, itot);
#define EM_SY(KS_)                                                                                                     \
    hipLaunchKernelGGL((k_synch_fused<KS_, 1, 4>), dim3((unsigned)(16 * nbx)), dim3(256), 0, st, nfreq, nlogb,         \
                       logb_min, inv_dlogb, cre, b_field, perm, hist, ioff, itot, kt_hi, kt_lo, col_shift, eps, fix)
    if (n == 64) EM_SY(2);
    else if (n == 128) EM_SY(4);
    else EM_SY(8);
#undef EM_SY
    // tracers beyond the contraction's range: the FP32 log-domain kernel
    // over the list (waves loop over the listed tracers)
    int gx = (ntr + 3) / 4;
    if (gx > 4 * ncu) gx = 4 * ncu;
    hipLaunchKernelGGL(k_emis_synch32<true>, dim3((unsigned)gx), dim3(256), 0, st, ntr, n, nfreq, nlogb, logb_min,
                       inv_dlogb, table, lt, (const int *)(fix + 1), cre, dp, b_field, eps, (const int *)fix);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int ck_emission_synch(int ntr, int n, int nfreq, int nlogb, double logb_min, double inv_dlogb,
                                 const double *table, double *logt, int *work, const double *cre, const double *dp,
                                 const double *b_field, double *eps, int fp64, void *stream)
{
    if (ntr <= 0 || n <= 3 || n > 512 || nfreq <= 0 || nlogb < 2) return -1;
    hipStream_t st = (hipStream_t)stream;
    const int chunk = em_sort_chunk(ntr);
    const dim3 grid((ntr + 3) / 4), g1((ntr + chunk - 1) / chunk);
    int *perm = nullptr;
    if (work) {
        int *key = work, *hist = work + ntr;
        perm = work + ntr + nlogb;
        if (ck_fill_i32(hist, 0, (size_t)nlogb, stream)) return -1;
        hipLaunchKernelGGL(k_synch_plane_count, g1, dim3(256), 0, st, ntr, chunk, b_field, logb_min, inv_dlogb, nlogb,
                           key, hist);
        hipLaunchKernelGGL(k_synch_plane_scan, dim3(1), dim3(256), 0, st, nlogb, hist);
        hipLaunchKernelGGL(k_synch_plane_place, g1, dim3(256), 0, st, ntr, chunk, nlogb, key, hist, perm);
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
        hipLaunchKernelGGL(k_emis_synch32<false>, grid, dim3(256), 0, st, ntr, n, nfreq, nlogb, logb_min, inv_dlogb, table,
                           (const float *)lt, (const int *)perm, cre, dp, b_field, eps, nullptr);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    hipLaunchKernelGGL(k_emis_synch<double>, grid, dim3(256), 0, st, ntr, n, nfreq, nlogb, logb_min, inv_dlogb,
                       table, (const double *)logt, (const int *)perm, cre, dp, b_field, eps);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
