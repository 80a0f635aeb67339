// czw_kernels.hip -- gfx950 kernels and C ABI (include/czw.h) of the wavelet
// block compressor.
//
// One workgroup of 1024 threads (16 wave64) owns one 32^3 block.  The block
// stays in LDS in its ORIGINAL orientation O[z][y][x] for the whole
// multilevel transform; each 1D pass loads one row per thread into
// registers, transforms it there and stores it back.  The reference instead
// runs every pass along its fastest axis and swaps axes in place
// (WaveletsOnInterval.h:205-276, never swapping back), so its coefficients end
// up in a piecewise axis-permuted "physical" layout.  Keeping the data still
// and changing the pass axis gives bit-identical values (the same rows in the
// same order); the physical layout is only needed for the mask bit positions
// and the survivor stream order, and is applied there by index arithmetic:
//
//   level | forward pass axes | inverse pass axes | physical (a0,a1,a2) of O
//   32    | x, y, z           | z, y, x           | (y, x, z)   outside [0,16)^3
//   16    | z, x, y           | y, x, z           | (x, z, y)   in [0,16)^3 \ [0,8)^3
//    8    | y, z, x           | x, z, y           | (z, y, x)   in [0,8)^3
//
// with physical index p = a2 + 32*(a1 + 32*a0).
//
// LDS layout: O[z][y][x] at z*SP + y*RP + x with RP = 33, SP = 32*33+1, so
// that rows along x (thread = y or z fastest), along y (thread = x fastest)
// and along z (thread = x fastest) all hit 32 distinct banks.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/czw.h"
#include "czw_wavelet.hpp"

namespace czw {

constexpr int NB3 = 32768;
constexpr int MASK_BYTES = 4096;
constexpr int RP = 33;
constexpr int SP = 32 * 33 + 1;
constexpr int LDS_WORDS = 32 * SP;
constexpr int NT = 512;            // threads per block workgroup (8 wave64)
constexpr int NROW = 1024 / NT;    // level-32 rows (and stream segments) per thread
constexpr size_t SLOT32 = MASK_BYTES + 4 * (size_t)NB3;

enum { LAYOUT_BLOCKS = 0, LAYOUT_FIELD = 1 };

template <int N, typename R>
CZW_DI void lds_ld(const R *A, int base, int stride, R *r)
{
#pragma unroll
    for (int k = 0; k < N; ++k) r[k] = A[base + k * stride];
}

template <int N, typename R>
CZW_DI void lds_st(R *A, int base, int stride, const R *r)
{
#pragma unroll
    for (int k = 0; k < N; ++k) A[base + k * stride] = r[k];
}

// Wave64 inclusive prefix sum.
CZW_DI uint32_t wave_incl_scan(uint32_t x, int lane)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

// Block-wide exclusive scan of one value per thread.  Returns the exclusive
// prefix; *total receives the sum.  Ends with the wave totals read, so the
// caller must barrier before wsum is written again.
CZW_DI uint32_t block_excl_scan(uint32_t v, uint32_t *wsum, uint32_t *total)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t inc = wave_incl_scan(v, lane);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) {
        const uint32_t s = wsum[i];
        base += (i < w) ? s : 0u;
        tot += s;
    }
    *total = tot;
    return base + inc - v;
}

// ---------------------------------------------------------------------------
// Stream segment g covers stream indices [32g, 32g+32) (g = t + NT*k for the
// k-th segment of thread t).  Stream order (FullWaveletTransform.h:59-96,
// 163-176): [0,512) is the [0,8)^3 corner as codes 0..7 of boxes of h=4
// (code 0 = the untouched 4^3 base), [512,4096) codes 1..7 of h=8,
// [4096,32768) codes 1..7 of h=16.  Inside a code box (origin
// h*(code&1, code>>1&1, code>>2&1) on (a2,a1,a0)) the order is a0, a1, a2
// with a2 fastest, so a segment is 32/h runs of h consecutive a2.  For each
// run the callback receives the run index q, the run length h, the LDS address
// of its first element, the LDS stride along the run, the mask word p>>5 and
// the bit shift of its first element.
template <typename F>
CZW_DI void for_each_run(int g, F &&f)
{
    if (g < 16) {  // h = 4: coarse base + level-8 details, O = (a0,a1,a2)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int s = 32 * g + 4 * q;
            const int code = s >> 6, k = s & 63;
            const int a0 = 4 * ((code >> 2) & 1) + (k >> 4);
            const int a1 = 4 * ((code >> 1) & 1) + ((k >> 2) & 3);
            const int o2 = 4 * (code & 1);
            f(q, 4, a0 * SP + a1 * RP + o2, 1, a1 + 32 * a0, o2);
        }
    } else if (g < 128) {  // h = 8: level-16 details, O.x=a0, O.z=a1, O.y=a2
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int s = 32 * g + 8 * q;
            const int code = s >> 9, k = s & 511;
            const int a0 = 8 * ((code >> 2) & 1) + (k >> 6);
            const int a1 = 8 * ((code >> 1) & 1) + ((k >> 3) & 7);
            const int o2 = 8 * (code & 1);
            f(q, 8, a1 * SP + o2 * RP + a0, RP, a1 + 32 * a0, o2);
        }
    } else {  // h = 16: level-32 details, O.y=a0, O.x=a1, O.z=a2
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int s = 32 * g + 16 * q;
            const int code = s >> 12, k = s & 4095;
            const int a0 = 16 * ((code >> 2) & 1) + (k >> 8);
            const int a1 = 16 * ((code >> 1) & 1) + ((k >> 4) & 15);
            const int o2 = 16 * (code & 1);
            f(q, 16, o2 * SP + a0 * RP + a1, SP, a1 + 32 * a0, o2);
        }
    }
}

// ---------------------------------------------------------------------------
// Forward level-32 y- and z-passes (the x-pass runs on the rows as loaded),
// with the level-16 z-pass fused into the level-32 z-pass.
template <int WT>
CZW_DI void fwt32_yz_f32(float *A, int t)
{
#pragma unroll
    for (int k = 0; k < NROW; ++k) {  // y-pass: row (z, x)
        const int q = t + NT * k, z = q >> 5, x = q & 31;
        float r[32];
        lds_ld<32>(A, z * SP + x, RP, r);
        row<32, true, WT>(r);
        lds_st<32>(A, z * SP + x, RP, r);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NROW; ++k) {  // z-pass: row (y, x); then level 16 on y, x < 16
        const int q = t + NT * k, y = q >> 5, x = q & 31;
        float r[32];
        lds_ld<32>(A, y * RP + x, SP, r);
        row<32, true, WT>(r);
        if (y < 16 && x < 16) row<16, true, WT>(r);
        lds_st<32>(A, y * RP + x, SP, r);
    }
    __syncthreads();
}

// Remaining forward passes: level 16 (x, y) and level 8 (y fused, z, x).
template <int WT>
CZW_DI void fwt16_8_f32(float *A, int t)
{
    float r[16];
    if (t < 256) {  // level 16 x-pass: row (z, y) in [0,16)^2
        const int z = t >> 4, y = t & 15;
        lds_ld<16>(A, z * SP + y * RP, 1, r);
        row<16, true, WT>(r);
        lds_st<16>(A, z * SP + y * RP, 1, r);
    }
    __syncthreads();
    if (t < 256) {  // level 16 y-pass, then level 8 y-pass: row (z, x)
        const int z = t >> 4, x = t & 15;
        lds_ld<16>(A, z * SP + x, RP, r);
        row<16, true, WT>(r);
        if (z < 8 && x < 8) row<8, true, WT>(r);
        lds_st<16>(A, z * SP + x, RP, r);
    }
    __syncthreads();
    if (t < 64) {  // level 8 z-pass: row (y, x) in [0,8)^2
        const int y = t >> 3, x = t & 7;
        lds_ld<8>(A, y * RP + x, SP, r);
        row<8, true, WT>(r);
        lds_st<8>(A, y * RP + x, SP, r);
    }
    __syncthreads();
    if (t < 64) {  // level 8 x-pass: row (z, y)
        const int z = t >> 3, y = t & 7;
        lds_ld<8>(A, z * SP + y * RP, 1, r);
        row<8, true, WT>(r);
        lds_st<8>(A, z * SP + y * RP, 1, r);
    }
    __syncthreads();
}

// Inverse passes of all levels except the final level-32 x-pass, which the
// caller fuses with the store: level 8 (x, z, y), level 16 (y fused, x, z),
// level 32 (z fused, y).
template <int WT>
CZW_DI void iwt_most_f32(float *A, int t)
{
    {
        float r[16];
        if (t < 64) {  // level 8 x-pass
            const int z = t >> 3, y = t & 7;
            lds_ld<8>(A, z * SP + y * RP, 1, r);
            row<8, false, WT>(r);
            lds_st<8>(A, z * SP + y * RP, 1, r);
        }
        __syncthreads();
        if (t < 64) {  // level 8 z-pass
            const int y = t >> 3, x = t & 7;
            lds_ld<8>(A, y * RP + x, SP, r);
            row<8, false, WT>(r);
            lds_st<8>(A, y * RP + x, SP, r);
        }
        __syncthreads();
        if (t < 256) {  // level 8 y-pass, then level 16 y-pass
            const int z = t >> 4, x = t & 15;
            lds_ld<16>(A, z * SP + x, RP, r);
            if (z < 8 && x < 8) row<8, false, WT>(r);
            row<16, false, WT>(r);
            lds_st<16>(A, z * SP + x, RP, r);
        }
        __syncthreads();
        if (t < 256) {  // level 16 x-pass
            const int z = t >> 4, y = t & 15;
            lds_ld<16>(A, z * SP + y * RP, 1, r);
            row<16, false, WT>(r);
            lds_st<16>(A, z * SP + y * RP, 1, r);
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < NROW; ++k) {  // level 16 z-pass, then level 32 z-pass
        const int q = t + NT * k, y = q >> 5, x = q & 31;
        float r[32];
        lds_ld<32>(A, y * RP + x, SP, r);
        if (y < 16 && x < 16) row<16, false, WT>(r);
        row<32, false, WT>(r);
        lds_st<32>(A, y * RP + x, SP, r);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NROW; ++k) {  // level 32 y-pass
        const int q = t + NT * k, z = q >> 5, x = q & 31;
        float r[32];
        lds_ld<32>(A, z * SP + x, RP, r);
        row<32, false, WT>(r);
        lds_st<32>(A, z * SP + x, RP, r);
    }
    __syncthreads();
}

struct FieldGeom {
    int nx, ny, nz;
};

// Global element offset of block b's row start (x = 0) for the x-row (y, z)
// of an HDF5-ordered field, and the stride along x.
CZW_DI size_t field_row(const FieldGeom &g, int b, int y, int z, size_t *xstride)
{
    const int bx = g.nx >> 5, by = g.ny >> 5;
    const int ix = b % bx, iy = (b / bx) % by, iz = b / (bx * by);
    *xstride = (size_t)g.ny * (size_t)g.nz;
    return ((size_t)(32 * ix) * g.ny + (size_t)(32 * iy + y)) * g.nz + (size_t)(32 * iz + z);
}

// ---------------------------------------------------------------------------
// x-row q (0..1023) of a block: in the block layout q = (z, y) with y
// fastest (16-byte loads along x); in the HDF5 field layout q = (y, z) with z
// fastest, so that each 4-byte load of a wave reads two runs of 32
// consecutive Z of the field.
template <int LAYOUT>
CZW_DI int xrow_lds_base(int q)
{
    if constexpr (LAYOUT == LAYOUT_BLOCKS) return (q >> 5) * SP + (q & 31) * RP;
    else return (q & 31) * SP + (q >> 5) * RP;
}

template <int LAYOUT>
CZW_DI void load_xrow_f32(const float *__restrict__ src, const FieldGeom &g, int b, int q, float *r)
{
    if constexpr (LAYOUT == LAYOUT_BLOCKS) {
        const float4 *p = reinterpret_cast<const float4 *>(src + (size_t)b * NB3 + (size_t)q * 32);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float4 v = p[i];
            r[4 * i] = v.x;
            r[4 * i + 1] = v.y;
            r[4 * i + 2] = v.z;
            r[4 * i + 3] = v.w;
        }
    } else {
        size_t xs;
        const float *p = src + field_row(g, b, q >> 5, q & 31, &xs);
#pragma unroll
        for (int x = 0; x < 32; ++x) r[x] = p[x * xs];
    }
}

template <int LAYOUT>
CZW_DI void store_xrow_f32(float *__restrict__ dst, const FieldGeom &g, int b, int q, const float *r)
{
    if constexpr (LAYOUT == LAYOUT_BLOCKS) {
        float4 *p = reinterpret_cast<float4 *>(dst + (size_t)b * NB3 + (size_t)q * 32);
#pragma unroll
        for (int i = 0; i < 8; ++i) p[i] = make_float4(r[4 * i], r[4 * i + 1], r[4 * i + 2], r[4 * i + 3]);
    } else {
        size_t xs;
        float *p = dst + field_row(g, b, q >> 5, q & 31, &xs);
#pragma unroll
        for (int x = 0; x < 32; ++x) p[x * xs] = r[x];
    }
}

// ---------------------------------------------------------------------------
// Compressor, binary32.  Mirrors WaveletCompressorGeneric::compress
// (WaveletCompressor.cpp:153-182): fwt, threshold + mask, zerobits, shuffle.
// Persistent: workgroup w handles blocks w, w+G, w+2G, ... (G = gridDim.x)
// and loads block b+G into registers while block b is transformed, so the
// HBM reads of one block overlap the arithmetic of the previous one.  Block
// b's payload goes to slots + b*SLOT32 and its size to nbytes[b].
template <int WT, int LAYOUT>
__global__ __launch_bounds__(NT) void k_compress_f32(const float *__restrict__ src, FieldGeom g, int nblocks,
                                                     float eps, uint32_t zkeep, int shuffle,
                                                     uint8_t *__restrict__ slots, uint32_t *__restrict__ nbytes)
{
    __shared__ float A[LDS_WORDS];
    __shared__ uint32_t M[1024];
    __shared__ uint32_t wsum[NT / 64];
    const int t = threadIdx.x;
    const int G = gridDim.x;
    float nx[NROW][32];
    int b = blockIdx.x;
    if (b < nblocks) {
#pragma unroll
        for (int k = 0; k < NROW; ++k) load_xrow_f32<LAYOUT>(src, g, b, t + NT * k, nx[k]);
    }
    for (; b < nblocks; b += G) {
#pragma unroll
        for (int k = 0; k < NROW; ++k) {
            M[t + NT * k] = 0u;
            float r[32];
#pragma unroll
            for (int i = 0; i < 32; ++i) r[i] = nx[k][i];
            row<32, true, WT>(r);  // level 32, x-pass
            lds_st<32>(A, xrow_lds_base<LAYOUT>(t + NT * k), 1, r);
        }
        __syncthreads();
        if (b + G < nblocks) {  // prefetch the next block while this one is transformed
#pragma unroll
            for (int k = 0; k < NROW; ++k) load_xrow_f32<LAYOUT>(src, g, b + G, t + NT * k, nx[k]);
        }
        fwt32_yz_f32<WT>(A, t);
        fwt16_8_f32<WT>(A, t);

        // threshold: keep bits of this thread's stream segments (base 4^3 always kept)
        float v[NROW][32];
        uint32_t keep[NROW];
#pragma unroll
        for (int k = 0; k < NROW; ++k) {
            const int sg = t + NT * k;
            keep[k] = 0;
            for_each_run(sg, [&](int q, int h, int addr, int stride, int word, int shift) {
                uint32_t bits = 0;
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    if (i < h) {
                        const float c = A[addr + i * stride];
                        v[k][q * h + i] = c;
                        const bool kp = (sg < 2) || (fabsf(c) > eps);
                        bits |= (uint32_t)kp << i;
                    }
                }
                keep[k] |= bits << (q * h);
                if (bits) atomicOr(&M[word], bits << shift);
            });
        }
        // segment counts: low 16 bits = segment t, high 16 bits = segment t+NT
        uint32_t tot;
        const uint32_t ex = block_excl_scan((uint32_t)__popc(keep[0]) | ((uint32_t)__popc(keep[1]) << 16), wsum, &tot);
        const uint32_t S = (tot & 0xffffu) + (tot >> 16);
        uint32_t first[NROW] = {ex & 0xffffu, (tot & 0xffffu) + (ex >> 16)};
        __syncthreads();  // every thread has read A: reuse it as the survivor stage

        uint32_t *stage = reinterpret_cast<uint32_t *>(A);
        uint8_t *stage8 = reinterpret_cast<uint8_t *>(A);
#pragma unroll
        for (int k = 0; k < NROW; ++k) {
            uint32_t idx = first[k];
#pragma unroll
            for (int j = 0; j < 32; ++j) {
                if ((keep[k] >> j) & 1u) {
                    const uint32_t w = __float_as_uint(v[k][j]) & zkeep;
                    if (shuffle) {
                        stage8[idx] = (uint8_t)w;
                        stage8[S + idx] = (uint8_t)(w >> 8);
                        stage8[2 * S + idx] = (uint8_t)(w >> 16);
                        stage8[3 * S + idx] = (uint8_t)(w >> 24);
                    } else {
                        stage[idx] = w;
                    }
                    ++idx;
                }
            }
        }
        __syncthreads();
        uint8_t *slot = slots + (size_t)b * SLOT32;
#pragma unroll
        for (int k = 0; k < NROW; ++k) reinterpret_cast<uint32_t *>(slot)[t + NT * k] = M[t + NT * k];
        uint32_t *dst = reinterpret_cast<uint32_t *>(slot + MASK_BYTES);
        for (uint32_t i = t; i < S; i += NT) dst[i] = stage[i];
        if (t == 0) nbytes[b] = MASK_BYTES + 4u * S;
        __syncthreads();  // stage and M are read before the next block reuses them
    }
}

// Exclusive scan of record sizes (4 + nbytes[b]) into offsets[0..n]; one
// workgroup of 1024 threads, each owning a contiguous run of ceil(n/1024)
// blocks so that every load is issued before the scan.
__global__ __launch_bounds__(1024) void k_scan_records(const uint32_t *__restrict__ nbytes, int n,
                                                       uint64_t *__restrict__ offsets)
{
    __shared__ uint64_t wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (n + 1023) / 1024;
    const int i0 = min(n, t * per), i1 = min(n, i0 + per);
    uint64_t v = 0;
    for (int i = i0; i < i1; ++i) v += (uint64_t)nbytes[i] + 4u;
    uint64_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint64_t run = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint64_t s = wsum[k];
        run += (k < w) ? s : 0u;
        tot += s;
    }
    run += inc - v;
    for (int i = i0; i < i1; ++i) {
        offsets[i] = run;
        run += (uint64_t)nbytes[i] + 4u;
    }
    if (t == 0) offsets[n] = tot;
}

// Copy each block's payload from its slot into the record stream, prefixed
// by its int32 size word (SerializerIO_WaveletCompression_MPI_Simple.h:237-240).
__global__ __launch_bounds__(256) void k_pack_records(const uint8_t *__restrict__ slots, size_t slot_bytes,
                                                      const uint32_t *__restrict__ nbytes,
                                                      const uint64_t *__restrict__ offsets,
                                                      uint8_t *__restrict__ records)
{
    const int b = blockIdx.x, t = threadIdx.x;
    const uint32_t n = nbytes[b];
    const uint32_t nd = n / 4;
    uint32_t *dst = reinterpret_cast<uint32_t *>(records + offsets[b]);
    const uint32_t *src = reinterpret_cast<const uint32_t *>(slots + (size_t)b * slot_bytes);
    if (t == 0) dst[0] = n;
    for (uint32_t base = t; base < nd; base += 256 * 4) {
        uint32_t v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t i = base + 256 * k;
            v[k] = i < nd ? src[i] : 0u;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t i = base + 256 * k;
            if (i < nd) dst[1 + i] = v[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Decompressor, binary32.  Mirrors WaveletCompressorGeneric::decompress
// (WaveletCompressor.cpp:232-259): mask, reshuffle, load (stream order,
// zeros for dropped coefficients; FullWaveletTransform.h:100-150), iwt.
// Persistent like the compressor: while block b is decoded, the record of
// block b+G (size word, mask, first PF*NT survivor words) is already in
// flight and offsets[b+2G] is being fetched.
constexpr int PF = 8;

struct RecPrefetch {
    uint32_t nb, mw[NROW], sv[PF];
};

CZW_DI void fetch_record(const uint8_t *__restrict__ records, uint64_t rec_bytes, uint64_t ob, int t,
                         RecPrefetch &r)
{
    if ((ob & 3u) || ob + 4 + MASK_BYTES > rec_bytes) {
        r.nb = 0xffffffffu;  // rejected by the size check
#pragma unroll
        for (int k = 0; k < NROW; ++k) r.mw[k] = 0u;
#pragma unroll
        for (int k = 0; k < PF; ++k) r.sv[k] = 0u;
        return;
    }
    const uint32_t *p = reinterpret_cast<const uint32_t *>(records + ob);
    const uint64_t nwords = (rec_bytes - ob) / 4;  // words of the buffer from this record on
    r.nb = p[0];
#pragma unroll
    for (int k = 0; k < NROW; ++k) r.mw[k] = p[1 + t + NT * k];
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const uint64_t i = 1 + 1024 + (uint64_t)k * NT + t;
        r.sv[k] = i < nwords ? p[i] : 0u;
    }
}

template <int WT, int LAYOUT>
__global__ __launch_bounds__(NT) void k_decompress_f32(const uint8_t *__restrict__ records, uint64_t rec_bytes,
                                                       const uint64_t *__restrict__ offsets, int nblocks,
                                                       int shuffle, float *__restrict__ dst, FieldGeom g,
                                                       uint32_t *__restrict__ err)
{
    __shared__ float A[LDS_WORDS];
    __shared__ uint32_t M[1024];
    __shared__ uint32_t wsum[NT / 64];
    __shared__ uint32_t coarse_ok;
    const int t = threadIdx.x;
    const int G = gridDim.x;
    int b = blockIdx.x;
    uint64_t ob = b < nblocks ? offsets[b] : 0;
    uint64_t ob_next = b + G < nblocks ? offsets[b + G] : 0;
    RecPrefetch pf;
    if (b < nblocks) fetch_record(records, rec_bytes, ob, t, pf);
    for (; b < nblocks; b += G) {
        const RecPrefetch cur = pf;
        const uint64_t ob_cur = ob;
        ob = ob_next;
        if (b + 2 * G < nblocks) ob_next = offsets[b + 2 * G];
        if (b + G < nblocks) fetch_record(records, rec_bytes, ob, t, pf);

        const uint32_t nb = cur.nb;
        if (nb < MASK_BYTES || ((nb - MASK_BYTES) & 3u) || nb > SLOT32 || ob_cur + 4 + nb > rec_bytes) {
            if (t == 0) atomicOr(err, (uint32_t)CZW_ERR_NBYTES);
            continue;
        }
        const uint32_t S = (nb - MASK_BYTES) >> 2;
        const uint32_t *pay = reinterpret_cast<const uint32_t *>(records + ob_cur + 4);
        __syncthreads();  // the previous block's last reads of A, M and coarse_ok are done
        if (t == 0) coarse_ok = 1u;
#pragma unroll
        for (int k = 0; k < NROW; ++k) M[t + NT * k] = cur.mw[k];
        uint32_t *stage = reinterpret_cast<uint32_t *>(A);
        const uint8_t *stage8 = reinterpret_cast<const uint8_t *>(A);
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const uint32_t i = (uint32_t)(k * NT + t);
            if (i < S) stage[i] = cur.sv[k];
        }
        for (uint32_t i = PF * NT + t; i < S; i += NT) stage[i] = pay[MASK_BYTES / 4 + i];
        __syncthreads();
        // the 4^3 base is always stored (FullWaveletTransform.h:163-176): its 64
        // mask bits are the low nibbles of words a1 + 32*a0, a0, a1 < 4
        if (t < 16 && (M[(t & 3) + 32 * (t >> 2)] & 0xfu) != 0xfu) coarse_ok = 0u;

        uint32_t keep[NROW];
#pragma unroll
        for (int k = 0; k < NROW; ++k) {
            const int sg = t + NT * k;
            keep[k] = 0;
            for_each_run(sg, [&](int q, int h, int, int, int word, int shift) {
                const uint32_t hm = (h == 16) ? 0xffffu : ((1u << h) - 1u);
                const uint32_t bits = (sg < 2) ? hm : ((M[word] >> shift) & hm);
                keep[k] |= bits << (q * h);
            });
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan((uint32_t)__popc(keep[0]) | ((uint32_t)__popc(keep[1]) << 16), wsum, &tot);
        // with all base bits set, the kept count equals the mask popcount
        // that WaveletCompressor.cpp:235-251 checks against the payload size
        if ((tot & 0xffffu) + (tot >> 16) != S || !coarse_ok) {
            if (t == 0) atomicOr(err, (uint32_t)CZW_ERR_POPCOUNT);
            continue;
        }
        const uint32_t first[NROW] = {ex & 0xffffu, (tot & 0xffffu) + (ex >> 16)};
        float v[NROW][32];
#pragma unroll
        for (int k = 0; k < NROW; ++k) {
            uint32_t idx = first[k];
#pragma unroll
            for (int j = 0; j < 32; ++j) {
                float c = 0.0f;
                if ((keep[k] >> j) & 1u) {
                    uint32_t w;
                    if (shuffle)
                        w = (uint32_t)stage8[idx] | ((uint32_t)stage8[S + idx] << 8) |
                            ((uint32_t)stage8[2 * S + idx] << 16) | ((uint32_t)stage8[3 * S + idx] << 24);
                    else
                        w = stage[idx];
                    c = __uint_as_float(w);
                    ++idx;
                }
                v[k][j] = c;
            }
        }
        __syncthreads();  // all survivors gathered: A becomes the coefficient array
#pragma unroll
        for (int k = 0; k < NROW; ++k) {
            for_each_run(t + NT * k, [&](int q, int h, int addr, int stride, int, int) {
#pragma unroll
                for (int i = 0; i < 16; ++i)
                    if (i < h) A[addr + i * stride] = v[k][q * h + i];
            });
        }
        __syncthreads();
        iwt_most_f32<WT>(A, t);
#pragma unroll
        for (int k = 0; k < NROW; ++k) {
            float r[32];
            lds_ld<32>(A, xrow_lds_base<LAYOUT>(t + NT * k), 1, r);
            row<32, false, WT>(r);  // level 32, x-pass
            store_xrow_f32<LAYOUT>(dst, g, b, t + NT * k, r);
        }
    }
}

// ---------------------------------------------------------------------------
// host-side dispatch

static int wt_index(int wtype)
{
    switch (wtype) {
    case 1: return 1;
    case 2: return 2;
    case 0:
    case 3: return 3;
    default: return -1;
    }
}

// Workgroups that fit on the current device at once for a kernel with the
// given block size (the persistent kernels use one 139 KB LDS block per CU).
static int resident_groups(const void *kernel, int threads)
{
    static int cus[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cus[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cus[dev] = n;
    }
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per <= 0) per = 1;
    return cus[dev] * per;
}

static bool zb_ok(int zb) { return zb == 0 || zb == 4 || zb == 8 || zb == 12 || zb == 16; }

template <int LAYOUT>
static int compress_slots_f32(const float *src, int nblocks, FieldGeom g, float eps, int wtype, int zerobits,
                              int shuffle, uint8_t *slots, uint32_t *nb, hipStream_t s)
{
    const int wt = wt_index(wtype);
    if (wt < 0 || !zb_ok(zerobits) || nblocks <= 0 || !src || !slots || !nb) return CZW_EINVAL;
    const uint32_t zkeep = zerobits ? ~((1u << zerobits) - 1u) : 0xffffffffu;
    const void *fn = wt == 1 ? (const void *)k_compress_f32<1, LAYOUT>
                   : wt == 2 ? (const void *)k_compress_f32<2, LAYOUT> : (const void *)k_compress_f32<3, LAYOUT>;
    const dim3 grid(std::min(nblocks, resident_groups(fn, NT))), blk(NT);
    if (wt == 1) hipLaunchKernelGGL((k_compress_f32<1, LAYOUT>), grid, blk, 0, s, src, g, nblocks, eps, zkeep, shuffle, slots, nb);
    if (wt == 2) hipLaunchKernelGGL((k_compress_f32<2, LAYOUT>), grid, blk, 0, s, src, g, nblocks, eps, zkeep, shuffle, slots, nb);
    if (wt == 3) hipLaunchKernelGGL((k_compress_f32<3, LAYOUT>), grid, blk, 0, s, src, g, nblocks, eps, zkeep, shuffle, slots, nb);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

static int pack_records(const uint8_t *slots, size_t slot_bytes, const uint32_t *nb, int nblocks, uint8_t *records,
                        uint64_t *offsets, hipStream_t s)
{
    if (nblocks <= 0 || !slots || !nb || !records || !offsets || (slot_bytes & 3)) return CZW_EINVAL;
    hipLaunchKernelGGL(k_scan_records, dim3(1), dim3(1024), 0, s, nb, nblocks, offsets);
    hipLaunchKernelGGL(k_pack_records, dim3(nblocks), dim3(256), 0, s, slots, slot_bytes, nb, offsets, records);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

template <int LAYOUT>
static int compress_f32(const float *src, int nblocks, FieldGeom g, float eps, int wtype, int zerobits, int shuffle,
                        void *records, uint64_t *offsets, void *workspace, hipStream_t s)
{
    if (!records || !offsets || !workspace || nblocks <= 0) return CZW_EINVAL;
    uint8_t *slots = static_cast<uint8_t *>(workspace);
    uint32_t *nb = reinterpret_cast<uint32_t *>(slots + SLOT32 * (size_t)nblocks);
    const int rc = compress_slots_f32<LAYOUT>(src, nblocks, g, eps, wtype, zerobits, shuffle, slots, nb, s);
    if (rc != CZW_OK) return rc;
    return pack_records(slots, SLOT32, nb, nblocks, static_cast<uint8_t *>(records), offsets, s);
}

template <int LAYOUT>
static int decompress_f32(const void *records, size_t rec_bytes, const uint64_t *offsets, int nblocks, FieldGeom g,
                          int wtype, int shuffle, float *dst, uint32_t *err, hipStream_t s)
{
    const int wt = wt_index(wtype);
    if (wt < 0 || nblocks <= 0 || !records || !offsets || !dst || !err) return CZW_EINVAL;
    const uint8_t *rec = static_cast<const uint8_t *>(records);
    const void *fn = wt == 1 ? (const void *)k_decompress_f32<1, LAYOUT>
                   : wt == 2 ? (const void *)k_decompress_f32<2, LAYOUT> : (const void *)k_decompress_f32<3, LAYOUT>;
    const dim3 grid(std::min(nblocks, resident_groups(fn, NT))), blk(NT);
    if (wt == 1) hipLaunchKernelGGL((k_decompress_f32<1, LAYOUT>), grid, blk, 0, s, rec, (uint64_t)rec_bytes, offsets, nblocks, shuffle, dst, g, err);
    if (wt == 2) hipLaunchKernelGGL((k_decompress_f32<2, LAYOUT>), grid, blk, 0, s, rec, (uint64_t)rec_bytes, offsets, nblocks, shuffle, dst, g, err);
    if (wt == 3) hipLaunchKernelGGL((k_decompress_f32<3, LAYOUT>), grid, blk, 0, s, rec, (uint64_t)rec_bytes, offsets, nblocks, shuffle, dst, g, err);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

static bool field_ok(int nx, int ny, int nz)
{
    return nx > 0 && ny > 0 && nz > 0 && !(nx & 31) && !(ny & 31) && !(nz & 31) &&
           (size_t)(nx / 32) * (ny / 32) * (nz / 32) < (size_t)INT32_MAX;
}

} // namespace czw

using namespace czw;

extern "C" {

int czw_version(void) { return 100; }

size_t czw_compress_workspace_bytes(int nblocks, int prec)
{
    if (nblocks <= 0 || (prec != 4 && prec != 8)) return 0;
    const size_t slot = MASK_BYTES + (size_t)prec * NB3;
    return slot * (size_t)nblocks + 4 * (size_t)nblocks;
}

size_t czw_records_bound(int nblocks, int prec)
{
    if (nblocks <= 0 || (prec != 4 && prec != 8)) return 0;
    return (4 + MASK_BYTES + (size_t)prec * NB3) * (size_t)nblocks;
}

int czw_compress_field(const void *field, int nx, int ny, int nz, int prec, float eps, int wtype, int zerobits,
                       int shuffle, void *records, uint64_t *offsets, void *workspace, hipStream_t stream)
{
    if (!field_ok(nx, ny, nz)) return CZW_EINVAL;
    const int nb = (nx / 32) * (ny / 32) * (nz / 32);
    const FieldGeom g{nx, ny, nz};
    if (prec == 4)
        return compress_f32<LAYOUT_FIELD>(static_cast<const float *>(field), nb, g, eps, wtype, zerobits, shuffle,
                                          records, offsets, workspace, stream);
    return prec == 8 ? CZW_EUNSUPPORTED : CZW_EINVAL;
}

int czw_compress_field_slots(const void *field, int nx, int ny, int nz, int prec, float eps, int wtype,
                             int zerobits, int shuffle, void *slots, uint32_t *nbytes, hipStream_t stream)
{
    if (!field_ok(nx, ny, nz)) return CZW_EINVAL;
    const int nb = (nx / 32) * (ny / 32) * (nz / 32);
    const FieldGeom g{nx, ny, nz};
    if (prec == 4)
        return compress_slots_f32<LAYOUT_FIELD>(static_cast<const float *>(field), nb, g, eps, wtype, zerobits,
                                                shuffle, static_cast<uint8_t *>(slots), nbytes, stream);
    return prec == 8 ? CZW_EUNSUPPORTED : CZW_EINVAL;
}

int czw_pack_records(const void *slots, size_t slot_bytes, const uint32_t *nbytes, int nblocks, void *records,
                     uint64_t *offsets, hipStream_t stream)
{
    return pack_records(static_cast<const uint8_t *>(slots), slot_bytes, nbytes, nblocks,
                        static_cast<uint8_t *>(records), offsets, stream);
}

size_t czw_slot_bytes(int prec) { return (prec == 4 || prec == 8) ? MASK_BYTES + (size_t)prec * NB3 : 0; }

int czw_compress_blocks(const void *blocks, int nblocks, int prec, float eps, int wtype, int zerobits, int shuffle,
                        void *records, uint64_t *offsets, void *workspace, hipStream_t stream)
{
    const FieldGeom g{0, 0, 0};
    if (prec == 4)
        return compress_f32<LAYOUT_BLOCKS>(static_cast<const float *>(blocks), nblocks, g, eps, wtype, zerobits,
                                           shuffle, records, offsets, workspace, stream);
    return prec == 8 ? CZW_EUNSUPPORTED : CZW_EINVAL;
}

int czw_decompress_field(const void *records, size_t records_bytes, const uint64_t *offsets, int nx, int ny, int nz,
                         int prec, int wtype, int shuffle, void *field, uint32_t *err, hipStream_t stream)
{
    if (!field_ok(nx, ny, nz)) return CZW_EINVAL;
    const int nb = (nx / 32) * (ny / 32) * (nz / 32);
    const FieldGeom g{nx, ny, nz};
    if (prec == 4)
        return decompress_f32<LAYOUT_FIELD>(records, records_bytes, offsets, nb, g, wtype, shuffle,
                                            static_cast<float *>(field), err, stream);
    return prec == 8 ? CZW_EUNSUPPORTED : CZW_EINVAL;
}

int czw_decompress_blocks(const void *records, size_t records_bytes, const uint64_t *offsets, int nblocks, int prec,
                          int wtype, int shuffle, void *blocks, uint32_t *err, hipStream_t stream)
{
    const FieldGeom g{0, 0, 0};
    if (prec == 4)
        return decompress_f32<LAYOUT_BLOCKS>(records, records_bytes, offsets, nblocks, g, wtype, shuffle,
                                             static_cast<float *>(blocks), err, stream);
    return prec == 8 ? CZW_EUNSUPPORTED : CZW_EINVAL;
}

} // extern "C"
