// czw_kernels.hip -- gfx950 kernels and C ABI (include/czw.h) of the wavelet
// block compressor.
//
// A persistent grid of 512-thread workgroups (8 wave64, one per CU slot the
// LDS allows) walks the blocks; a workgroup owns one 32^3 block at a time
// while the next block's input is prefetched into registers.  The block
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

#include "czw_common.hpp"

namespace czw {

constexpr int RP = 33;
constexpr int SP = 32 * 33 + 1;
constexpr int LDS_WORDS = 32 * SP;
constexpr size_t SLOT32 = MASK_BYTES + 4 * (size_t)NB3;



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
// One level-32 pass over this thread's NROW rows (row k starts at base[k],
// elements `stride` apart): every row is loaded before the first one is
// written back, so the loads of the second row overlap the arithmetic of the
// first (the compiler cannot hoist them past stores to the same array).
template <bool FWD, int WT>
CZW_DI void pass_rows_f32(float *A, const int (&base)[NROW], int stride)
{
    float r[NROW][32];
#pragma unroll
    for (int k = 0; k < NROW; ++k) lds_ld<32>(A, base[k], stride, r[k]);
#pragma unroll
    for (int k = 0; k < NROW; ++k) {
        row<32, FWD, WT>(r[k]);
        lds_st<32>(A, base[k], stride, r[k]);
    }
}

// Rows of the level-32 passes owned by thread t: y-pass rows (z, x) and
// z-pass rows (y, x) at q = t + NT*k, x fastest.
CZW_DI void ypass_bases(int t, int (&base)[NROW])
{
#pragma unroll
    for (int k = 0; k < NROW; ++k) base[k] = ((t + NT * k) >> 5) * SP + ((t + NT * k) & 31);
}

template <int WT>
CZW_DI void iwt_most_f32(float *A, int t)
{
    iwt16_8<WT, RP, SP>(A, t);
    {  // level 16 z-pass, then level 32 z-pass: rows (y, x)
        float r[NROW][32];
#pragma unroll
        for (int k = 0; k < NROW; ++k) lds_ld<32>(A, ((t + NT * k) >> 5) * RP + ((t + NT * k) & 31), SP, r[k]);
#pragma unroll
        for (int k = 0; k < NROW; ++k) {
            const int q = t + NT * k, y = q >> 5, x = q & 31;
            if (y < 16 && x < 16) row<16, false, WT>(r[k]);
            row<32, false, WT>(r[k]);
            lds_st<32>(A, y * RP + x, SP, r[k]);
        }
    }
    __syncthreads();
    {  // level 32 y-pass
        int base[NROW];
        ypass_bases(t, base);
        pass_rows_f32<false, WT>(A, base, RP);
    }
    __syncthreads();
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
    } else if constexpr (LAYOUT == LAYOUT_FIELD) {
        size_t xs;
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(src + field_row(g, b, 0, 0, &xs));
        const int vo = ((q >> 5) * g.nz + (q & 31)) * 4, so = (int)xs * 4;
#pragma unroll
        for (int x = 0; x < 32; ++x) r[x] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, vo, x * so, 0));
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
// LDS address of bit j (0..31) of stream segment g (see for_each_run).
CZW_DI int seg_lds_addr(int g, int j)
{
    if (g < 16) {
        const int s = 32 * g + j, code = s >> 6, k = s & 63;
        const int a0 = 4 * ((code >> 2) & 1) + (k >> 4), a1 = 4 * ((code >> 1) & 1) + ((k >> 2) & 3);
        return a0 * SP + a1 * RP + 4 * (code & 1) + (k & 3);
    } else if (g < 128) {
        const int s = 32 * g + j, code = s >> 9, k = s & 511;
        const int a0 = 8 * ((code >> 2) & 1) + (k >> 6), a1 = 8 * ((code >> 1) & 1) + ((k >> 3) & 7);
        return a1 * SP + (8 * (code & 1) + (k & 7)) * RP + a0;
    } else {
        const int s = 32 * g + j, code = s >> 12, k = s & 4095;
        const int a0 = 16 * ((code >> 2) & 1) + (k >> 8), a1 = 16 * ((code >> 1) & 1) + ((k >> 4) & 15);
        return (16 * (code & 1) + (k & 15)) * SP + a0 * RP + a1;
    }
}

// Slot layout written by the compressor (block b at slots + b*SLOT32):
//   [mask 4096 B][survivors of the [0,16)^3 corner, stream order, <= 4096
//   words][level-32 survivors, stream order], both before the byte shuffle.
// nbytes[b] = payload bytes 4096 + 4*S, nbytes[nblocks + b] = corner count.
constexpr int CORNER_MAX = 4096;
constexpr int NRUN32 = 1792;  // level-32 stream runs: 7 codes x 16 x 16 runs of 16
constexpr int STAGE32_AT = 16 * SP;                      // the z >= 16 half of A
constexpr int STAGE32 = ((LDS_WORDS - STAGE32_AT - 64) / 4) * 4;  // staged survivors (64 dummy words follow)


//
// Block order: workgroup w starts with block w; the next block is w + G
// (static) or, when `queue` is given, the next ticket G + atomicAdd(queue, 1)
// of a work queue whose counter the caller has cleared.  In LAYOUT_BATCH the
// blocks of nf fields (table `fields`, batch-wide numbering) form one queue.
template <int WT, int LAYOUT>
__global__ __launch_bounds__(NT) void k_compress_f32(const float *__restrict__ src, FieldGeom g, int nblocks,
                                                     float eps, uint32_t zkeep, uint8_t *__restrict__ slots,
                                                     uint32_t *__restrict__ nbytes,
                                                     const czw_field_t *__restrict__ fields, int nf,
                                                     uint32_t *__restrict__ queue)
{
    constexpr int RL = row_layout<LAYOUT>();
    __shared__ __align__(16) float A[LDS_WORDS];
    __shared__ uint32_t M[1024];
    __shared__ uint32_t cnt[2048];  // level-32 run counts, then their exclusive offsets
    __shared__ uint32_t wsum[NT / 64];
    __shared__ uint32_t s32_total;
    __shared__ int ticket[2];
    __shared__ czw_field_t T[LAYOUT == LAYOUT_BATCH ? BATCH_MAX_FIELDS : 1];
    const int t = threadIdx.x;
    const int G = gridDim.x;
    load_table<LAYOUT>(T, fields, nf);
    float nx[NROW][32];
    int b = blockIdx.x;
    float eps_b = eps;
    if (b < nblocks) {
        const BlockRef<float> br = locate<LAYOUT, float>(T, nf, b, src, nullptr, g, eps);
        eps_b = br.eps;
#pragma unroll
        for (int k = 0; k < NROW; ++k) load_xrow_f32<RL>(br.src, br.g, br.lb, t + NT * k, nx[k]);
    }
    for (int it = 0; b < nblocks; ++it) {
        uint8_t *slot = slots + (size_t)b * SLOT32;
        if (queue && t == 0) ticket[it & 1] = G + (int)atomicAdd(queue, 1u);
#pragma unroll
        for (int k = 0; k < NROW; ++k) {  // level 32, x-pass
            float r[32];
#pragma unroll
            for (int i = 0; i < 32; ++i) r[i] = nx[k][i];
            row<32, true, WT>(r);
            lds_st<32>(A, xrow_lds_base<RL>(t + NT * k), 1, r);
        }
        if (t < 2048 - NRUN32) cnt[NRUN32 + t] = 0u;
        __syncthreads();
        // Prefetch the next block's rows in two halves, each well under the
        // 63 outstanding loads a wave may have (more would stall the wave
        // until the first returns).  The last block reloads itself, so nx is
        // always written and is dead between its use and its reload.
        const int bnext = queue ? ticket[it & 1] : b + G;
        const int bn = min(bnext, nblocks - 1);
        const float eps = eps_b;
        {
            const BlockRef<float> br = locate<LAYOUT, float>(T, nf, bn, src, nullptr, g, eps);
            eps_b = br.eps;
            load_xrow_f32<RL>(br.src, br.g, br.lb, t, nx[0]);
        }
        {  // level 32, y-pass: rows (z, x)
            int base[NROW];
            ypass_bases(t, base);
            pass_rows_f32<true, WT>(A, base, RP);
        }
        __syncthreads();
        // level 32, z-pass: row (y, x) = physical mask word x + 32*y, bit z
        // (level-32 layout (a0,a1,a2) = (y,x,z)).  Outside the 16^2 corner of
        // (y,x) all 32 results are final; inside it only z >= 16 is, and the
        // level-16 z-pass continues on z < 16.  Keep bits are taken here, in
        // registers; the values stay in registers until they are written.
        float r[NROW][32];
        uint32_t kb[NROW];
#pragma unroll
        for (int k = 0; k < NROW; ++k) lds_ld<32>(A, ((t + NT * k) >> 5) * RP + ((t + NT * k) & 31), SP, r[k]);
#pragma unroll
        for (int k = 0; k < NROW; ++k) {
            const int q = t + NT * k, y = q >> 5, x = q & 31;
            const bool corner = (y < 16) && (x < 16);
            row<32, true, WT>(r[k]);
            uint32_t bits = 0;
#pragma unroll
            for (int z = 0; z < 32; ++z) bits |= (uint32_t)(fabsf(r[k][z]) > eps) << z;
            if (corner) {
                float c[16];
#pragma unroll
                for (int z = 0; z < 16; ++z) c[z] = r[k][z];
                row<16, true, WT>(c);
                lds_st<16>(A, y * RP + x, SP, c);
                bits &= 0xffff0000u;
            }
            kb[k] = bits;
            M[x + 32 * y] = bits;
            // stream run (code c, i0 = y&15, i1 = x&15) of each z-half; code
            // bits: z >= 16, x >= 16, y >= 16 (FullWaveletTransform.h:68-72)
            const int c0 = 2 * (x >> 4) + 4 * (y >> 4);
            const int base = (y & 15) * 16 + (x & 15);
            if (!corner) cnt[(c0 - 1) * 256 + base] = __popc(bits & 0xffffu);
            cnt[c0 * 256 + base] = __popc(bits >> 16);
        }
        __syncthreads();
        {  // exclusive scan of the run counts in stream order
            uint32_t c4[4], sum = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                c4[i] = cnt[4 * t + i];
                sum += c4[i];
            }
            uint32_t S32;
            uint32_t run = block_excl_scan(sum, wsum, &S32);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                cnt[4 * t + i] = run;
                run += c4[i];
            }
            if (t == 0) s32_total = S32;
        }
        __syncthreads();
        const uint32_t S32 = s32_total;
        // level-32 survivors, from registers, zerobits applied.  The z >= 16
        // half of A is dead after the z-pass (the remaining passes only touch
        // the [0,16)^3 corner), so survivors are compacted there branch-free
        // (unkept values go to a per-lane dummy word) and then copied out with
        // 16-byte stores.  A block with more survivors than that half holds
        // writes them straight to the slot instead.
        uint32_t *d32 = reinterpret_cast<uint32_t *>(slot + MASK_BYTES) + CORNER_MAX;
        const bool staged = S32 <= (uint32_t)STAGE32;
        uint32_t *st32 = staged ? reinterpret_cast<uint32_t *>(A) + STAGE32_AT : d32;
        const uint32_t dummy = (uint32_t)STAGE32 + (t & 63);
#pragma unroll
        for (int k = 0; k < NROW; ++k) {
            const int q = t + NT * k, y = q >> 5, x = q & 31;
            const int c0 = 2 * (x >> 4) + 4 * (y >> 4);
            const int base = (y & 15) * 16 + (x & 15);
            uint32_t i0 = (c0 > 0) ? cnt[(c0 - 1) * 256 + base] : 0u;
            uint32_t i1 = cnt[c0 * 256 + base];
            if (staged) {
#pragma unroll
                for (int z = 0; z < 32; ++z) {
                    const uint32_t bit = (kb[k] >> z) & 1u;
                    uint32_t &i = z < 16 ? i0 : i1;
                    st32[bit ? i : dummy] = __float_as_uint(r[k][z]) & zkeep;
                    i += bit;
                }
            } else {
#pragma unroll
                for (int z = 0; z < 32; ++z) {
                    uint32_t &i = z < 16 ? i0 : i1;
                    if ((kb[k] >> z) & 1u) d32[i++] = __float_as_uint(r[k][z]) & zkeep;
                }
            }
        }
        {  // second half of the prefetch
            const BlockRef<float> br = locate<LAYOUT, float>(T, nf, bn, src, nullptr, g, eps);
            load_xrow_f32<RL>(br.src, br.g, br.lb, t + NT, nx[1]);
        }
        __syncthreads();  // stage complete, wsum free again
        if (staged) {
            const uint4 *s4 = reinterpret_cast<const uint4 *>(st32);
            uint4 *d4 = reinterpret_cast<uint4 *>(d32);
            for (uint32_t i = t; i < S32 / 4; i += NT) d4[i] = s4[i];
            for (uint32_t i = (S32 & ~3u) + t; i < S32; i += NT) d32[i] = st32[i];
        }
        fwt16_8<WT, RP, SP>(A, t);

        // the [0,16)^3 corner: stream positions [0,4096) in pieces of 8 per
        // thread (base 4^3 always kept); values re-read from A
        uint32_t keep = 0;
        float v[8];
        corner_piece<RP, SP>(t, [&](int q, int h, int addr, int stride, int word, int shift) {
            uint32_t bits = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (i < h) {
                    const float c = A[addr + i * stride];
                    v[q * h + i] = c;
                    bits |= (uint32_t)((t < 8) || (fabsf(c) > eps)) << i;
                }
            }
            keep |= bits << (q * h);
            atomicOr(&M[word], bits << shift);
        });
        uint32_t Sc;
        uint32_t idx = block_excl_scan(__popc(keep), wsum, &Sc);
        {
            uint32_t *dc = reinterpret_cast<uint32_t *>(slot + MASK_BYTES);
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if ((keep >> j) & 1u) dc[idx++] = __float_as_uint(v[j]) & zkeep;
        }
        __syncthreads();  // mask complete
#pragma unroll
        for (int k = 0; k < NROW; ++k) reinterpret_cast<uint32_t *>(slot)[t + NT * k] = M[t + NT * k];
        if (t == 0) {
            nbytes[b] = MASK_BYTES + 4u * (Sc + S32);
            nbytes[nblocks + b] = Sc;
        }
        __syncthreads();  // A, M, cnt and wsum are read before the next block reuses them
        b = bnext;
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

int scan_records(const uint32_t *nbytes, int n, uint64_t *offsets, hipStream_t s)
{
    hipLaunchKernelGGL(k_scan_records, dim3(1), dim3(1024), 0, s, nbytes, n, offsets);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

// Frame each block's payload into the record stream
// (SerializerIO_WaveletCompression_MPI_Simple.h:237-240): size word, mask,
// corner survivors then level-32 survivors, with the byte-plane shuffle
// (WaveletCompressor.cpp:56-73: byte p of survivor i at p*S + i) if asked.
__global__ __launch_bounds__(256) void k_pack_records(const uint8_t *__restrict__ slots, size_t slot_bytes,
                                                      const uint32_t *__restrict__ nbytes, int nblocks, int shuffle,
                                                      const uint64_t *__restrict__ offsets,
                                                      uint8_t *__restrict__ records)
{
    const int b = blockIdx.x, t = threadIdx.x;
    const uint32_t n = nbytes[b], sc = nbytes[nblocks + b];
    const uint32_t S = (n - MASK_BYTES) / 4;
    uint32_t *dst = reinterpret_cast<uint32_t *>(records + offsets[b]);
    uint8_t *dst8 = records + offsets[b] + 4 + MASK_BYTES;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(slots + (size_t)b * slot_bytes);
    const uint32_t *corner = src + MASK_BYTES / 4, *l32 = corner + CORNER_MAX;
    if (t == 0) dst[0] = n;
    for (uint32_t i = t; i < MASK_BYTES / 4; i += 256) dst[1 + i] = src[i];
    for (uint32_t base = t; base < S; base += 256 * 4) {
        uint32_t v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t i = base + 256 * k;
            v[k] = i < S ? (i < sc ? corner[i] : l32[i - sc]) : 0u;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t i = base + 256 * k;
            if (i < S) {
                if (shuffle) {
                    dst8[i] = (uint8_t)v[k];
                    dst8[S + i] = (uint8_t)(v[k] >> 8);
                    dst8[2 * S + i] = (uint8_t)(v[k] >> 16);
                    dst8[3 * S + i] = (uint8_t)(v[k] >> 24);
                } else {
                    dst[1 + MASK_BYTES / 4 + i] = v[k];
                }
            }
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

// Survivor words sv[k] = survivor k*NT + t were loaded for the indices below
// `lim`; the record's extent is taken from the next record's offset when the
// records are in order (a hint only: survivors past lim are read later), so
// the prefetch does not read into the following records.
struct RecPrefetch {
    uint32_t nb, lim, mw[NROW], sv[PF];
};

CZW_DI void fetch_record(const uint8_t *__restrict__ records, uint64_t rec_bytes, uint64_t ob, uint64_t ob_end,
                         int t, RecPrefetch &r)
{
    if ((ob & 3u) || ob + 4 + MASK_BYTES > rec_bytes) {
        r.nb = 0xffffffffu;  // rejected by the size check
        r.lim = 0u;
#pragma unroll
        for (int k = 0; k < NROW; ++k) r.mw[k] = 0u;
#pragma unroll
        for (int k = 0; k < PF; ++k) r.sv[k] = 0u;
        return;
    }
    const uint32_t *p = reinterpret_cast<const uint32_t *>(records + ob);
    uint64_t nwords = (rec_bytes - ob) / 4;  // words of the buffer from this record on
    if (ob_end > ob + 4 + MASK_BYTES && (ob_end - ob) / 4 < nwords) nwords = (ob_end - ob) / 4;
    r.lim = nwords > 1 + 1024 ? (uint32_t)min(nwords - 1 - 1024, (uint64_t)(PF * NT)) : 0u;
    r.nb = p[0];
#pragma unroll
    for (int k = 0; k < NROW; ++k) r.mw[k] = p[1 + t + NT * k];
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const uint32_t i = (uint32_t)(k * NT + t);
        r.sv[k] = i < r.lim ? p[1 + 1024 + i] : 0u;
    }
}

// Undo the byte-plane shuffle of S staged survivors in place (byte p of
// survivor i at byte p*S + i, WaveletCompressor.cpp:75-93): each thread
// gathers its survivors i = t + NT*j into registers, then all write back.
template <int MAXPT>
CZW_DI void unshuffle_stage(uint32_t *stage, uint32_t S)
{
    const int t = threadIdx.x;
    const uint8_t *b8 = reinterpret_cast<const uint8_t *>(stage);
    uint32_t u[MAXPT];
#pragma unroll
    for (int j = 0; j < MAXPT; ++j) {
        const uint32_t i = (uint32_t)(t + NT * j);
        u[j] = i < S ? ((uint32_t)b8[i] | ((uint32_t)b8[S + i] << 8) | ((uint32_t)b8[2 * S + i] << 16) |
                        ((uint32_t)b8[3 * S + i] << 24))
                     : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < MAXPT; ++j) {
        const uint32_t i = (uint32_t)(t + NT * j);
        if (i < S) stage[i] = u[j];
    }
    __syncthreads();
}

// Decompression of one block whose survivors fit the z >= 16 half of A
// (S <= STAGE32): survivors are staged there; the corner positions are
// written into A branch-free (value or +0); the level-32 survivors are
// gathered straight into registers for the fused level-16/level-32 z-pass.
// Returns false (uniformly) when the record is inconsistent.
template <int WT>
CZW_DI bool decode_block_fast_f32(float *A, uint32_t *M, uint32_t *cnt, uint32_t *wsum, uint32_t *flag,
                                  const RecPrefetch &cur, const uint32_t *pay, uint32_t S, int shuffle,
                                  float (&r)[NROW][32])
{
    const int t = threadIdx.x;
    uint32_t *stage = reinterpret_cast<uint32_t *>(A) + STAGE32_AT;
    const uint8_t *stage8 = reinterpret_cast<const uint8_t *>(stage);
    if (t == 0) *flag = 1u;
#pragma unroll
    for (int k = 0; k < NROW; ++k) M[t + NT * k] = cur.mw[k];
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const uint32_t i = (uint32_t)(k * NT + t);
        if (i < S) stage[i] = i < cur.lim ? cur.sv[k] : pay[MASK_BYTES / 4 + i];
    }
    for (uint32_t i = PF * NT + t; i < S; i += NT) stage[i] = pay[MASK_BYTES / 4 + i];
    if (t < 64) stage[STAGE32 + t] = 0u;  // the dummy words read for dropped coefficients
    __syncthreads();
    if (shuffle) unshuffle_stage<(STAGE32 + NT - 1) / NT>(stage, S);
    // the 4^3 base is always stored (FullWaveletTransform.h:163-176): its 64
    // mask bits are the low nibbles of words a1 + 32*a0, a0, a1 < 4
    if (t < 16 && (M[(t & 3) + 32 * (t >> 2)] & 0xfu) != 0xfu) *flag = 0u;
    // counts in stream order: corner piece t at cnt[t], level-32 run rho at cnt[512 + rho]
    uint32_t keep = 0;
    corner_piece<RP, SP>(t, [&](int q, int h, int, int, int word, int shift) {
        const uint32_t hm = (1u << h) - 1u;
        keep |= ((t < 8) ? hm : ((M[word] >> shift) & hm)) << (q * h);
    });
    cnt[t] = __popc(keep);
    uint32_t mw[NROW];
#pragma unroll
    for (int k = 0; k < NROW; ++k) {
        const int q = t + NT * k, y = q >> 5, x = q & 31;
        const bool corner = (y < 16) && (x < 16);
        mw[k] = M[x + 32 * y] & (corner ? 0xffff0000u : 0xffffffffu);
        const int c0 = 2 * (x >> 4) + 4 * (y >> 4), base = (y & 15) * 16 + (x & 15);
        if (!corner) cnt[512 + (c0 - 1) * 256 + base] = __popc(mw[k] & 0xffffu);
        cnt[512 + c0 * 256 + base] = __popc(mw[k] >> 16);
    }
    if (t < 2560 - 512 - NRUN32) cnt[512 + NRUN32 + t] = 0u;
    __syncthreads();
    {  // exclusive scan of the 2560 counts, 5 consecutive per thread
        uint32_t c5[5], sum = 0;
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            c5[i] = cnt[5 * t + i];
            sum += c5[i];
        }
        uint32_t tot;
        uint32_t run = block_excl_scan(sum, wsum, &tot);
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            cnt[5 * t + i] = run;
            run += c5[i];
        }
        if (tot != S) *flag = 0u;
    }
    __syncthreads();
    if (!*flag) return false;
    // the 64 dummy words after the stage hold +0: a dropped coefficient reads
    // its lane's dummy word
    auto sv = [&](uint32_t i) -> float { return __uint_as_float(stage[i]); };
    const uint32_t dummy = (uint32_t)STAGE32 + (t & 63);
    {  // corner positions: survivor or +0
        uint32_t idx = cnt[t];
        corner_piece<RP, SP>(t, [&](int q, int h, int addr, int stride, int, int) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (i < h) {
                    const uint32_t bit = (keep >> (q * h + i)) & 1u;
                    const float v = sv(bit ? idx : dummy);
                    A[addr + i * stride] = v;
                    idx += bit;
                }
            }
        });
    }
    // level-32 survivors of this thread's z-columns, straight into registers
#pragma unroll
    for (int k = 0; k < NROW; ++k) {
        const int q = t + NT * k, y = q >> 5, x = q & 31;
        const int c0 = 2 * (x >> 4) + 4 * (y >> 4), base = (y & 15) * 16 + (x & 15);
        uint32_t i0 = (c0 > 0) ? cnt[512 + (c0 - 1) * 256 + base] : 0u;
        uint32_t i1 = cnt[512 + c0 * 256 + base];
#pragma unroll
        for (int z = 0; z < 32; ++z) {
            const uint32_t bit = (mw[k] >> z) & 1u;
            uint32_t &i = z < 16 ? i0 : i1;
            const float v = sv(bit ? i : dummy);
            r[k][z] = v;
            i += bit;
        }
    }
    __syncthreads();  // stage consumed; corner of A complete
    return true;
}

//
// Block order as in the compressor, three deep: the record of the next block
// b1 and the offset of the one after, b2, are in flight while b is decoded.
// Workgroup w starts with w, w + G, w + 2G; then b2 advances by G (static) or
// takes the ticket 3G + atomicAdd(queue, 1) of a cleared work queue.
template <int WT, int LAYOUT>
__global__ __launch_bounds__(NT) void k_decompress_f32(const uint8_t *__restrict__ records, uint64_t rec_bytes,
                                                       const uint64_t *__restrict__ offsets, int nblocks,
                                                       int shuffle, float *__restrict__ dst, FieldGeom g,
                                                       uint32_t *__restrict__ err,
                                                       const czw_field_t *__restrict__ fields, int nf,
                                                       uint32_t *__restrict__ queue)
{
    constexpr int RL = row_layout<LAYOUT>();
    __shared__ __align__(16) float A[LDS_WORDS];
    __shared__ uint32_t M[1024];
    __shared__ uint32_t cnt[2560];
    __shared__ uint32_t wsum[NT / 64];
    __shared__ uint32_t flag;
    __shared__ int ticket[2];
    __shared__ czw_field_t T[LAYOUT == LAYOUT_BATCH ? BATCH_MAX_FIELDS : 1];
    const int t = threadIdx.x;
    const int G = gridDim.x;
    load_table<LAYOUT>(T, fields, nf);
    // offset of a block's record and the next block's (an extent hint; 0 = none)
    auto off_at = [&](int i) -> uint64_t { return i < nblocks ? offsets[i] : 0; };
    int b = blockIdx.x, b1 = b + G, b2 = b + 2 * G, bq = 0;
    uint64_t ob0 = off_at(b);                          // block b
    uint64_t ob1 = off_at(b1), ob1_end = off_at(b1 + 1);  // block b1
    RecPrefetch pf;
    if (b < nblocks) fetch_record(records, rec_bytes, ob0, off_at(b + 1), t, pf);
    for (int it = 0; b < nblocks; ++it, b = b1, b1 = b2, b2 = bq) {
        if (queue) {
            if (t == 0) ticket[it & 1] = 3 * G + (int)atomicAdd(queue, 1u);
            __syncthreads();
            bq = ticket[it & 1];
        } else {
            bq = b2 + G;
        }
        const RecPrefetch cur = pf;
        const uint64_t ob_cur = ob0;
        if (b1 < nblocks) fetch_record(records, rec_bytes, ob1, ob1_end, t, pf);
        ob0 = ob1;
        if (b2 < nblocks) {
            ob1 = offsets[b2];
            ob1_end = off_at(b2 + 1);
        }

        const uint32_t nb = cur.nb;
        if (nb < MASK_BYTES || ((nb - MASK_BYTES) & 3u) || nb > SLOT32 || ob_cur + 4 + nb > rec_bytes) {
            if (t == 0) atomicOr(err, (uint32_t)CZW_ERR_NBYTES);
            continue;
        }
        const uint32_t S = (nb - MASK_BYTES) >> 2;
        const uint32_t *pay = reinterpret_cast<const uint32_t *>(records + ob_cur + 4);
        __syncthreads();  // the previous block's last reads of A, M, cnt and flag are done
        float r[NROW][32];
        if (S <= (uint32_t)STAGE32) {
            if (!decode_block_fast_f32<WT>(A, M, cnt, wsum, &flag, cur, pay, S, shuffle, r)) {
                if (t == 0) atomicOr(err, (uint32_t)CZW_ERR_POPCOUNT);
                continue;
            }
            iwt16_8<WT, RP, SP>(A, t);
            // level 16 z-pass (corner columns), then level 32 z-pass; both
            // threads' corner loads are issued before the first store
#pragma unroll
            for (int k = 0; k < NROW; ++k) {
                const int q = t + NT * k, y = q >> 5, x = q & 31;
                if (y < 16 && x < 16) lds_ld<16>(A, y * RP + x, SP, r[k]);
            }
#pragma unroll
            for (int k = 0; k < NROW; ++k) {
                const int q = t + NT * k, y = q >> 5, x = q & 31;
                if (y < 16 && x < 16) row<16, false, WT>(r[k]);
                row<32, false, WT>(r[k]);
                lds_st<32>(A, y * RP + x, SP, r[k]);
            }
            __syncthreads();
        } else {
            // dense block: survivors staged over A, every position gathered
            // into registers, then placed; the full inverse follows
            if (t == 0) flag = 1u;
#pragma unroll
            for (int k = 0; k < NROW; ++k) M[t + NT * k] = cur.mw[k];
            uint32_t *stage = reinterpret_cast<uint32_t *>(A);
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const uint32_t i = (uint32_t)(k * NT + t);  // < S: dense blocks hold more than PF*NT survivors
                stage[i] = i < cur.lim ? cur.sv[k] : pay[MASK_BYTES / 4 + i];
            }
            for (uint32_t i = PF * NT + t; i < S; i += NT) stage[i] = pay[MASK_BYTES / 4 + i];
            __syncthreads();
            if (t < 16 && (M[(t & 3) + 32 * (t >> 2)] & 0xfu) != 0xfu) flag = 0u;
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
            const uint32_t ex =
                block_excl_scan((uint32_t)__popc(keep[0]) | ((uint32_t)__popc(keep[1]) << 16), wsum, &tot);
            // with all base bits set, the kept count equals the mask popcount
            // that WaveletCompressor.cpp:235-251 checks against the payload size
            if ((tot & 0xffffu) + (tot >> 16) != S || !flag) {
                if (t == 0) atomicOr(err, (uint32_t)CZW_ERR_POPCOUNT);
                continue;
            }
            const uint32_t first[NROW] = {ex & 0xffffu, (tot & 0xffffu) + (ex >> 16)};
#pragma unroll
            for (int k = 0; k < NROW; ++k) {
                uint32_t idx = first[k];
#pragma unroll
                for (int j = 0; j < 32; ++j) {
                    float c = 0.0f;
                    if ((keep[k] >> j) & 1u) {
                        const uint8_t *b8 = reinterpret_cast<const uint8_t *>(stage);
                        c = __uint_as_float(shuffle ? ((uint32_t)b8[idx] | ((uint32_t)b8[S + idx] << 8) |
                                                       ((uint32_t)b8[2 * S + idx] << 16) |
                                                       ((uint32_t)b8[3 * S + idx] << 24))
                                                    : stage[idx]);
                        ++idx;
                    }
                    r[k][j] = c;
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < NROW; ++k) {
                for_each_run(t + NT * k, [&](int q, int h, int addr, int stride, int, int) {
#pragma unroll
                    for (int i = 0; i < 16; ++i)
                        if (i < h) A[addr + i * stride] = r[k][q * h + i];
                });
            }
            __syncthreads();
            iwt_most_f32<WT>(A, t);
        }
        if (S <= (uint32_t)STAGE32) {  // level 32 y-pass
            int base[NROW];
            ypass_bases(t, base);
            pass_rows_f32<false, WT>(A, base, RP);
            __syncthreads();
        }
        if constexpr (RL == LAYOUT_BLOCKS) {
#pragma unroll
            for (int k = 0; k < NROW; ++k) {
                float rr[32];
                lds_ld<32>(A, xrow_lds_base<RL>(t + NT * k), 1, rr);
                row<32, false, WT>(rr);  // level 32, x-pass
                store_xrow_f32<RL>(dst, g, b, t + NT * k, rr);
            }
        } else {
            // level 32 x-pass back into A, then 16-byte stores along the
            // field's contiguous Z: four times fewer store instructions than
            // storing each x-row element on its own
            {
                int base[NROW];
#pragma unroll
                for (int k = 0; k < NROW; ++k) base[k] = (t & 31) * SP + ((t + NT * k) >> 5) * RP;
                pass_rows_f32<false, WT>(A, base, 1);
            }
            __syncthreads();
            const BlockRef<float> br = locate<LAYOUT, float>(T, nf, b, nullptr, dst, g, 0.0f);
            size_t xs;
            float *p0 = br.dst + field_row(br.g, br.lb, 0, 0, &xs);
            const int nz = br.g.nz;
            if constexpr (RL == LAYOUT_FIELD) {
                // x = i >> 8 is the same for the 64 lanes of a wave: a scalar offset
                const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(p0);
                const int so = (int)xs * 4;
#pragma unroll 4
                for (int i = t; i < 32 * 32 * 8; i += NT) {  // (x, y, z4) with z4 fastest
                    const int z4 = i & 7, y = (i >> 3) & 31, x = __builtin_amdgcn_readfirstlane(i >> 8);
                    const int a = (4 * z4) * SP + y * RP + x;
                    const czw_u32x4 v = {__float_as_uint(A[a]), __float_as_uint(A[a + SP]), __float_as_uint(A[a + 2 * SP]),
                                         __float_as_uint(A[a + 3 * SP])};
                    __builtin_amdgcn_raw_buffer_store_b128(v, rs, (y * nz + 4 * z4) * 4, x * so, 0);
                }
            } else {
#pragma unroll 4
                for (int i = t; i < 32 * 32 * 8; i += NT) {  // (x, y, z4) with z4 fastest
                    const int z4 = i & 7, y = (i >> 3) & 31, x = i >> 8;
                    const int a = (4 * z4) * SP + y * RP + x;
                    const float4 v = make_float4(A[a], A[a + SP], A[a + 2 * SP], A[a + 3 * SP]);
                    *reinterpret_cast<float4 *>(p0 + x * xs + (size_t)y * nz + 4 * z4) = v;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host-side dispatch


template <int LAYOUT>
static int compress_slots_f32(const float *src, int nblocks, FieldGeom g, float eps, int wtype, int zerobits,
                              uint8_t *slots, uint32_t *nb, hipStream_t s)
{
    const int wt = wt_index(wtype);
    if (wt < 0 || !zb_ok(zerobits) || nblocks <= 0 || !src || !slots || !nb) return CZW_EINVAL;
    const uint32_t zkeep = zerobits ? ~((1u << zerobits) - 1u) : 0xffffffffu;
    const void *fn = wt == 1 ? (const void *)k_compress_f32<1, LAYOUT>
                   : wt == 2 ? (const void *)k_compress_f32<2, LAYOUT> : (const void *)k_compress_f32<3, LAYOUT>;
    const dim3 grid(std::min(nblocks, resident_groups(fn, NT))), blk(NT);
    const czw_field_t *nf0 = nullptr;
    uint32_t *nq = nullptr;
    if (wt == 1) hipLaunchKernelGGL((k_compress_f32<1, LAYOUT>), grid, blk, 0, s, src, g, nblocks, eps, zkeep, slots, nb, nf0, 0, nq);
    if (wt == 2) hipLaunchKernelGGL((k_compress_f32<2, LAYOUT>), grid, blk, 0, s, src, g, nblocks, eps, zkeep, slots, nb, nf0, 0, nq);
    if (wt == 3) hipLaunchKernelGGL((k_compress_f32<3, LAYOUT>), grid, blk, 0, s, src, g, nblocks, eps, zkeep, slots, nb, nf0, 0, nq);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

static int pack_records(const uint8_t *slots, size_t slot_bytes, const uint32_t *nb, int nblocks, int shuffle,
                        uint8_t *records, uint64_t *offsets, hipStream_t s)
{
    if (nblocks <= 0 || !slots || !nb || !records || !offsets || (slot_bytes & 3)) return CZW_EINVAL;
    if (scan_records(nb, nblocks, offsets, s) != CZW_OK) return CZW_EHIP;
    hipLaunchKernelGGL(k_pack_records, dim3(nblocks), dim3(256), 0, s, slots, slot_bytes, nb, nblocks, shuffle, offsets,
                       records);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

template <int LAYOUT>
static int compress_f32(const float *src, int nblocks, FieldGeom g, float eps, int wtype, int zerobits, int shuffle,
                        void *records, uint64_t *offsets, void *workspace, hipStream_t s)
{
    if (!records || !offsets || !workspace || nblocks <= 0) return CZW_EINVAL;
    uint8_t *slots = static_cast<uint8_t *>(workspace);
    uint32_t *nb = reinterpret_cast<uint32_t *>(slots + SLOT32 * (size_t)nblocks);
    const int rc = compress_slots_f32<LAYOUT>(src, nblocks, g, eps, wtype, zerobits, slots, nb, s);
    if (rc != CZW_OK) return rc;
    return pack_records(slots, SLOT32, nb, nblocks, shuffle, static_cast<uint8_t *>(records), offsets, s);
}

static int compress_f64(int layout, const void *src, int nblocks, FieldGeom g, float eps, int wtype, int zerobits,
                        int shuffle, void *records, uint64_t *offsets, void *workspace, hipStream_t s)
{
    if (!records || !offsets || !workspace || nblocks <= 0) return CZW_EINVAL;
    uint8_t *slots = static_cast<uint8_t *>(workspace);
    uint32_t *nb = reinterpret_cast<uint32_t *>(slots + (MASK_BYTES + 8 * (size_t)NB3) * (size_t)nblocks);
    const int rc = compress_slots_f64(layout, src, nblocks, g, eps, wtype, zerobits, slots, nb, nullptr, 0, 0,
                                      nullptr, s);
    if (rc != CZW_OK) return rc;
    return pack_records_f64(slots, nb, nblocks, shuffle, static_cast<uint8_t *>(records), offsets, s);
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
    const czw_field_t *nf0 = nullptr;
    uint32_t *nq = nullptr;
    if (wt == 1) hipLaunchKernelGGL((k_decompress_f32<1, LAYOUT>), grid, blk, 0, s, rec, (uint64_t)rec_bytes, offsets, nblocks, shuffle, dst, g, err, nf0, 0, nq);
    if (wt == 2) hipLaunchKernelGGL((k_decompress_f32<2, LAYOUT>), grid, blk, 0, s, rec, (uint64_t)rec_bytes, offsets, nblocks, shuffle, dst, g, err, nf0, 0, nq);
    if (wt == 3) hipLaunchKernelGGL((k_decompress_f32<3, LAYOUT>), grid, blk, 0, s, rec, (uint64_t)rec_bytes, offsets, nblocks, shuffle, dst, g, err, nf0, 0, nq);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}


// ---------------------------------------------------------------------------
// batched mode

static_assert(sizeof(czw_field_t) == 40, "czw_field_t layout");
static_assert(BATCH_MAX_FIELDS == CZW_BATCH_MAX_FIELDS, "field table size");

// Clears n words; used instead of a memset so that captured graphs replay it.
__global__ void k_fill_u32(uint32_t *__restrict__ p, int n, uint32_t v)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

static uint32_t *batch_queue(const czw_batch_t *B)
{
    const size_t slot = MASK_BYTES + (size_t)B->prec * NB3;
    return reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(B->workspace) +
                                        (slot + 4 * size_words(B->prec)) * (size_t)B->nblocks);
}

static bool batch_ok(const czw_batch_t *B)
{
    return B && B->fields && B->nfields > 0 && B->nfields <= BATCH_MAX_FIELDS && B->nblocks > 0 && B->workspace &&
           B->records && B->offsets && B->groups >= 0 && wt_index(B->wtype) > 0;
}

static int batch_compress(const czw_batch_t *B, hipStream_t s)
{
    if (!batch_ok(B) || !zb_ok(B->zerobits)) return CZW_EINVAL;
    if (B->prec != 4 && B->prec != 8) return CZW_EINVAL;
    const int wt = wt_index(B->wtype);
    uint8_t *slots = static_cast<uint8_t *>(B->workspace);
    const size_t slot_bytes = MASK_BYTES + (size_t)B->prec * NB3;
    uint32_t *nb = reinterpret_cast<uint32_t *>(slots + slot_bytes * (size_t)B->nblocks);
    uint32_t *q = B->queue ? batch_queue(B) : nullptr;
    if (q) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, q, 1, 0u);
    if (B->prec == 8) {
        const int rc = compress_slots_f64(LAYOUT_BATCH, nullptr, B->nblocks, FieldGeom{0, 0, 0}, 0.0f, B->wtype,
                                          B->zerobits, slots, nb, B->fields, B->nfields, B->groups, q, s);
        if (rc != CZW_OK) return rc;
        return pack_records_f64(slots, nb, B->nblocks, B->shuffle, static_cast<uint8_t *>(B->records), B->offsets, s);
    }
    const uint32_t zkeep = B->zerobits ? ~((1u << B->zerobits) - 1u) : 0xffffffffu;
    const void *fn = wt == 1 ? (const void *)k_compress_f32<1, LAYOUT_BATCH>
                   : wt == 2 ? (const void *)k_compress_f32<2, LAYOUT_BATCH> : (const void *)k_compress_f32<3, LAYOUT_BATCH>;
    const int groups = B->groups ? B->groups : resident_groups(fn, NT);
    const dim3 grid(std::min(B->nblocks, groups)), blk(NT);
    const float *src0 = nullptr;
    const FieldGeom g{0, 0, 0};
    if (wt == 1) hipLaunchKernelGGL((k_compress_f32<1, LAYOUT_BATCH>), grid, blk, 0, s, src0, g, B->nblocks, 0.0f, zkeep, slots, nb, B->fields, B->nfields, q);
    if (wt == 2) hipLaunchKernelGGL((k_compress_f32<2, LAYOUT_BATCH>), grid, blk, 0, s, src0, g, B->nblocks, 0.0f, zkeep, slots, nb, B->fields, B->nfields, q);
    if (wt == 3) hipLaunchKernelGGL((k_compress_f32<3, LAYOUT_BATCH>), grid, blk, 0, s, src0, g, B->nblocks, 0.0f, zkeep, slots, nb, B->fields, B->nfields, q);
    if (hipGetLastError() != hipSuccess) return CZW_EHIP;
    return pack_records(slots, SLOT32, nb, B->nblocks, B->shuffle, static_cast<uint8_t *>(B->records), B->offsets, s);
}

static int batch_decompress(const czw_batch_t *B, hipStream_t s)
{
    if (!batch_ok(B) || !B->err) return CZW_EINVAL;
    if (B->prec != 4 && B->prec != 8) return CZW_EINVAL;
    const int wt = wt_index(B->wtype);
    uint32_t *q = B->queue ? batch_queue(B) + 1 : nullptr;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, B->err, 1, 0u);
    if (q) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, q, 1, 0u);
    if (B->prec == 8)
        return decompress_f64(LAYOUT_BATCH, B->records, B->records_bytes, B->offsets, B->nblocks, FieldGeom{0, 0, 0},
                              B->wtype, B->shuffle, nullptr, B->err, B->fields, B->nfields, B->groups, q, s);
    const uint8_t *rec = static_cast<const uint8_t *>(B->records);
    const void *fn = wt == 1 ? (const void *)k_decompress_f32<1, LAYOUT_BATCH>
                   : wt == 2 ? (const void *)k_decompress_f32<2, LAYOUT_BATCH> : (const void *)k_decompress_f32<3, LAYOUT_BATCH>;
    const int groups = B->groups ? B->groups : resident_groups(fn, NT);
    const dim3 grid(std::min(B->nblocks, groups)), blk(NT);
    float *dst0 = nullptr;
    const FieldGeom g{0, 0, 0};
    const uint64_t rb = (uint64_t)B->records_bytes;
    if (wt == 1) hipLaunchKernelGGL((k_decompress_f32<1, LAYOUT_BATCH>), grid, blk, 0, s, rec, rb, B->offsets, B->nblocks, B->shuffle, dst0, g, B->err, B->fields, B->nfields, q);
    if (wt == 2) hipLaunchKernelGGL((k_decompress_f32<2, LAYOUT_BATCH>), grid, blk, 0, s, rec, rb, B->offsets, B->nblocks, B->shuffle, dst0, g, B->err, B->fields, B->nfields, q);
    if (wt == 3) hipLaunchKernelGGL((k_decompress_f32<3, LAYOUT_BATCH>), grid, blk, 0, s, rec, rb, B->offsets, B->nblocks, B->shuffle, dst0, g, B->err, B->fields, B->nfields, q);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

} // namespace czw

using namespace czw;

extern "C" {

int czw_version(void) { return 100; }

size_t czw_compress_workspace_bytes(int nblocks, int prec)
{
    if (nblocks <= 0 || (prec != 4 && prec != 8)) return 0;
    const size_t slot = MASK_BYTES + (size_t)prec * NB3;
    return (slot + 4 * size_words(prec)) * (size_t)nblocks;
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
    const bool buf = field_fits_buffer(ny, nz, prec);
    if (prec == 4)
        return buf ? compress_f32<LAYOUT_FIELD>(static_cast<const float *>(field), nb, g, eps, wtype, zerobits,
                                                shuffle, records, offsets, workspace, stream)
                   : compress_f32<LAYOUT_FIELD_WIDE>(static_cast<const float *>(field), nb, g, eps, wtype, zerobits,
                                                     shuffle, records, offsets, workspace, stream);
    if (prec == 8)
        return compress_f64(buf ? LAYOUT_FIELD : LAYOUT_FIELD_WIDE, field, nb, g, eps, wtype, zerobits, shuffle,
                            records, offsets, workspace, stream);
    return CZW_EINVAL;
}

int czw_compress_field_range(const void *field, int nx, int ny, int nz, int prec, float eps, int wtype,
                             int zerobits, int shuffle, int block0, int nblocks, void *records, uint64_t *offsets,
                             void *workspace, hipStream_t stream)
{
    if (!field_ok(nx, ny, nz)) return CZW_EINVAL;
    const int total = (nx / 32) * (ny / 32) * (nz / 32);
    if (block0 < 0 || nblocks <= 0 || block0 > total - nblocks) return CZW_EINVAL;
    const FieldGeom g{nx, ny, nz, block0};
    const bool buf = field_fits_buffer(ny, nz, prec);
    if (prec == 4)
        return buf ? compress_f32<LAYOUT_FIELD>(static_cast<const float *>(field), nblocks, g, eps, wtype, zerobits,
                                                shuffle, records, offsets, workspace, stream)
                   : compress_f32<LAYOUT_FIELD_WIDE>(static_cast<const float *>(field), nblocks, g, eps, wtype,
                                                     zerobits, shuffle, records, offsets, workspace, stream);
    if (prec == 8)
        return compress_f64(buf ? LAYOUT_FIELD : LAYOUT_FIELD_WIDE, field, nblocks, g, eps, wtype, zerobits, shuffle,
                            records, offsets, workspace, stream);
    return CZW_EINVAL;
}

int czw_compress_field_slots(const void *field, int nx, int ny, int nz, int prec, float eps, int wtype,
                             int zerobits, void *slots, uint32_t *nbytes, hipStream_t stream)
{
    if (!field_ok(nx, ny, nz)) return CZW_EINVAL;
    const int nb = (nx / 32) * (ny / 32) * (nz / 32);
    const FieldGeom g{nx, ny, nz};
    const bool buf = field_fits_buffer(ny, nz, prec);
    if (prec == 4)
        return buf ? compress_slots_f32<LAYOUT_FIELD>(static_cast<const float *>(field), nb, g, eps, wtype, zerobits,
                                                      static_cast<uint8_t *>(slots), nbytes, stream)
                   : compress_slots_f32<LAYOUT_FIELD_WIDE>(static_cast<const float *>(field), nb, g, eps, wtype,
                                                           zerobits, static_cast<uint8_t *>(slots), nbytes, stream);
    if (prec == 8)
        return compress_slots_f64(buf ? LAYOUT_FIELD : LAYOUT_FIELD_WIDE, field, nb, g, eps, wtype, zerobits,
                                  static_cast<uint8_t *>(slots), nbytes, nullptr, 0, 0, nullptr, stream);
    return CZW_EINVAL;
}

int czw_pack_records(const void *slots, size_t slot_bytes, const uint32_t *nbytes, int nblocks, int shuffle,
                     void *records, uint64_t *offsets, hipStream_t stream)
{
    if (slot_bytes == MASK_BYTES + 8 * (size_t)NB3)
        return pack_records_f64(static_cast<const uint8_t *>(slots), nbytes, nblocks, shuffle,
                                static_cast<uint8_t *>(records), offsets, stream);
    if (slot_bytes != SLOT32) return CZW_EINVAL;
    return pack_records(static_cast<const uint8_t *>(slots), slot_bytes, nbytes, nblocks, shuffle,
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
    if (prec == 8) return compress_f64(LAYOUT_BLOCKS, blocks, nblocks, g, eps, wtype, zerobits, shuffle, records,
                                       offsets, workspace, stream);
    return CZW_EINVAL;
}

int czw_decompress_field(const void *records, size_t records_bytes, const uint64_t *offsets, int nx, int ny, int nz,
                         int prec, int wtype, int shuffle, void *field, uint32_t *err, hipStream_t stream)
{
    if (!field_ok(nx, ny, nz)) return CZW_EINVAL;
    const int nb = (nx / 32) * (ny / 32) * (nz / 32);
    const FieldGeom g{nx, ny, nz};
    const bool buf = field_fits_buffer(ny, nz, prec);
    if (prec == 4)
        return buf ? decompress_f32<LAYOUT_FIELD>(records, records_bytes, offsets, nb, g, wtype, shuffle,
                                                  static_cast<float *>(field), err, stream)
                   : decompress_f32<LAYOUT_FIELD_WIDE>(records, records_bytes, offsets, nb, g, wtype, shuffle,
                                                       static_cast<float *>(field), err, stream);
    if (prec == 8)
        return decompress_f64(buf ? LAYOUT_FIELD : LAYOUT_FIELD_WIDE, records, records_bytes, offsets, nb, g, wtype,
                              shuffle, field, err, nullptr, 0, 0, nullptr, stream);
    return CZW_EINVAL;
}

int czw_decompress_blocks(const void *records, size_t records_bytes, const uint64_t *offsets, int nblocks, int prec,
                          int wtype, int shuffle, void *blocks, uint32_t *err, hipStream_t stream)
{
    const FieldGeom g{0, 0, 0};
    if (prec == 4)
        return decompress_f32<LAYOUT_BLOCKS>(records, records_bytes, offsets, nblocks, g, wtype, shuffle,
                                             static_cast<float *>(blocks), err, stream);
    if (prec == 8)
        return decompress_f64(LAYOUT_BLOCKS, records, records_bytes, offsets, nblocks, g, wtype, shuffle, blocks,
                              err, nullptr, 0, 0, nullptr, stream);
    return CZW_EINVAL;
}

int czw_batch_prepare(czw_field_t *fields, int nfields, int prec, int *nblocks)
{
    if (!fields || !nblocks || nfields <= 0 || nfields > BATCH_MAX_FIELDS || (prec != 4 && prec != 8))
        return CZW_EINVAL;
    int64_t total = 0;
    for (int f = 0; f < nfields; ++f) {
        czw_field_t &d = fields[f];
        if (!field_ok(d.nx, d.ny, d.nz) || !field_fits_buffer(d.ny, d.nz, prec) || !(d.eps >= 0.0f))
            return CZW_EINVAL;
        d.block0 = (int32_t)total;
        d.nblocks = (d.nx / 32) * (d.ny / 32) * (d.nz / 32);
        total += d.nblocks;
        if (total >= INT32_MAX / 4) return CZW_EINVAL;
    }
    *nblocks = (int)total;
    return CZW_OK;
}

size_t czw_batch_workspace_bytes(int nblocks, int prec)
{
    const size_t w = czw_compress_workspace_bytes(nblocks, prec);
    return w ? w + 256 : 0;
}

int czw_batch_compress(const czw_batch_t *batch, hipStream_t stream) { return batch_compress(batch, stream); }

int czw_batch_decompress(const czw_batch_t *batch, hipStream_t stream) { return batch_decompress(batch, stream); }

int czw_batch_autotune(czw_batch_t *batch, int stage, hipStream_t stream, float *best_ms)
{
    if (!batch_ok(batch) || (stage != 1 && stage != 2)) return CZW_EINVAL;
    const void *fn = (const void *)k_compress_f32<1, LAYOUT_BATCH>;
    const int R = resident_groups(fn, NT);
    const int shapes[4][2] = {{R, 0}, {R, 1}, {R / 2 > 0 ? R / 2 : 1, 0}, {R / 2 > 0 ? R / 2 : 1, 1}};
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return CZW_EHIP;
    czw_batch_t trial = *batch;
    float best = 1e30f;
    int bg = R, bq = 0, rc = CZW_OK;
    for (int k = 0; k < 4 && rc == CZW_OK; ++k) {
        trial.groups = shapes[k][0];
        trial.queue = shapes[k][1];
        auto run = [&]() { return stage == 1 ? batch_compress(&trial, stream) : batch_decompress(&trial, stream); };
        if ((rc = run()) != CZW_OK) break;  // warm-up
        (void)hipEventRecord(e0, stream);
        for (int r = 0; r < 3 && rc == CZW_OK; ++r) rc = run();
        (void)hipEventRecord(e1, stream);
        if (hipEventSynchronize(e1) != hipSuccess) rc = CZW_EHIP;
        float ms = 0.0f;
        if (rc == CZW_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms / 3 < best) {
            best = ms / 3;
            bg = trial.groups;
            bq = trial.queue;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != CZW_OK) return rc;
    batch->groups = bg;
    batch->queue = bq;
    if (best_ms) *best_ms = best;
    return CZW_OK;
}

int czw_batch_graph(const czw_batch_t *batch, int stages, void **graph)
{
    if (!batch_ok(batch) || !graph || stages < 1 || stages > 3) return CZW_EINVAL;
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return CZW_EHIP;
    int rc = CZW_OK;
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed) != hipSuccess) rc = CZW_EHIP;
    if (rc == CZW_OK && (stages & 1)) rc = batch_compress(batch, cs);
    if (rc == CZW_OK && (stages & 2)) rc = batch_decompress(batch, cs);
    // end the capture whatever happened, so the stream is usable again
    if (hipStreamEndCapture(cs, &g) != hipSuccess && rc == CZW_OK) rc = CZW_EHIP;
    if (rc == CZW_OK && hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) rc = CZW_EHIP;
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
    if (rc == CZW_OK) *graph = ex;
    return rc;
}

int czw_graph_launch(void *graph, hipStream_t stream)
{
    if (!graph) return CZW_EINVAL;
    return hipGraphLaunch(static_cast<hipGraphExec_t>(graph), stream) == hipSuccess ? CZW_OK : CZW_EHIP;
}

void czw_graph_destroy(void *graph)
{
    if (graph) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph));
}

} // extern "C"
