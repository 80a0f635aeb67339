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
#include <type_traits>

#include "czw_common.hpp"

namespace czw {

constexpr int RP = 33;
constexpr int SP = 32 * 33 + 1;
constexpr int LDS_WORDS = 32 * SP;
constexpr size_t SLOT32 = MASK_BYTES + 4 * (size_t)NB3;
// the compressor's workgroup: CT threads, CR level-32 rows per thread
constexpr int CT = 512;
constexpr int CR = 1024 / CT;



// ---------------------------------------------------------------------------
// One level-32 pass over this thread's CR rows (row k starts at base[k],
// elements `stride` apart): every row is loaded before the first one is
// written back, so the loads of the second row overlap the arithmetic of the
// first (the compiler cannot hoist them past stores to the same array).
template <bool FWD, int WT>
CZW_DI void pass_rows_f32(float *A, const int (&base)[CR], int stride)
{
    float r[CR][32];
#pragma unroll
    for (int k = 0; k < CR; ++k) lds_ld<32>(A, base[k], stride, r[k]);
#pragma unroll
    for (int k = 0; k < CR; ++k) {
        row<32, FWD, WT>(r[k]);
        lds_st<32>(A, base[k], stride, r[k]);
    }
}

// Rows of the level-32 passes owned by thread t: y-pass rows (z, x) and
// z-pass rows (y, x) at q = t + CT*k, x fastest.
CZW_DI void ypass_bases(int t, int (&base)[CR])
{
#pragma unroll
    for (int k = 0; k < CR; ++k) base[k] = ((t + CT * k) >> 5) * SP + ((t + CT * k) & 31);
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

// The CR x-rows of thread t.  Block layout: q = t + CT*k.  Field layouts:
// the rows (y, z0) and (y, z0 + 1) with y = t >> 4, z0 = 2*(t & 15), so that
// one 8-byte load reads both rows' values at one x (half the load
// instructions of 4-byte loads; their issue, not their bytes, is what a
// block's loads cost).
template <int LAYOUT>
CZW_DI int xrow_of(int t, int k)
{
    if constexpr (LAYOUT == LAYOUT_BLOCKS) return t + CT * k;
    else return ((t >> 4) << 5) | (2 * (t & 15) + k);
}

// x in [X0, X1) of thread t's CR x-rows into r[k][X0 .. X1).
template <int LAYOUT, int X0, int X1>
CZW_DI void load_xrows_f32(const float *__restrict__ src, const FieldGeom &g, int b, int t, float (&r)[CR][32])
{
    static_assert(CR == 2, "x-rows come in pairs");
    if constexpr (LAYOUT == LAYOUT_BLOCKS) {
#pragma unroll
        for (int k = 0; k < CR; ++k) {
            const czw_f32x4 *p = reinterpret_cast<const czw_f32x4 *>(src + (size_t)b * NB3 + (size_t)(t + CT * k) * 32);
#pragma unroll
            for (int i = X0 / 4; i < X1 / 4; ++i) {
                const czw_f32x4 v = __builtin_nontemporal_load(p + i);
                r[k][4 * i] = v.x;
                r[k][4 * i + 1] = v.y;
                r[k][4 * i + 2] = v.z;
                r[k][4 * i + 3] = v.w;
            }
        }
    } else if constexpr (LAYOUT == LAYOUT_FIELD) {
        size_t xs;
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(src + field_row(g, b, 0, 0, &xs));
        const int vo = ((t >> 4) * g.nz + 2 * (t & 15)) * 4, so = (int)xs * 4;
#pragma unroll
        for (int x = X0; x < X1; ++x) {
            const czw_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, vo, x * so, CZW_FIELD_CPOL);
            r[0][x] = __uint_as_float(v.x);
            r[1][x] = __uint_as_float(v.y);
        }
    } else {
        size_t xs;
        const float *p = src + field_row(g, b, t >> 4, 2 * (t & 15), &xs);
#pragma unroll
        for (int x = X0; x < X1; ++x) {
            const czw_f32x2 v = __builtin_nontemporal_load(reinterpret_cast<const czw_f32x2 *>(p + x * xs));
            r[0][x] = v.x;
            r[1][x] = v.y;
        }
    }
}

// ---------------------------------------------------------------------------
// Compressor, binary32.  Mirrors WaveletCompressorGeneric::compress
// (WaveletCompressor.cpp:153-182): fwt, threshold + mask, zerobits, shuffle.
// Persistent: workgroup w handles blocks w, w+G, w+2G, ... (G = gridDim.x)
// and loads block b+G into registers while block b is transformed, so the
// HBM reads of one block overlap the arithmetic of the previous one.  Block
// b's payload goes to slots + b*SLOT32 and its size to nbytes[b].

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
__global__ __launch_bounds__(CT) void k_compress_f32(const float *__restrict__ src, FieldGeom g, int nblocks,
                                                     float eps, uint32_t zkeep, uint8_t *__restrict__ slots,
                                                     uint32_t *__restrict__ nbytes,
                                                     const czw_field_t *__restrict__ fields, int nf,
                                                     uint32_t *__restrict__ queue)
{
    constexpr int RL = row_layout<LAYOUT>();
    __shared__ __align__(16) float A[LDS_WORDS];
    __shared__ uint32_t M[1024];
    __shared__ uint32_t cnt[2048];  // level-32 run counts, then their exclusive offsets
    __shared__ uint32_t wsum[CT / 64];
    __shared__ uint32_t s32_total;
    __shared__ int ticket[2];
    __shared__ czw_field_t T[LAYOUT == LAYOUT_BATCH ? BATCH_MAX_FIELDS : 1];
    const int t = threadIdx.x;
    const int G = gridDim.x;
    load_table<LAYOUT>(T, fields, nf);
    float nx[CR][32];
    int b = blockIdx.x;
    float eps_b = eps;
    if (b < nblocks) {
        const BlockRef<float> br = locate<LAYOUT, float>(T, nf, b, src, nullptr, g, eps);
        eps_b = br.eps;
        load_xrows_f32<RL, 0, 32>(br.src, br.g, br.lb, t, nx);
    }
    for (int it = 0; b < nblocks; ++it) {
        uint8_t *slot = slots + (size_t)b * SLOT32;
        if (queue && t == 0) ticket[it & 1] = G + (int)atomicAdd(queue, 1u);
#pragma unroll
        for (int k = 0; k < CR; ++k) {  // level 32, x-pass
            float r[32];
#pragma unroll
            for (int i = 0; i < 32; ++i) r[i] = nx[k][i];
            row<32, true, WT>(r);
            lds_st<32>(A, xrow_lds_base<RL>(xrow_of<RL>(t, k)), 1, r);
        }
        if (t < 2048 - NRUN32) cnt[NRUN32 + t] = 0u;
        __syncthreads();
        // Prefetch the next block's rows in four quarters of x, issued in
        // four phases: a CU keeps only some 48 KB of loads in flight, and a
        // wave that issues past that waits, so 32 KB at a time keeps the
        // waves computing while the loads run.  The last block reloads
        // itself, so nx is always written and is dead between its use and
        // its reload.
        const int bnext = queue ? ticket[it & 1] : b + G;
        const int bn = min(bnext, nblocks - 1);
        const float eps = eps_b;
        auto prefetch = [&](auto x0c) {
            constexpr int X0 = decltype(x0c)::value;
            const BlockRef<float> br = locate<LAYOUT, float>(T, nf, bn, src, nullptr, g, eps);
            if constexpr (X0 == 0) eps_b = br.eps;
            load_xrows_f32<RL, X0, X0 + 8>(br.src, br.g, br.lb, t, nx);
        };
        prefetch(std::integral_constant<int, 0>());
        {  // level 32, y-pass: rows (z, x)
            int base[CR];
            ypass_bases(t, base);
            pass_rows_f32<true, WT>(A, base, RP);
        }
        __syncthreads();
        prefetch(std::integral_constant<int, 8>());
        // level 32, z-pass: row (y, x) = physical mask word x + 32*y, bit z
        // (level-32 layout (a0,a1,a2) = (y,x,z)).  Outside the 16^2 corner of
        // (y,x) all 32 results are final; inside it only z >= 16 is, and the
        // level-16 z-pass continues on z < 16.  Keep bits are taken here, in
        // registers; the values stay in registers until they are written.
        float r[CR][32];
        uint32_t kb[CR];
#pragma unroll
        for (int k = 0; k < CR; ++k) lds_ld<32>(A, ((t + CT * k) >> 5) * RP + ((t + CT * k) & 31), SP, r[k]);
#pragma unroll
        for (int k = 0; k < CR; ++k) {
            const int q = t + CT * k, y = q >> 5, x = q & 31;
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
            constexpr int PER = 2048 / CT;
            uint32_t c4[PER], sum = 0;
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                c4[i] = cnt[PER * t + i];
                sum += c4[i];
            }
            uint32_t S32;
            uint32_t run = block_excl_scan<CT>(sum, wsum, &S32);
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                cnt[PER * t + i] = run;
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
        for (int k = 0; k < CR; ++k) {
            const int q = t + CT * k, y = q >> 5, x = q & 31;
            const int c0 = 2 * (x >> 4) + 4 * (y >> 4);
            const int base = (y & 15) * 16 + (x & 15);
            uint32_t i0 = (c0 > 0) ? cnt[(c0 - 1) * 256 + base] : 0u;
            uint32_t i1 = cnt[c0 * 256 + base];
            if (staged) {
                // Byte addresses into A, each run walked from its last
                // survivor down: with m = 0 or -1 the keep bit, the target is
                // (m & a) | (~m & dummy) (v_bfi_b32) and the next a is a + 4*m
                // (v_lshl_add_u32), three VALU per position with the bit
                // extract.  Both are written as instructions, as the compiler
                // otherwise rebuilds the addresses from a running index.
                // Zerobits are applied by the copy-out.
                char *Ab = reinterpret_cast<char *>(A);
                const uint32_t n0 = __popc(kb[k] & 0xffffu), n1 = __popc(kb[k] >> 16);
                uint32_t a0 = (STAGE32_AT + i0 + n0 - 1) * 4u, a1 = (STAGE32_AT + i1 + n1 - 1) * 4u;
                const uint32_t ad = (STAGE32_AT + dummy) * 4u;
#pragma unroll
                for (int z = 31; z >= 0; --z) {
                    const uint32_t m = (uint32_t)((int32_t)(kb[k] << (31 - z)) >> 31);
                    uint32_t &a = z < 16 ? a0 : a1;
                    uint32_t at;
                    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(at) : "v"(m), "v"(a), "v"(ad));
                    *reinterpret_cast<float *>(Ab + at) = r[k][z];
                    asm("v_lshl_add_u32 %0, %1, 2, %2" : "=v"(a) : "v"(m), "v"(a));
                }
            } else {
#pragma unroll
                for (int z = 0; z < 32; ++z) {
                    uint32_t &i = z < 16 ? i0 : i1;
                    if ((kb[k] >> z) & 1u) d32[i++] = __float_as_uint(r[k][z]) & zkeep;
                }
            }
        }
        prefetch(std::integral_constant<int, 16>());
        __syncthreads();  // stage complete, wsum free again
        // Levels 16 and 8 on the [0,16)^3 corner (fwt16_8's passes) keep 256,
        // then 64 threads busy.  Meanwhile threads 256.. copy the staged
        // survivors out (the staging area lies outside the corner), and
        // every thread issues its last prefetch quarter in a pass where it
        // has no rows.
        {
            float c[16];
            if (t < 256) {  // level 16 x-pass: row (z, y) in [0,16)^2
                const int z = t >> 4, y = t & 15;
                lds_ld<16>(A, z * SP + y * RP, 1, c);
                row<16, true, WT>(c);
                lds_st<16>(A, z * SP + y * RP, 1, c);
            } else {
                if (staged) {
                    const uint4 *s4 = reinterpret_cast<const uint4 *>(st32);
                    uint4 *d4 = reinterpret_cast<uint4 *>(d32);
                    for (uint32_t i = t - 256; i < S32 / 4; i += CT - 256) {
                        uint4 v = s4[i];
                        v.x &= zkeep;
                        v.y &= zkeep;
                        v.z &= zkeep;
                        v.w &= zkeep;
                        d4[i] = v;
                    }
                    for (uint32_t i = (S32 & ~3u) + t - 256; i < S32; i += CT - 256) d32[i] = st32[i] & zkeep;
                }
                prefetch(std::integral_constant<int, 24>());
            }
            __syncthreads();
            if (t < 256) {  // level 16 y-pass, then level 8 y-pass: row (z, x)
                const int z = t >> 4, x = t & 15;
                lds_ld<16>(A, z * SP + x, RP, c);
                row<16, true, WT>(c);
                if (z < 8 && x < 8) row<8, true, WT>(c);
                lds_st<16>(A, z * SP + x, RP, c);
            }
            __syncthreads();
            if (t < 64) {  // level 8 z-pass: row (y, x) in [0,8)^2
                const int y = t >> 3, x = t & 7;
                lds_ld<8>(A, y * RP + x, SP, c);
                row<8, true, WT>(c);
                lds_st<8>(A, y * RP + x, SP, c);
            } else if (t < 256) {
                prefetch(std::integral_constant<int, 24>());
            }
            __syncthreads();
            if (t < 64) {  // level 8 x-pass: row (z, y)
                const int z = t >> 3, y = t & 7;
                lds_ld<8>(A, z * SP + y * RP, 1, c);
                row<8, true, WT>(c);
                lds_st<8>(A, z * SP + y * RP, 1, c);
            }
            __syncthreads();
            if (t < 64) prefetch(std::integral_constant<int, 24>());
        }

        // the [0,16)^3 corner: stream positions [0,4096) in pieces of 8 per
        // thread (base 4^3 always kept); values re-read from A
        uint32_t keep = 0;
        float v[8];
        if (t < 512) corner_piece<RP, SP>(t, [&](int q, int h, int addr, int stride, int word, int shift) {
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
        uint32_t idx = block_excl_scan<CT>(__popc(keep), wsum, &Sc);
        {
            uint32_t *dc = reinterpret_cast<uint32_t *>(slot + MASK_BYTES);
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if ((keep >> j) & 1u) dc[idx++] = __float_as_uint(v[j]) & zkeep;
        }
        __syncthreads();  // mask complete
#pragma unroll
        for (int k = 0; k < CR; ++k) reinterpret_cast<uint32_t *>(slot)[t + CT * k] = M[t + CT * k];
        if (t == 0) {
            nbytes[b] = MASK_BYTES + 4u * (Sc + S32);
            nbytes[nblocks + b] = Sc;
        }
        __syncthreads();  // A, M, cnt and wsum are read before the next block reuses them
        b = bnext;
    }
}

// Exclusive scan of record sizes (4 + nbytes[b]) into offsets[0..n]; one
// workgroup of 1024 threads over chunks of 8192 sizes staged in LDS, so that
// the global loads and the 8-byte offset stores are both coalesced.  A
// chunk's sizes sum to at most 8192 records of <= 266244 bytes: u32.
__global__ __launch_bounds__(1024) void k_scan_records(const uint32_t *__restrict__ nbytes, int n,
                                                       uint64_t *__restrict__ offsets)
{
    constexpr int CH = 8192;
    __shared__ uint32_t P[CH];
    __shared__ uint32_t wsum[16];
    const int t = threadIdx.x;
    uint64_t carry = 0;
    for (int c0 = 0; c0 < n; c0 += CH) {
        // unconditional loads at an index clamped into [0, n), then a select,
        // so that all of a thread's loads are in flight at once
        uint32_t q[CH / 1024];
#pragma unroll
        for (int j = 0; j < CH / 1024; ++j) q[j] = nbytes[min(c0 + t + 1024 * j, n - 1)];
#pragma unroll
        for (int j = 0; j < CH / 1024; ++j) P[t + 1024 * j] = c0 + t + 1024 * j < n ? q[j] + 4u : 0u;
        __syncthreads();
        uint32_t v[CH / 1024], sum = 0;
#pragma unroll
        for (int j = 0; j < CH / 1024; ++j) {
            v[j] = P[(CH / 1024) * t + j];
            sum += v[j];
        }
        uint32_t tot;
        uint32_t run = block_excl_scan<1024>(sum, wsum, &tot);
        __syncthreads();  // every thread has read its sizes
#pragma unroll
        for (int j = 0; j < CH / 1024; ++j) {
            P[(CH / 1024) * t + j] = run;
            run += v[j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < CH / 1024; ++j) {
            const int i = c0 + t + 1024 * j;
            if (i < n) offsets[i] = carry + P[t + 1024 * j];
        }
        carry += tot;
        __syncthreads();  // P and wsum are reused by the next chunk
    }
    if (t == 0) offsets[n] = carry;
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
// PK_B blocks per workgroup, the first PK_W survivor words of each block
// loaded before any store.  With fused != 0 the workgroup finds its records'
// offsets itself (sum of the sizes before it: O(nblocks^2) cached reads in
// all, used for up to PK_FUSED_MAX blocks) and writes offsets[]; otherwise
// offsets[] comes from k_scan_records.
constexpr int PK_T = 512, PK_W = 4, PK_B = 2, PK_FUSED_MAX = 4096;
__global__ __launch_bounds__(PK_T) void k_pack_records(const uint8_t *__restrict__ slots, size_t slot_bytes,
                                                       const uint32_t *__restrict__ nbytes, int nblocks, int shuffle,
                                                       int fused, uint64_t *__restrict__ offsets,
                                                       uint8_t *__restrict__ records)
{
    __shared__ uint64_t wsum[PK_T / 64];
    __shared__ uint64_t start[PK_B];
    const int b0 = blockIdx.x * PK_B, t = threadIdx.x;
    const int nbk = min(PK_B, nblocks - b0);
    // the slots' loads do not depend on the offsets: issued first, they are
    // in flight while the offsets are found
    constexpr int MW = 1024 / PK_T;
    uint32_t m[PK_B][MW], v[PK_B][PK_W];
#pragma unroll
    for (int kb = 0; kb < PK_B; ++kb) {
        if (kb < nbk) {
            const int b = b0 + kb;
            const uint32_t S = (nbytes[b] - MASK_BYTES) / 4, sc = nbytes[nblocks + b];
            const uint32_t *src = reinterpret_cast<const uint32_t *>(slots + (size_t)b * slot_bytes);
            const uint32_t *corner = src + MASK_BYTES / 4, *l32 = corner + CORNER_MAX;
#pragma unroll
            for (int k = 0; k < MW; ++k) m[kb][k] = src[t + PK_T * k];
#pragma unroll
            for (int k = 0; k < PK_W; ++k) {
                const uint32_t i = t + PK_T * k;
                v[kb][k] = i < S ? (i < sc ? corner[i] : l32[i - sc]) : 0u;
            }
        }
    }
    if (fused) {
        uint64_t part = 0;
        for (int c = 0; c < b0; c += 8 * PK_T) {
            // unconditional loads at an index clamped into [0, b0), then a
            // select: a load under a branch waits for itself before the next
            uint32_t q[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) q[k] = nbytes[min(c + t + PK_T * k, b0 - 1)];
#pragma unroll
            for (int k = 0; k < 8; ++k) q[k] = c + t + PK_T * k < b0 ? q[k] + 4u : 0u;
#pragma unroll
            for (int k = 0; k < 8; ++k) part += q[k];
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) part += __shfl_xor(part, d, 64);
        if ((t & 63) == 0) wsum[t >> 6] = part;
        __syncthreads();
        if (t == 0) {
            uint64_t o = 0;
#pragma unroll
            for (int k = 0; k < PK_T / 64; ++k) o += wsum[k];
            for (int k = 0; k < nbk; ++k) {
                start[k] = o;
                offsets[b0 + k] = o;
                o += (uint64_t)nbytes[b0 + k] + 4u;
            }
            if (b0 + nbk == nblocks) offsets[nblocks] = o;
        }
    } else if (t < nbk) {
        start[t] = offsets[b0 + t];
    }
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < PK_B; ++kb) {
        if (kb < nbk) {
            const int b = b0 + kb;
            const uint32_t n = nbytes[b];
            const uint32_t S = (n - MASK_BYTES) / 4, sc = nbytes[nblocks + b];
            uint32_t *dst = reinterpret_cast<uint32_t *>(records + start[kb]);
            uint8_t *dst8 = records + start[kb] + 4 + MASK_BYTES;
            const uint32_t *src = reinterpret_cast<const uint32_t *>(slots + (size_t)b * slot_bytes);
            const uint32_t *corner = src + MASK_BYTES / 4, *l32 = corner + CORNER_MAX;
            auto put = [&](uint32_t i, uint32_t w) {
                if (shuffle) {
                    dst8[i] = (uint8_t)w;
                    dst8[S + i] = (uint8_t)(w >> 8);
                    dst8[2 * S + i] = (uint8_t)(w >> 16);
                    dst8[3 * S + i] = (uint8_t)(w >> 24);
                } else {
                    dst[1 + MASK_BYTES / 4 + i] = w;
                }
            };
            if (t == 0) dst[0] = n;
#pragma unroll
            for (int k = 0; k < MW; ++k) dst[1 + t + PK_T * k] = m[kb][k];
#pragma unroll
            for (int k = 0; k < PK_W; ++k) {
                const uint32_t i = t + PK_T * k;
                if (i < S) put(i, v[kb][k]);
            }
            for (uint32_t c0 = PK_T * PK_W; c0 < S; c0 += PK_T * PK_W) {
                uint32_t w[PK_W];
#pragma unroll
                for (int k = 0; k < PK_W; ++k) {
                    const uint32_t i = c0 + t + PK_T * k;
                    w[k] = i < S ? (i < sc ? corner[i] : l32[i - sc]) : 0u;
                }
#pragma unroll
                for (int k = 0; k < PK_W; ++k) {
                    const uint32_t i = c0 + t + PK_T * k;
                    if (i < S) put(i, w[k]);
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
    const dim3 grid(std::min(nblocks, resident_groups(fn, CT))), blk(CT);
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
    const int fused = nblocks <= PK_FUSED_MAX;
    if (!fused && scan_records(nb, nblocks, offsets, s) != CZW_OK) return CZW_EHIP;
    hipLaunchKernelGGL(k_pack_records, dim3((nblocks + PK_B - 1) / PK_B), dim3(PK_T), 0, s, slots, slot_bytes, nb,
                       nblocks, shuffle, fused, offsets, records);
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
    // at most the resident groups: a groups value tuned for the decoder (two
    // per CU) would otherwise queue a second persistent wave behind the first
    const int R = resident_groups(fn, CT);
    const int groups = B->groups ? std::min(B->groups, R) : R;
    const dim3 grid(std::min(B->nblocks, groups)), blk(CT);
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
    uint32_t *q = B->queue ? batch_queue(B) + 1 : nullptr;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, B->err, 1, 0u);
    if (q) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, q, 1, 0u);
    return decompress_half(B->prec, LAYOUT_BATCH, B->records, B->records_bytes, B->offsets, B->nblocks,
                           FieldGeom{0, 0, 0}, B->wtype, B->shuffle, nullptr, B->err, B->fields, B->nfields, B->groups,
                           q, s);
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
    if (prec != 4 && prec != 8) return CZW_EINVAL;
    return decompress_half(prec, buf ? LAYOUT_FIELD : LAYOUT_FIELD_WIDE, records, records_bytes, offsets, nb, g,
                           wtype, shuffle, field, err, nullptr, 0, 0, nullptr, stream);
}

int czw_decompress_blocks(const void *records, size_t records_bytes, const uint64_t *offsets, int nblocks, int prec,
                          int wtype, int shuffle, void *blocks, uint32_t *err, hipStream_t stream)
{
    const FieldGeom g{0, 0, 0};
    if (prec != 4 && prec != 8) return CZW_EINVAL;
    return decompress_half(prec, LAYOUT_BLOCKS, records, records_bytes, offsets, nblocks, g, wtype, shuffle, blocks,
                           err, nullptr, 0, 0, nullptr, stream);
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
    const int R = resident_groups(fn, CT);  // the compressor's: one group per CU
    // the compressor is tried at R and R/2 groups; the decoder, which fits two
    // groups per CU for binary32, at 2R and R
    const int G1 = stage == 1 ? R : 2 * R, G2 = stage == 1 ? (R / 2 > 0 ? R / 2 : 1) : R;
    const int shapes[4][2] = {{G1, 0}, {G1, 1}, {G2, 0}, {G2, 1}};
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
