// czw_half.hip -- block compressor and decompressor that keep one x-half of a
// 32^3 block in LDS at a time (WaveletCompressor.cpp:153-259).  The
// compressor serves Real = double (the reference's double build: a 256 KB
// block does not fit in the 160 KB of LDS); the decompressor serves double
// and float (a 64 KB half lets two workgroups share a CU, so one's memory
// phases overlap the other's arithmetic):
//
//   compress:   level-32 x-pass on the thread's two x-rows in registers;
//               x >= 16 half -> LDS: y-pass, z-pass, threshold, survivors
//               (all final: level-32 codes 2, 3, 6, 7);
//               x <  16 half -> LDS: y-pass, z-pass (+ level-16 z-pass on the
//               [0,16)^2 columns), survivors of codes 1, 4, 5;
//               the [0,16)^3 corner stays in LDS for levels 16 and 8.
//   decompress: the mirror image, x < 16 half first (it holds the corner),
//               then the x >= 16 half, then the level-32 x-pass in registers.
//
// Because the two halves are handled one after the other, the survivors of
// each half go to their own region of the block's slot with per-code counts
// (see SLOT64 below); the packing kernel concatenates the regions in stream
// order (codes 0..7, FullWaveletTransform.h:59-96).  The arithmetic is the
// reference's: predictors in binary64, rounded to Real (czw_wavelet.hpp).
//
// LDS layout of a half: H[z][y][x & 15] at z*SPD + y*RPD + (x & 15) with
// RPD = 17, SPD = 545 (elements): x-rows for either thread order, y- and
// z-rows and the corner passes are at most 2-way bank conflicted for both
// element sizes.
#include <type_traits>

#include "czw_common.hpp"

namespace czw {

constexpr int RPD = 17;
constexpr int SPD = 32 * RPD + 1;
constexpr int HALF_WORDS = 32 * SPD;  // elements
template <typename R>
constexpr size_t slot_of() { return MASK_BYTES + sizeof(R) * (size_t)NB3; }
// Survivor regions of a slot, in elements after the 4096-byte mask: the
// [0,16)^3 corner (stream positions [0,4096), <= 4096 values), then the
// x < 16 half's level-32 codes 1, 4, 5 (<= 3 x 4096), then the x >= 16
// half's codes 2, 3, 6, 7 (<= 4 x 4096): 8 x 4096 elements in all.
constexpr int REG_CORNER = 0;
constexpr int REG_H0 = 4096;
constexpr int REG_H1 = 4 * 4096;
constexpr int NCNT = 8;  // per-block counts: corner, then level-32 codes 1..7

// x-row q (0..1023) of a block: (z, y) = (q >> 5, q & 31) in the block
// layout, (y, z) = (q >> 5, q & 31) in the HDF5 field layout (as for
// binary32); its x-half in H starts here.
template <int RL, int RP_ = RPD, int SP_ = SPD>
CZW_DI int xrow_half_base(int q)
{
    if constexpr (RL == LAYOUT_BLOCKS) return (q >> 5) * SP_ + (q & 31) * RP_;
    else return (q & 31) * SP_ + (q >> 5) * RP_;
}

// The decoder's x-row k of thread t: z = t & 31 and y = (t >> 5) + 16k in
// either layout, so that a half-wave's LDS accesses to its rows' x-halves
// step through z (one bank apart with SP_ odd).
template <int RL>
CZW_DI int dec_xrow(int t, int k)
{
    if constexpr (RL == LAYOUT_BLOCKS) return ((t & 31) << 5) | ((t >> 5) + 16 * k);
    else return t + NT * k;
}

template <typename R>
CZW_DI R from_bits(czw_u32x2 v);
template <>
CZW_DI double from_bits<double>(czw_u32x2 v)
{
    return __longlong_as_double((long long)(((unsigned long long)v.y << 32) | v.x));
}

// x in [x0, x0 + n) of x-row q into r[x0 .. x0 + n).
template <typename R, int RL, int N = 32>
CZW_DI void load_xrow_h(const R *__restrict__ src, const FieldGeom &g, int b, int q, R *r, int x0 = 0)
{
    if constexpr (RL == LAYOUT_BLOCKS) {
        const R *p = src + (size_t)b * NB3 + (size_t)q * 32 + x0;
        if constexpr (sizeof(R) == 8) {
#pragma unroll
            for (int i = 0; i < N / 2; ++i) {
                const czw_f64x2 v = __builtin_nontemporal_load(reinterpret_cast<const czw_f64x2 *>(p) + i);
                r[x0 + 2 * i] = v.x;
                r[x0 + 2 * i + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int i = 0; i < N / 4; ++i) {
                const czw_f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const czw_f32x4 *>(p) + i);
                r[x0 + 4 * i] = v.x;
                r[x0 + 4 * i + 1] = v.y;
                r[x0 + 4 * i + 2] = v.z;
                r[x0 + 4 * i + 3] = v.w;
            }
        }
    } else if constexpr (RL == LAYOUT_FIELD) {
        size_t xs;
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(src + field_row(g, b, 0, 0, &xs));
        const int vo = ((q >> 5) * g.nz + (q & 31)) * (int)sizeof(R), so = (int)xs * (int)sizeof(R);
#pragma unroll
        for (int x = x0; x < x0 + N; ++x) {
            if constexpr (sizeof(R) == 8)
                r[x] = from_bits<double>(__builtin_amdgcn_raw_buffer_load_b64(rs, vo, x * so, CZW_FIELD_CPOL));
            else
                r[x] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, vo, x * so, CZW_FIELD_CPOL));
        }
    } else {
        size_t xs;
        const R *p = src + field_row(g, b, q >> 5, q & 31, &xs);
#pragma unroll
        for (int x = x0; x < x0 + N; ++x) r[x] = __builtin_nontemporal_load(p + x * xs);
    }
}

template <typename R, int RL>
CZW_DI void store_xrow_h(R *__restrict__ dst, const FieldGeom &g, int b, int q, const R *r)
{
    if constexpr (RL == LAYOUT_BLOCKS) {
        R *p = dst + (size_t)b * NB3 + (size_t)q * 32;
        if constexpr (sizeof(R) == 8) {
#pragma unroll
            for (int i = 0; i < 16; ++i)
                __builtin_nontemporal_store(czw_f64x2{r[2 * i], r[2 * i + 1]}, reinterpret_cast<czw_f64x2 *>(p) + i);
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i)
                __builtin_nontemporal_store(czw_f32x4{r[4 * i], r[4 * i + 1], r[4 * i + 2], r[4 * i + 3]},
                                            reinterpret_cast<czw_f32x4 *>(p) + i);
        }
    } else if constexpr (RL == LAYOUT_FIELD) {
        size_t xs;
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(dst + field_row(g, b, 0, 0, &xs));
        const int vo = ((q >> 5) * g.nz + (q & 31)) * (int)sizeof(R), so = (int)xs * (int)sizeof(R);
#pragma unroll
        for (int x = 0; x < 32; ++x) {
            if constexpr (sizeof(R) == 8) {
                const unsigned long long u = (unsigned long long)__double_as_longlong(r[x]);
                const czw_u32x2 v = {(unsigned)u, (unsigned)(u >> 32)};
                __builtin_amdgcn_raw_buffer_store_b64(v, rs, vo, x * so, CZW_FIELD_CPOL);
            } else {
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(r[x]), rs, vo, x * so, CZW_FIELD_CPOL);
            }
        }
    } else {
        size_t xs;
        R *p = dst + field_row(g, b, q >> 5, q & 31, &xs);
#pragma unroll
        for (int x = 0; x < 32; ++x) __builtin_nontemporal_store(r[x], p + x * xs);
    }
}

// float_zero_bits3 (WaveletCompressor.cpp:34-43) on the first 32-bit word of
// a little-endian Real: the whole float, or a double's low mantissa word.
CZW_DI double zero_low_bits(double v, uint32_t zkeep)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return __longlong_as_double((long long)(u & (0xffffffff00000000ull | (unsigned long long)zkeep)));
}
CZW_DI float zero_low_bits(float v, uint32_t zkeep) { return __uint_as_float(__float_as_uint(v) & zkeep); }
CZW_DI float absval(float v) { return fabsf(v); }
CZW_DI double absval(double v) { return fabs(v); }

// Stage of the x < 16 half's survivors: the z >= 16 planes of H.
constexpr int STAGE0_AT = 16 * SPD;
constexpr int STAGE0 = HALF_WORDS - STAGE0_AT - 64;  // 64 dummy words follow
static_assert(16384 + 64 <= HALF_WORDS, "room for the x >= 16 half's stage and its dummy words");

// One z-row's survivors into an LDS stage st at run offsets i0 (z < 16) and
// i1 (z >= 16), zerobits applied.  Groups of eight positions that no row of
// the wave keeps are skipped on a scalar test; a kept group runs branch-free,
// unkept values going to the lane's dummy word (a test per position costs
// taken branches).  Wavelet type 2 (the lifting step holds more registers)
// keeps the per-position form: the grouped one spills there (binary64
// compress +10%).
template <typename R, int WT>
CZW_DI void stage_row(R *st, R *dummy, uint32_t bits, uint32_t i0, uint32_t i1, const R *r, uint32_t zkeep)
{
    if constexpr (WT == 2) {
        (void)dummy;
#pragma unroll
        for (int z = 0; z < 32; ++z) {
            uint32_t &i = z < 16 ? i0 : i1;
            if ((bits >> z) & 1u) st[i++] = zero_low_bits(r[z], zkeep);
        }
        return;
    }
    const uint32_t wk = wave_or(bits);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        if (!((wk >> (8 * c)) & 0xffu)) continue;
#pragma unroll
        for (int z = 8 * c; z < 8 * c + 8; ++z) {
            const uint32_t bit = (bits >> z) & 1u;
            uint32_t &i = z < 16 ? i0 : i1;
            R *p = bit ? st + i : dummy;
            *p = zero_low_bits(r[z], zkeep);
            i += bit;
        }
    }
}

// n staged values (LDS, 16-byte aligned) to global memory (16-byte aligned),
// 16 bytes per store where possible.
template <typename R>
CZW_DI void copy_out(R *__restrict__ dst, const R *src, uint32_t n, int t)
{
    constexpr uint32_t V = 16 / sizeof(R);
    const uint32_t nv = n / V;
    for (uint32_t i = t; i < nv; i += NT) reinterpret_cast<uint4 *>(dst)[i] = reinterpret_cast<const uint4 *>(src)[i];
    for (uint32_t i = nv * V + t; i < n; i += NT) dst[i] = src[i];
}

// Exclusive scan of cnt[0..1024) in place (two entries per thread); returns
// the total.  Barriers before and after are the caller's.
CZW_DI uint32_t scan1024(uint32_t *cnt, uint32_t *wsum, int t)
{
    const uint32_t c0 = cnt[2 * t], c1 = cnt[2 * t + 1];
    uint32_t total;
    const uint32_t run = block_excl_scan(c0 + c1, wsum, &total);
    cnt[2 * t] = run;
    cnt[2 * t + 1] = run + c0;
    return total;
}

// ---------------------------------------------------------------------------
// Compressor.  Persistent like the binary32 one (static stride or work queue,
// LAYOUT_BATCH field tables).  Block b's survivors go to the regions of slot
// b, its payload size to nbytes[b] and its 8 region counts to
// counts[8b .. 8b+8).
template <typename R, int WT, int LAYOUT>
__global__ __launch_bounds__(NT, sizeof(R) == 4 ? 4 : 2) void k_compress_h(const R *__restrict__ src, FieldGeom g, int nblocks,
                                                     float eps, uint32_t zkeep, uint8_t *__restrict__ slots,
                                                     uint32_t *__restrict__ nbytes, uint32_t *__restrict__ counts,
                                                     const czw_field_t *__restrict__ fields, int nf,
                                                     uint32_t *__restrict__ queue)
{
    constexpr int RL = row_layout<LAYOUT>();
    __shared__ __align__(16) R H[HALF_WORDS];
    __shared__ uint32_t M[1024];
    __shared__ uint32_t cnt[1024];
    __shared__ uint32_t wsum[NT / 64];
    __shared__ int ticket[2];
    __shared__ czw_field_t T[LAYOUT == LAYOUT_BATCH ? BATCH_MAX_FIELDS : 1];
    const int t = threadIdx.x;
    const int G = gridDim.x;
    load_table<LAYOUT>(T, fields, nf);
    load_one<LAYOUT>(T, src, nullptr, g, eps);
    // pass-row owners: z-pass row (y, x & 15) = (t >> 4, t & 15); y-pass row
    // (z, x & 15) = (yz, t & 15) with yz = (t >> 5) + 16*((t >> 4) & 1), so
    // that a half-wave's 32 lanes take z and z + 16 (16 banks apart, as SPD
    // is 1 modulo 32) instead of z and z + 1 (one bank apart, 2-way conflicts
    // in binary32)
    const int xh = t & 15, sl = t >> 4, yz = (t >> 5) + 16 * ((t >> 4) & 1);
    for (int b = blockIdx.x, it = 0; b < nblocks; ++it) {
        if (queue && t == 0) ticket[it & 1] = G + (int)atomicAdd(queue, 1u);
        const BlockRef<R> br = locate<table_layout<LAYOUT>(), R>(T, one_field<LAYOUT>() ? 1 : nf, b, src, nullptr, g, eps);
        const R e = (R)br.eps;  // fabs(c) > threshold in the coefficient type (FullWaveletTransform.h:84)
        uint8_t *slot = slots + (size_t)b * slot_of<R>();
        R *reg = reinterpret_cast<R *>(slot + MASK_BYTES);
        R xr[NROW][32];
#pragma unroll
        for (int k = 0; k < NROW; ++k) {  // level 32, x-pass, in registers
            load_xrow_h<R, RL>(br.src, br.g, br.lb, t + NT * k, xr[k]);
            row<32, true, WT, R>(xr[k]);
        }

        // ---- x >= 16: level-32 y- and z-passes; every coefficient is final
#pragma unroll
        for (int k = 0; k < NROW; ++k) lds_st<16>(H, xrow_half_base<RL>(t + NT * k), 1, &xr[k][16]);
        __syncthreads();
        {
            R r[32];
            lds_ld<32>(H, yz * SPD + xh, RPD, r);  // y-pass: row (z, x)
            row<32, true, WT, R>(r);
            lds_st<32>(H, yz * SPD + xh, RPD, r);
        }
        __syncthreads();
        uint32_t S1;
        {
            const int y = sl, x = 16 + xh;
            R r[32];
            lds_ld<32>(H, y * RPD + xh, SPD, r);  // z-pass: row (y, x)
            row<32, true, WT, R>(r);
            uint32_t bits = 0;
#pragma unroll
            for (int z = 0; z < 32; ++z) bits |= (uint32_t)(absval(r[z]) > e) << z;
            M[x + 32 * y] = bits;
            // region order: codes 2, 3 (y < 16), 6, 7 (y >= 16); code bit 0 = z >= 16
            const int ri = 2 * (y >> 4), base = (y & 15) * 16 + xh;
            cnt[ri * 256 + base] = __popc(bits & 0xffffu);
            cnt[(ri + 1) * 256 + base] = __popc(bits >> 16);
            __syncthreads();
            S1 = scan1024(cnt, wsum, t);
            __syncthreads();
            if (t == 0) {
                uint32_t *cb = counts + (size_t)NCNT * b;
                cb[2] = cnt[256] - cnt[0];
                cb[3] = cnt[512] - cnt[256];
                cb[6] = cnt[768] - cnt[512];
                cb[7] = S1 - cnt[768];
            }
            // every column is in registers: H becomes the stage of this half's
            // survivors (at most 16384), copied out with 16-byte stores
            stage_row<R, WT>(H, H + 16384 + (t & 63), bits, cnt[ri * 256 + base], cnt[(ri + 1) * 256 + base], r, zkeep);
        }
        __syncthreads();
        copy_out<R>(reg + REG_H1, H, S1, t);
        __syncthreads();  // H and cnt are free again

        // ---- x < 16: level-32 y- and z-passes, level-16 z-pass on the corner columns
#pragma unroll
        for (int k = 0; k < NROW; ++k) lds_st<16>(H, xrow_half_base<RL>(t + NT * k), 1, &xr[k][0]);
        __syncthreads();
        {
            R r[32];
            lds_ld<32>(H, yz * SPD + xh, RPD, r);
            row<32, true, WT, R>(r);
            lds_st<32>(H, yz * SPD + xh, RPD, r);
        }
        __syncthreads();
        uint32_t S0;
        {
            const int y = sl, x = xh;
            const bool corner = y < 16;
            R r[32];
            lds_ld<32>(H, y * RPD + xh, SPD, r);
            row<32, true, WT, R>(r);
            uint32_t bits = 0;
#pragma unroll
            for (int z = 0; z < 32; ++z) bits |= (uint32_t)(absval(r[z]) > e) << z;
            if (corner) {
                R c[16];
#pragma unroll
                for (int z = 0; z < 16; ++z) c[z] = r[z];
                row<16, true, WT, R>(c);
                lds_st<16>(H, y * RPD + xh, SPD, c);
                bits &= 0xffff0000u;  // z < 16 is decided with the corner below
            }
            M[x + 32 * y] = bits;
            // region order: code 1 (y < 16, z >= 16), code 4 (y >= 16, z < 16), code 5
            const int base = (y & 15) * 16 + xh;
            if (corner) {
                cnt[base] = __popc(bits >> 16);
            } else {
                cnt[256 + base] = __popc(bits & 0xffffu);
                cnt[512 + base] = __popc(bits >> 16);
            }
            if (t < 256) cnt[768 + t] = 0u;
            __syncthreads();
            S0 = scan1024(cnt, wsum, t);
            __syncthreads();
            if (t == 0) {
                uint32_t *cb = counts + (size_t)NCNT * b;
                cb[1] = cnt[256] - cnt[0];
                cb[4] = cnt[512] - cnt[256];
                cb[5] = S0 - cnt[512];
            }
            // the z >= 16 planes of H are free (the corner needs z < 16 only): they
            // stage this half's survivors when they fit, else these go out directly
            uint32_t i0 = corner ? 0u : cnt[256 + base], i1 = corner ? cnt[base] : cnt[512 + base];
            if (S0 <= (uint32_t)STAGE0) {
                stage_row<R, WT>(H + STAGE0_AT, H + STAGE0_AT + STAGE0 + (t & 63), bits, i0, i1, r, zkeep);
            } else {
                R *st = reg + REG_H0;
#pragma unroll
                for (int z = 0; z < 32; ++z) {
                    uint32_t &i = z < 16 ? i0 : i1;
                    if ((bits >> z) & 1u) st[i++] = zero_low_bits(r[z], zkeep);
                }
            }
        }
        // levels 16 and 8 on the corner (the barrier inside orders the column stores above)
        __syncthreads();
        if (S0 <= (uint32_t)STAGE0) copy_out<R>(reg + REG_H0, H + STAGE0_AT, S0, t);
        fwt16_8<WT, RPD, SPD>(H, t);

        // the [0,16)^3 corner: stream positions [0,4096) in pieces of 8 (base 4^3 always kept)
        uint32_t keep = 0;
        R v[8];
        corner_piece<RPD, SPD>(t, [&](int q, int h, int addr, int stride, int word, int shift) {
            uint32_t bits = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (i < h) {
                    const R c = H[addr + i * stride];
                    v[q * h + i] = c;
                    bits |= (uint32_t)((t < 8) || (absval(c) > e)) << i;
                }
            }
            keep |= bits << (q * h);
            atomicOr(&M[word], bits << shift);
        });
        uint32_t Sc;
        uint32_t idx = block_excl_scan(__popc(keep), wsum, &Sc);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if ((keep >> j) & 1u) reg[REG_CORNER + idx++] = zero_low_bits(v[j], zkeep);
        __syncthreads();  // mask complete
#pragma unroll
        for (int k = 0; k < NROW; ++k) reinterpret_cast<uint32_t *>(slot)[t + NT * k] = M[t + NT * k];
        if (t == 0) {
            counts[(size_t)NCNT * b] = Sc;
            nbytes[b] = MASK_BYTES + (uint32_t)sizeof(R) * (Sc + S0 + S1);
        }
        const int bnext = queue ? ticket[it & 1] : b + G;
        __syncthreads();  // M, cnt, wsum and ticket are read before the next block reuses them
        b = bnext;
    }
}

// Frame each block's payload into the record stream: size word, mask, then
// the survivors in stream order (corner, codes 1..7 from their regions), with
// the byte-plane shuffle for sizeof(Real) if asked (WaveletCompressor.cpp:56-73).
// Records start at 4-byte offsets, so doubles are written as two words.
template <typename R>
__global__ __launch_bounds__(256) void k_pack_records_h(const uint8_t *__restrict__ slots,
                                                        const uint32_t *__restrict__ nbytes,
                                                        const uint32_t *__restrict__ counts, int shuffle,
                                                        const uint64_t *__restrict__ offsets,
                                                        uint8_t *__restrict__ records)
{
    constexpr int W = (int)sizeof(R) / 4;  // words per survivor
    const int b = blockIdx.x, t = threadIdx.x;
    const uint32_t n = nbytes[b];
    const uint32_t S = (n - MASK_BYTES) / (uint32_t)sizeof(R);
    const uint32_t *cb = counts + (size_t)NCNT * b;
    uint32_t c[NCNT];
#pragma unroll
    for (int i = 0; i < NCNT; ++i) c[i] = cb[i];
    // region start of each stream segment (corner, codes 1..7)
    const uint32_t from[NCNT] = {(uint32_t)REG_CORNER, (uint32_t)REG_H0, (uint32_t)REG_H1, REG_H1 + c[2],
                                 REG_H0 + c[1], REG_H0 + c[1] + c[4], REG_H1 + c[2] + c[3],
                                 REG_H1 + c[2] + c[3] + c[6]};
    uint32_t st[NCNT + 1];
    st[0] = 0;
#pragma unroll
    for (int i = 0; i < NCNT; ++i) st[i + 1] = st[i] + c[i];
    const uint8_t *slot = slots + (size_t)b * slot_of<R>();
    const uint32_t *reg = reinterpret_cast<const uint32_t *>(slot + MASK_BYTES);
    uint32_t *dst = reinterpret_cast<uint32_t *>(records + offsets[b]);
    uint8_t *dst8 = records + offsets[b] + 4 + MASK_BYTES;
    if (t == 0) dst[0] = n;
    {
        uint32_t m[MASK_BYTES / 4 / 256];
#pragma unroll
        for (int k = 0; k < MASK_BYTES / 4 / 256; ++k) m[k] = reinterpret_cast<const uint32_t *>(slot)[t + 256 * k];
#pragma unroll
        for (int k = 0; k < MASK_BYTES / 4 / 256; ++k) dst[1 + t + 256 * k] = m[k];
    }
    // U survivors per thread per round, all loaded before the first store
    // (one memory round trip per round instead of per survivor)
    constexpr int U = 8;
    for (uint32_t i0 = t; i0 < S; i0 += 256 * U) {
        uint32_t w[U][W];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + 256 * u;
            if (i < S) {
                int seg = 0;
#pragma unroll
                for (int j = 1; j < NCNT; ++j) seg += (i >= st[j]) ? 1 : 0;
                const size_t e = from[seg] + (i - st[seg]);
                if constexpr (W == 2) {
                    const uint2 v = reinterpret_cast<const uint2 *>(reg)[e];
                    w[u][0] = v.x;
                    w[u][1] = v.y;
                } else {
                    w[u][0] = reg[e];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + 256 * u;
            if (i < S) {
                if (shuffle) {
#pragma unroll
                    for (int p = 0; p < 4 * W; ++p)
                        dst8[(size_t)p * S + i] = (uint8_t)(w[u][p >> 2] >> (8 * (p & 3)));
                } else {
#pragma unroll
                    for (int k = 0; k < W; ++k) dst[1 + MASK_BYTES / 4 + (size_t)W * i + k] = w[u][k];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Decompressor (WaveletCompressor.cpp:232-259 with FullWaveletTransform.h:
// 100-150 and the inverse transform), one block per workgroup iteration.
// Survivors are read straight from the record (they are at 4-byte offsets;
// byte planes when shuffled).  SHUF is a template parameter so that the
// per-position gathers carry no test of it.
// LDS the binary32 decoder has left for staging a block's level-32
// survivors: two workgroups per CU (2 x 81920 bytes of the 160 KB) with the
// 64 KB half, mask and counts (and a batch's 64-entry field table).
// Binary64 (one workgroup per CU, denser records) reads them from the record.
template <typename R, int LAYOUT>
constexpr int DEC_LS_CAP()
{
    return sizeof(R) != 4 ? 1 : LAYOUT == LAYOUT_BATCH ? 1088 : 1728;
}

// v where bit is 1, +0 where it is 0, by a mask (no select on v's value).
CZW_DI float keep_if(float v, uint32_t bit) { return __uint_as_float(__float_as_uint(v) & (0u - bit)); }
CZW_DI double keep_if(double v, uint32_t bit)
{
    return __longlong_as_double(__double_as_longlong(v) & (long long)(0ull - (unsigned long long)bit));
}

template <typename R, bool SHUF>
struct Payload {
    const uint32_t *w;  // unshuffled: survivor i = words W*i .. W*i + W - 1
    const uint8_t *b8;  // shuffled: byte p of survivor i at p*S + i
    uint32_t S;
    CZW_DI R get(uint32_t i) const
    {
        if constexpr (sizeof(R) == 8) {
            unsigned long long u;
            if constexpr (SHUF) {
                u = 0;
#pragma unroll
                for (int p = 0; p < 8; ++p) u |= (unsigned long long)b8[(size_t)p * S + i] << (8 * p);
            } else {
                u = (unsigned long long)w[2 * i] | ((unsigned long long)w[2 * i + 1] << 32);
            }
            return __longlong_as_double((long long)u);
        } else {
            uint32_t u;
            if constexpr (SHUF) {
                u = (uint32_t)b8[i] | ((uint32_t)b8[S + i] << 8) | ((uint32_t)b8[2 * (size_t)S + i] << 16) |
                    ((uint32_t)b8[3 * (size_t)S + i] << 24);
            } else {
                u = w[i];
            }
            return __uint_as_float(u);
        }
    }
};

template <typename R, int WT, int LAYOUT, bool SHUF>
__global__ __launch_bounds__(NT, sizeof(R) == 4 ? 4 : 2) void k_decompress_h(const uint8_t *__restrict__ records, uint64_t rec_bytes,
                                                       const uint64_t *__restrict__ offsets, int nblocks,
                                                       R *__restrict__ dst, FieldGeom g,
                                                       uint32_t *__restrict__ err,
                                                       const czw_field_t *__restrict__ fields, int nf,
                                                       uint32_t *__restrict__ queue)
{
    constexpr int RL = row_layout<LAYOUT>();
    // LDS layout of a half: binary32 rows of 16 (no pad) in planes of 513,
    // so that the y-pass (z and z + 16 in a half-wave) and the z-pass (y and
    // y + 1) both hit 32 distinct banks; binary64 keeps RPD/SPD (two banks
    // per element, a quarter-wave per pass)
    constexpr int RPd = sizeof(R) == 4 ? 16 : RPD;
    constexpr int SPd = 32 * RPd + 1;
    __shared__ R H[32 * SPd];
    __shared__ uint32_t M[1024];
    __shared__ uint16_t cnt[2560];  // corner pieces [0,512), level-32 runs 512 + (code-1)*256 + run
    __shared__ R Ls[DEC_LS_CAP<R, LAYOUT>()];  // a block's level-32 survivors, when they fit
    __shared__ uint32_t wsum[NT / 64];
    __shared__ uint32_t flag;
    __shared__ int ticket[2];
    __shared__ czw_field_t T[LAYOUT == LAYOUT_BATCH ? BATCH_MAX_FIELDS : 1];
    const int t = threadIdx.x;
    const int G = gridDim.x;
    load_table<LAYOUT>(T, fields, nf);
    load_one<LAYOUT>(T, nullptr, dst, g, 0.0f);
    // z-pass rows (y, x & 15) = (t >> 4, t & 15); y-pass rows (yz, t & 15), see k_compress_h
    const int xh = t & 15, sl = t >> 4, yz = (t >> 5) + 16 * ((t >> 4) & 1);
    int bn = 0;
    // The record offset of the next block is read one iteration ahead, and a
    // block's size word and mask are read together, so that a block starts
    // after one memory round trip instead of three dependent ones.
    uint64_t obn = (int)blockIdx.x < nblocks ? offsets[blockIdx.x] : 0;
    for (int b = blockIdx.x, it = 0; b < nblocks; ++it, b = bn) {
        if (queue) {
            if (t == 0) ticket[it & 1] = G + (int)atomicAdd(queue, 1u);
            __syncthreads();
            bn = ticket[it & 1];
        } else {
            bn = b + G;
        }
        const uint64_t ob = obn;
        if (bn < nblocks) obn = offsets[bn];
        if ((ob & 3u) || ob + 4 + MASK_BYTES > rec_bytes) {
            if (t == 0) atomicOr(err, (uint32_t)CZW_ERR_NBYTES);
            continue;
        }
        // the mask lies inside the record (checked above); it is stored
        // before the size word is checked, so that both loads are in flight
        // together (a bad size leaves M to be overwritten by the next block)
        const uint32_t *p = reinterpret_cast<const uint32_t *>(records + ob);
        uint32_t mk[NROW];
#pragma unroll
        for (int k = 0; k < NROW; ++k) mk[k] = p[1 + t + NT * k];
        const uint32_t nb = p[0];
        __syncthreads();  // the previous block's last reads of H, M, cnt and flag are done
#pragma unroll
        for (int k = 0; k < NROW; ++k) M[t + NT * k] = mk[k];
        if (nb < MASK_BYTES || ((nb - MASK_BYTES) % sizeof(R)) || nb > slot_of<R>() || ob + 4 + nb > rec_bytes) {
            if (t == 0) atomicOr(err, (uint32_t)CZW_ERR_NBYTES);
            continue;
        }
        const uint32_t S = (nb - MASK_BYTES) / (uint32_t)sizeof(R);
        const Payload<R, SHUF> pay{p + 1 + MASK_BYTES / 4, records + ob + 4 + MASK_BYTES, S};
        if (t == 0) flag = 1u;
        __syncthreads();
        // the 4^3 base is always stored (FullWaveletTransform.h:163-176)
        if (t < 16 && (M[(t & 3) + 32 * (t >> 2)] & 0xfu) != 0xfu) flag = 0u;
        uint32_t keep = 0;
        corner_piece<RPd, SPd>(t, [&](int q, int h, int, int, int word, int shift) {
            const uint32_t hm = (1u << h) - 1u;
            keep |= ((t < 8) ? hm : ((M[word] >> shift) & hm)) << (q * h);
        });
        cnt[t] = __popc(keep);
#pragma unroll
        for (int k = 0; k < NROW; ++k) {
            const int q = t + NT * k, y = q >> 5, x = q & 31;
            const bool corner = (y < 16) && (x < 16);
            const uint32_t mw = M[x + 32 * y] & (corner ? 0xffff0000u : 0xffffffffu);
            const int c0 = 2 * (x >> 4) + 4 * (y >> 4), base = (y & 15) * 16 + (x & 15);
            if (!corner) cnt[512 + (c0 - 1) * 256 + base] = __popc(mw & 0xffffu);
            cnt[512 + c0 * 256 + base] = __popc(mw >> 16);
        }
        if (t < 2560 - 512 - 1792) cnt[512 + 1792 + t] = 0u;
        __syncthreads();
        {  // exclusive scan of the 2560 counts in stream order, 5 per thread
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
            // with all base bits set, the kept count equals the mask popcount
            // that WaveletCompressor.cpp:235-251 checks against the payload size
            if (tot != S) flag = 0u;
        }
        __syncthreads();
        if (!flag) {
            if (t == 0) atomicOr(err, (uint32_t)CZW_ERR_POPCOUNT);
            continue;
        }
        // The level-32 survivors [sc0, S) (sc0 = the corner's count) go to LDS
        // when they fit: loaded here in one coalesced pass, stored behind the
        // corner gather, so that the z-pass gathers below read LDS instead of
        // the record.  Otherwise those gathers read the record.
        const uint32_t sc0 = cnt[512];
        const uint32_t s32 = S - sc0;
        const bool staged = DEC_LS_CAP<R, LAYOUT>() > 1 && s32 <= (uint32_t)DEC_LS_CAP<R, LAYOUT>();
        constexpr int LS_PER = (DEC_LS_CAP<R, LAYOUT>() + NT - 1) / NT;
        R sv[LS_PER];
        if (staged) {
#pragma unroll
            for (int u = 0; u < LS_PER; ++u)
                if (t + NT * u < (int)s32) sv[u] = pay.get(sc0 + t + NT * u);
        }
        {  // corner positions: survivor or +0; all eight loads are issued
            // before the first LDS store, so that they overlap
            uint32_t idx = cnt[t];
            R v[8];
            corner_piece<RPd, SPd>(t, [&](int q, int h, int, int, int, int) {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    if (i < h) {
                        const bool bit = (keep >> (q * h + i)) & 1u;
                        v[q * h + i] = bit ? pay.get(idx) : (R)0;
                        idx += bit;
                    }
                }
            });
            corner_piece<RPd, SPd>(t, [&](int q, int h, int addr, int stride, int, int) {
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (i < h) H[addr + i * stride] = v[q * h + i];
            });
        }
        if (staged) {
#pragma unroll
            for (int u = 0; u < LS_PER; ++u)
                if (t + NT * u < (int)s32) Ls[t + NT * u] = sv[u];
        }
        __syncthreads();
        iwt16_8<WT, RPd, SPd>(H, t);  // level 8 (x, z, y), level 16 (y, x)
        // Survivors z in [z0, 32) of one z-row (mask word mw, clear below z0;
        // run offsets i0 for z < 16, i1 for z >= 16), or +0.  From the record:
        // one masked load per kept position.  From LDS: positions that no row
        // of the wave keeps are skipped on a scalar test; at the others every
        // lane reads (index from a popcount, clamped into Ls) and the mask bit
        // selects, so that no read waits for the one before.
        auto gather = [&](auto from_lds, uint32_t mw, uint32_t i0, uint32_t i1, int z0, R *r) {
            if constexpr (decltype(from_lds)::value) {
                const uint32_t wm = wave_or(mw);
                // the run offsets (LDS reads) are consumed before the first
                // branch, so that no wait inside one covers the gathers
                uint32_t j0 = i0 - sc0, j1 = i1 - sc0;
                asm volatile("" : "+v"(j0), "+v"(j1));
                // v[z] starts as an arbitrary but fixed value (a frozen
                // undefined one, no instruction): positions that the wave
                // skips are then well defined and the mask bit clears them
                R v[32];
#pragma unroll
                for (int z = 0; z < 32; ++z) {
                    v[z] = __builtin_nondeterministic_value((R)0);
                    if (z < z0 || !((wm >> z) & 1u)) continue;
                    const uint32_t below = mw & ((1u << z) - 1u) & (z < 16 ? 0xffffu : 0xffff0000u);
                    const uint32_t i = (z < 16 ? j0 : j1) + (uint32_t)__popc(below);
                    v[z] = Ls[min(i, (uint32_t)DEC_LS_CAP<R, LAYOUT>() - 1u)];
                }
#pragma unroll
                for (int z = 0; z < 32; ++z)
                    if (z >= z0) r[z] = keep_if(v[z], (mw >> z) & 1u);
            } else {
#pragma unroll
                for (int z = 0; z < 32; ++z) {
                    if (z < z0) continue;
                    uint32_t &i = z < 16 ? i0 : i1;
                    const bool bit = (mw >> z) & 1u;
                    r[z] = bit ? pay.get(i) : (R)0;
                    i += bit;
                }
            }
        };

        // ---- x < 16: level-16 z-pass on the corner columns, level-32 z- and y-passes
        {
            const int y = sl, x = xh, base = (y & 15) * 16 + xh;
            const uint32_t mw = M[x + 32 * y];
            R r[32];
            if (y < 16) {
                R c[16];
                lds_ld<16>(H, y * RPd + xh, SPd, c);
                row<16, false, WT, R>(c);
#pragma unroll
                for (int z = 0; z < 16; ++z) r[z] = c[z];
                const uint32_t i1 = cnt[512 + base];  // code 1
                if (staged) gather(std::true_type(), mw, 0u, i1, 16, r);
                else gather(std::false_type(), mw, 0u, i1, 16, r);
            } else {
                const uint32_t i0 = cnt[512 + 3 * 256 + base], i1 = cnt[512 + 4 * 256 + base];  // codes 4, 5
                if (staged) gather(std::true_type(), mw, i0, i1, 0, r);
                else gather(std::false_type(), mw, i0, i1, 0, r);
            }
            row<32, false, WT, R>(r);
            lds_st<32>(H, y * RPd + xh, SPd, r);
        }
        __syncthreads();
        {
            R r[32];
            lds_ld<32>(H, yz * SPd + xh, RPd, r);  // y-pass: row (z, x)
            row<32, false, WT, R>(r);
            lds_st<32>(H, yz * SPd + xh, RPd, r);
        }
        __syncthreads();
        R lo[NROW][16];
#pragma unroll
        for (int k = 0; k < NROW; ++k) lds_ld<16>(H, xrow_half_base<RL, RPd, SPd>(dec_xrow<RL>(t, k)), 1, lo[k]);
        __syncthreads();

        // ---- x >= 16: level-32 z- and y-passes (codes 2, 3 for y < 16; 6, 7 for y >= 16)
        {
            const int y = sl, x = 16 + xh, base = (y & 15) * 16 + xh;
            const int c0 = 2 + 4 * (y >> 4);
            const uint32_t mw = M[x + 32 * y];
            const uint32_t i0 = cnt[512 + (c0 - 1) * 256 + base], i1 = cnt[512 + c0 * 256 + base];
            R r[32];
            if (staged) gather(std::true_type(), mw, i0, i1, 0, r);
            else gather(std::false_type(), mw, i0, i1, 0, r);
            row<32, false, WT, R>(r);
            lds_st<32>(H, y * RPd + xh, SPd, r);
        }
        __syncthreads();
        {
            R r[32];
            lds_ld<32>(H, yz * SPd + xh, RPd, r);
            row<32, false, WT, R>(r);
            lds_st<32>(H, yz * SPd + xh, RPd, r);
        }
        __syncthreads();
        // level-32 x-pass on whole rows in registers, then the store
        const BlockRef<R> br = locate<table_layout<LAYOUT>(), R>(T, one_field<LAYOUT>() ? 1 : nf, b, nullptr, dst, g, 0.0f);
#pragma unroll
        for (int k = 0; k < NROW; ++k) {
            const int q = dec_xrow<RL>(t, k);
            R r[32];
#pragma unroll
            for (int i = 0; i < 16; ++i) r[i] = lo[k][i];
            lds_ld<16>(H, xrow_half_base<RL, RPd, SPd>(q), 1, &r[16]);
            row<32, false, WT, R>(r);
            store_xrow_h<R, RL>(br.dst, br.g, br.lb, q, r);
        }
    }
}

// ---------------------------------------------------------------------------
// host side

template <typename R, int LAYOUT>
static int launch_compress_h(int wt, const R *src, FieldGeom g, int nblocks, float eps, uint32_t zkeep,
                             uint8_t *slots, uint32_t *nb, uint32_t *cnt, const czw_field_t *fields, int nf,
                             int groups, uint32_t *queue, hipStream_t s)
{
    const void *fn = wt == 1 ? (const void *)k_compress_h<R, 1, LAYOUT>
                   : wt == 2 ? (const void *)k_compress_h<R, 2, LAYOUT> : (const void *)k_compress_h<R, 3, LAYOUT>;
    // never more than are resident: a second persistent wave would start only
    // after the first had drained the queue
    const int res = resident_groups(fn, NT);
    const dim3 grid(std::min(nblocks, groups ? std::min(groups, res) : res)), blk(NT);
    if (wt == 1) hipLaunchKernelGGL((k_compress_h<R, 1, LAYOUT>), grid, blk, 0, s, src, g, nblocks, eps, zkeep, slots, nb, cnt, fields, nf, queue);
    if (wt == 2) hipLaunchKernelGGL((k_compress_h<R, 2, LAYOUT>), grid, blk, 0, s, src, g, nblocks, eps, zkeep, slots, nb, cnt, fields, nf, queue);
    if (wt == 3) hipLaunchKernelGGL((k_compress_h<R, 3, LAYOUT>), grid, blk, 0, s, src, g, nblocks, eps, zkeep, slots, nb, cnt, fields, nf, queue);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

template <typename R>
static int compress_slots_t(int layout, const void *src, int nblocks, FieldGeom g, float eps, int wt,
                            uint32_t zkeep, uint8_t *slots, uint32_t *nbytes, const czw_field_t *fields, int nf,
                            int groups, uint32_t *queue, hipStream_t s)
{
    uint32_t *cnt = nbytes + nblocks;
    const R *d = static_cast<const R *>(src);
    switch (layout) {
    case LAYOUT_BLOCKS:
        return launch_compress_h<R, LAYOUT_BLOCKS>(wt, d, g, nblocks, eps, zkeep, slots, nbytes, cnt, nullptr, 0,
                                                   groups, queue, s);
    case LAYOUT_FIELD:
        return launch_compress_h<R, LAYOUT_FIELD>(wt, d, g, nblocks, eps, zkeep, slots, nbytes, cnt, nullptr, 0,
                                                  groups, queue, s);
    case LAYOUT_FIELD_WIDE:
        return launch_compress_h<R, LAYOUT_FIELD_WIDE>(wt, d, g, nblocks, eps, zkeep, slots, nbytes, cnt, nullptr,
                                                       0, groups, queue, s);
    default:
        return launch_compress_h<R, LAYOUT_BATCH>(wt, d, g, nblocks, eps, zkeep, slots, nbytes, cnt, fields, nf,
                                                  groups, queue, s);
    }
}

// The binary32 compressor is k_compress_f32 (czw_kernels.hip); the half-block
// compressor and its packer serve binary64, and binary32 batches as the
// second tile a batch's launch shape can name.
int compress_slots_f64(int layout, const void *src, int nblocks, FieldGeom g, float eps, int wtype, int zerobits,
                       uint8_t *slots, uint32_t *nbytes, const czw_field_t *fields, int nf, int groups,
                       uint32_t *queue, hipStream_t s)
{
    const int wt = wt_index(wtype);
    if (wt < 0 || !zb_ok(zerobits) || nblocks <= 0 || !slots || !nbytes) return CZW_EINVAL;
    if (layout != LAYOUT_BATCH && !src) return CZW_EINVAL;
    const uint32_t zkeep = zerobits ? ~((1u << zerobits) - 1u) : 0xffffffffu;
    return compress_slots_t<double>(layout, src, nblocks, g, eps, wt, zkeep, slots, nbytes, fields, nf, groups,
                                    queue, s);
}

int pack_records_f64(const uint8_t *slots, const uint32_t *nbytes, int nblocks, int shuffle, uint8_t *records,
                     uint64_t *offsets, hipStream_t s)
{
    if (nblocks <= 0 || !slots || !nbytes || !records || !offsets) return CZW_EINVAL;
    if (scan_records(nbytes, nblocks, offsets, s) != CZW_OK) return CZW_EHIP;
    hipLaunchKernelGGL(k_pack_records_h<double>, dim3(nblocks), dim3(256), 0, s, slots, nbytes, nbytes + nblocks,
                       shuffle, offsets, records);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

template <typename R, int LAYOUT, bool SHUF>
static int launch_decompress_hs(int wt, const uint8_t *rec, uint64_t rb, const uint64_t *offsets, int nblocks, R *dst,
                                FieldGeom g, uint32_t *err, const czw_field_t *fields, int nf, int groups,
                                uint32_t *queue, hipStream_t s)
{
    const void *fn = wt == 1 ? (const void *)k_decompress_h<R, 1, LAYOUT, SHUF>
                   : wt == 2 ? (const void *)k_decompress_h<R, 2, LAYOUT, SHUF>
                             : (const void *)k_decompress_h<R, 3, LAYOUT, SHUF>;
    const int res = resident_groups(fn, NT);
    const dim3 grid(std::min(nblocks, groups ? std::min(groups, res) : res)), blk(NT);
    if (wt == 1) hipLaunchKernelGGL((k_decompress_h<R, 1, LAYOUT, SHUF>), grid, blk, 0, s, rec, rb, offsets, nblocks, dst, g, err, fields, nf, queue);
    if (wt == 2) hipLaunchKernelGGL((k_decompress_h<R, 2, LAYOUT, SHUF>), grid, blk, 0, s, rec, rb, offsets, nblocks, dst, g, err, fields, nf, queue);
    if (wt == 3) hipLaunchKernelGGL((k_decompress_h<R, 3, LAYOUT, SHUF>), grid, blk, 0, s, rec, rb, offsets, nblocks, dst, g, err, fields, nf, queue);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

template <typename R, int LAYOUT>
static int launch_decompress_h(int wt, const uint8_t *rec, uint64_t rb, const uint64_t *offsets, int nblocks,
                               int shuffle, R *dst, FieldGeom g, uint32_t *err, const czw_field_t *fields, int nf,
                               int groups, uint32_t *queue, hipStream_t s)
{
    if (shuffle)
        return launch_decompress_hs<R, LAYOUT, true>(wt, rec, rb, offsets, nblocks, dst, g, err, fields, nf, groups,
                                                     queue, s);
    return launch_decompress_hs<R, LAYOUT, false>(wt, rec, rb, offsets, nblocks, dst, g, err, fields, nf, groups, queue,
                                                  s);
}

template <typename R>
static int decompress_t(int layout, const uint8_t *rec, uint64_t rb, const uint64_t *offsets, int nblocks,
                        FieldGeom g, int wt, int shuffle, void *dst, uint32_t *err, const czw_field_t *fields, int nf,
                        int groups, uint32_t *queue, hipStream_t s)
{
    R *d = static_cast<R *>(dst);
    switch (layout) {
    case LAYOUT_BLOCKS:
        return launch_decompress_h<R, LAYOUT_BLOCKS>(wt, rec, rb, offsets, nblocks, shuffle, d, g, err, nullptr, 0,
                                                     groups, queue, s);
    case LAYOUT_FIELD:
        return launch_decompress_h<R, LAYOUT_FIELD>(wt, rec, rb, offsets, nblocks, shuffle, d, g, err, nullptr, 0,
                                                    groups, queue, s);
    case LAYOUT_FIELD_WIDE:
        return launch_decompress_h<R, LAYOUT_FIELD_WIDE>(wt, rec, rb, offsets, nblocks, shuffle, d, g, err, nullptr,
                                                         0, groups, queue, s);
    default:
        return launch_decompress_h<R, LAYOUT_BATCH>(wt, rec, rb, offsets, nblocks, shuffle, d, g, err, fields, nf,
                                                    groups, queue, s);
    }
}

int decompress_half(int prec, int layout, const void *records, size_t rec_bytes, const uint64_t *offsets,
                    int nblocks, FieldGeom g, int wtype, int shuffle, void *dst, uint32_t *err,
                    const czw_field_t *fields, int nf, int groups, uint32_t *queue, hipStream_t s)
{
    const int wt = wt_index(wtype);
    if (wt < 0 || nblocks <= 0 || !records || !offsets || !err) return CZW_EINVAL;
    if (layout != LAYOUT_BATCH && !dst) return CZW_EINVAL;
    const uint8_t *rec = static_cast<const uint8_t *>(records);
    if (prec == 8)
        return decompress_t<double>(layout, rec, (uint64_t)rec_bytes, offsets, nblocks, g, wt, shuffle, dst, err,
                                    fields, nf, groups, queue, s);
    if (prec == 4)
        return decompress_t<float>(layout, rec, (uint64_t)rec_bytes, offsets, nblocks, g, wt, shuffle, dst, err,
                                   fields, nf, groups, queue, s);
    return CZW_EINVAL;
}


// Binary32 batches with the half-block tile: slots in the same region layout
// as binary64 (nbytes holds 9 words per block), framed by k_pack_records_h.
int compress_slots_h32_batch(const czw_field_t *fields, int nf, int nblocks, int wtype, int zerobits, uint8_t *slots,
                             uint32_t *nbytes, int groups, uint32_t *queue, hipStream_t s)
{
    const int wt = wt_index(wtype);
    if (wt < 0 || !zb_ok(zerobits) || nblocks <= 0 || !slots || !nbytes || !fields) return CZW_EINVAL;
    const uint32_t zkeep = zerobits ? ~((1u << zerobits) - 1u) : 0xffffffffu;
    return launch_compress_h<float, LAYOUT_BATCH>(wt, nullptr, FieldGeom{0, 0, 0}, nblocks, 0.0f, zkeep, slots, nbytes,
                                                  nbytes + nblocks, fields, nf, groups, queue, s);
}

int pack_records_h32(const uint8_t *slots, const uint32_t *nbytes, int nblocks, int shuffle, uint8_t *records,
                     uint64_t *offsets, hipStream_t s)
{
    if (nblocks <= 0 || !slots || !nbytes || !records || !offsets) return CZW_EINVAL;
    if (scan_records(nbytes, nblocks, offsets, s) != CZW_OK) return CZW_EHIP;
    hipLaunchKernelGGL(k_pack_records_h<float>, dim3(nblocks), dim3(256), 0, s, slots, nbytes, nbytes + nblocks,
                       shuffle, offsets, records);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

int resident_compress_h32(int wtype)
{
    const int wt = wt_index(wtype);
    const void *fn = wt == 1 ? (const void *)k_compress_h<float, 1, LAYOUT_BATCH>
                   : wt == 2 ? (const void *)k_compress_h<float, 2, LAYOUT_BATCH>
                             : (const void *)k_compress_h<float, 3, LAYOUT_BATCH>;
    return resident_groups(fn, NT);
}

// Resident workgroups of the batch-layout kernels a launch would use.
int resident_compress_f64(int layout, int wtype)
{
    (void)layout;  // the batch table is the only layout asked about
    const int wt = wt_index(wtype);
    const void *fn = wt == 1 ? (const void *)k_compress_h<double, 1, LAYOUT_BATCH>
                   : wt == 2 ? (const void *)k_compress_h<double, 2, LAYOUT_BATCH>
                             : (const void *)k_compress_h<double, 3, LAYOUT_BATCH>;
    return resident_groups(fn, NT);
}

template <typename R, bool SHUF>
static int resident_dec(int wt)
{
    const void *fn = wt == 1 ? (const void *)k_decompress_h<R, 1, LAYOUT_BATCH, SHUF>
                   : wt == 2 ? (const void *)k_decompress_h<R, 2, LAYOUT_BATCH, SHUF>
                             : (const void *)k_decompress_h<R, 3, LAYOUT_BATCH, SHUF>;
    return resident_groups(fn, NT);
}

int resident_decompress(int prec, int layout, int wtype, int shuffle)
{
    (void)layout;
    const int wt = wt_index(wtype);
    if (prec == 8) return shuffle ? resident_dec<double, true>(wt) : resident_dec<double, false>(wt);
    return shuffle ? resident_dec<float, true>(wt) : resident_dec<float, false>(wt);
}

}  // namespace czw
