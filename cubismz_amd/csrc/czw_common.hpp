// czw_common.hpp -- pieces shared by the binary32 and binary64 block
// kernels: sizes, the batch field table, field geometry, LDS row helpers,
// block-wide scans, the [0,16)^3 corner passes and stream pieces, and host
// launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/czw.h"
#include "czw_wavelet.hpp"

namespace czw {

constexpr int NB3 = 32768;
constexpr int MASK_BYTES = 4096;
constexpr int NT = 512;            // threads per block workgroup (8 wave64)
constexpr int NROW = 1024 / NT;    // level-32 rows (and stream segments) per thread

// LAYOUT_FIELD addresses a block's rows with buffer instructions: 32-bit
// byte offsets from the block origin, the x step in a scalar register (no
// per-element address arithmetic).  Fields whose 32 x-planes exceed that
// range use LAYOUT_FIELD_WIDE (64-bit addresses).
enum { LAYOUT_BLOCKS = 0, LAYOUT_FIELD = 1, LAYOUT_BATCH = 2, LAYOUT_FIELD_WIDE = 3 };
constexpr int BATCH_MAX_FIELDS = 64;  // field table entries one launch keeps in LDS

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

// Wave64 inclusive prefix sum in DPP steps (no LDS round trips): row shifts by
// 1, 2, 4 and 8 scan each 16-lane row (lanes shifted in from outside the row
// add 0), then row_bcast:15 adds row 0's total to row 1 and row 2's to row 3,
// and row_bcast:31 adds lane 31's prefix to rows 2 and 3.
CZW_DI uint32_t wave_incl_scan(uint32_t x, int lane)
{
    (void)lane;
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1, 3
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2, 3
    return x;
}

// OR of v over the wave, returned in a scalar register: row shifts by 1, 2,
// 4 and 8 leave each 16-lane row's OR in its last lane.
CZW_DI uint32_t wave_or(uint32_t v)
{
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
    return (uint32_t)(__builtin_amdgcn_readlane((int)v, 15) | __builtin_amdgcn_readlane((int)v, 31) |
                      __builtin_amdgcn_readlane((int)v, 47) | __builtin_amdgcn_readlane((int)v, 63));
}

// Block-wide exclusive scan of one value per thread.  Returns the exclusive
// prefix; *total receives the sum.  Ends with the wave totals read, so the
// caller must barrier before wsum is written again.
template <int NTH = NT>
CZW_DI uint32_t block_excl_scan(uint32_t v, uint32_t *wsum, uint32_t *total)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t inc = wave_incl_scan(v, lane);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NTH / 64; ++i) {
        const uint32_t s = wsum[i];
        base += (i < w) ? s : 0u;
        tot += s;
    }
    *total = tot;
    return base + inc - v;
}

// An HDF5-ordered field [nx][ny][nz]; a launch's block b is the field's
// block b0 + b (b0 > 0 when a launch covers a range of the field's blocks).
struct FieldGeom {
    int nx, ny, nz;
    int b0 = 0;
};

// Global element offset of block b's row start (x = 0) for the x-row (y, z)
// of an HDF5-ordered field, and the stride along x.
CZW_DI size_t field_row(const FieldGeom &g, int b, int y, int z, size_t *xstride)
{
    b += g.b0;
    const int bx = g.nx >> 5, by = g.ny >> 5;
    const int ix = b % bx, iy = (b / bx) % by, iz = b / (bx * by);
    *xstride = (size_t)g.ny * (size_t)g.nz;
    return ((size_t)(32 * ix) * g.ny + (size_t)(32 * iy + y)) * g.nz + (size_t)(32 * iz + z);
}

// A wave-uniform pointer as scalar registers.
template <typename T>
CZW_DI T *uniform_ptr(T *p)
{
    const unsigned long long u = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return reinterpret_cast<T *>(((unsigned long long)hi << 32) | lo);
}

// The block a workgroup works on, resolved from the batch-wide index b: the
// field it belongs to (source or destination, geometry, threshold) and its
// index lb inside that field.  Outside LAYOUT_BATCH this is the launch's one
// field and lb = b.
template <typename R>
struct BlockRef {
    const R *src;
    R *dst;
    FieldGeom g;
    int lb;
    float eps;
};

// Field table of a batch launch, copied into LDS once per workgroup.
template <int LAYOUT>
CZW_DI void load_table(czw_field_t *T, const czw_field_t *__restrict__ fields, int nf)
{
    if constexpr (LAYOUT == LAYOUT_BATCH) {
        constexpr int W = sizeof(czw_field_t) / 4;
        const uint32_t *s = reinterpret_cast<const uint32_t *>(fields);
        uint32_t *d = reinterpret_cast<uint32_t *>(T);
        for (int i = threadIdx.x; i < nf * W; i += blockDim.x) d[i] = s[i];
        __syncthreads();
    }
}

template <int LAYOUT, typename R>
CZW_DI BlockRef<R> locate(const czw_field_t *T, int nf, int b, const R *src, R *dst, FieldGeom g, float eps)
{
    if constexpr (LAYOUT != LAYOUT_BATCH) {
        return BlockRef<R>{src, dst, g, b, eps};
    } else {
        int lo = 0, hi = nf - 1;  // the last field whose first block is <= b
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (T[mid].block0 <= b) lo = mid;
            else hi = mid - 1;
        }
        // the table entry is the same for the whole workgroup: scalar registers,
        // so that the row addresses and the buffer offsets built from them are too
        const czw_field_t &f = T[lo];
        auto u32 = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
        const float e = __int_as_float(u32(__float_as_int(f.eps)));
        return BlockRef<R>{uniform_ptr(static_cast<const R *>(f.src)), uniform_ptr(static_cast<R *>(f.dst)),
                           FieldGeom{u32(f.nx), u32(f.ny), u32(f.nz)}, u32(b - f.block0), e};
    }
}

// A single HDF5-ordered field goes through a one-entry table as well: its
// geometry then comes from LDS into scalar registers when a block starts,
// rather than occupying scalar registers through the passes (which spilled
// them).  Block b of the launch is block b + g.b0 of the field.
template <int LAYOUT>
constexpr bool one_field()
{
    return LAYOUT == LAYOUT_FIELD || LAYOUT == LAYOUT_FIELD_WIDE;
}
template <int LAYOUT>
constexpr int table_layout()
{
    return one_field<LAYOUT>() ? LAYOUT_BATCH : LAYOUT;
}
template <int LAYOUT>
CZW_DI void load_one(czw_field_t *T, const void *src, void *dst, const FieldGeom &g, float eps)
{
    if constexpr (one_field<LAYOUT>()) {
        if (threadIdx.x == 0) T[0] = czw_field_t{src, dst, g.nx, g.ny, g.nz, eps, -g.b0, 0};
        __syncthreads();
    }
}

// Load/store layout of a block's rows: batch fields are HDF5-ordered fields.
template <int LAYOUT>
constexpr int row_layout()
{
    return LAYOUT == LAYOUT_BLOCKS ? LAYOUT_BLOCKS : (LAYOUT == LAYOUT_FIELD_WIDE ? LAYOUT_FIELD_WIDE : LAYOUT_FIELD);
}

// Cache policy of the field's loads (compressors) and stores (decoders):
// non-temporal.  The field is streamed once; keeping its 512 MiB out of the
// caches leaves the slots and the records that the next stage reads there
// (the whole 512^3 compress + pack + decompress step: -15%).
constexpr int CZW_FIELD_CPOL = 2;  // nt on gfx950 buffer instructions

// Whether every offset inside a block of an [nx][ny][nz] field with Real of
// prec bytes fits the buffer range used by LAYOUT_FIELD.
inline bool field_fits_buffer(int ny, int nz, int prec)
{
    return 32ull * (unsigned long long)ny * (unsigned long long)nz * (unsigned long long)prec < (1ull << 31);
}

// Raw buffer resource over [base, base + 2 GiB) (stride 0, no swizzle).
CZW_DI __amdgpu_buffer_rsrc_t buffer_rsrc(const void *base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(uniform_ptr(base)), 0, 0x7fffffff, 0x00020000);
}

typedef unsigned czw_u32x2 __attribute__((ext_vector_type(2)));
// element vectors for __builtin_nontemporal_load / _store
typedef float czw_f32x2 __attribute__((ext_vector_type(2)));
typedef float czw_f32x4 __attribute__((ext_vector_type(4)));
typedef double czw_f64x2 __attribute__((ext_vector_type(2)));
typedef unsigned czw_u32x4 __attribute__((ext_vector_type(4)));

// Level-16 y-pass row (z, x) of thread t < 256 on the [0,16)^3 corner: the
// first wave takes exactly the 64 rows z < 8, x < 8 that level 8 also
// transforms, so the level-8 passes stay inside that wave; the other 192
// rows go to threads 64..255.
CZW_DI void corner_yrow(int t, int &z, int &x)
{
    if (t < 64) {
        z = t >> 3;
        x = t & 7;
    } else if (t < 128) {
        z = (t - 64) >> 3;
        x = 8 + (t & 7);
    } else {
        z = 8 + ((t - 128) >> 4);
        x = t & 15;
    }
}

// Ordering point inside one wave between LDS passes whose rows cross lanes:
// the LDS executes a wave's instructions in issue order, so only the
// compiler must be kept from moving accesses across it.
CZW_DI void wave_lds_order()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

// Remaining forward passes on the [0,16)^3 corner of a block held in LDS at
// z*SP_ + y*RP_ + x: level 16 (x, y) and level 8 (y fused, z, x).  The
// level-8 passes run on the first wave alone, behind its level-16 y-rows,
// with no workgroup barrier between them.
template <int WT, int RP_, int SP_, typename R>
CZW_DI void fwt16_8(R *A, int t)
{
    R r[16];
    if (t < 256) {  // level 16 x-pass: row (z, y) in [0,16)^2
        const int z = t >> 4, y = t & 15;
        lds_ld<16>(A, z * SP_ + y * RP_, 1, r);
        row<16, true, WT, R>(r);
        lds_st<16>(A, z * SP_ + y * RP_, 1, r);
    }
    __syncthreads();
    if (t < 256) {  // level 16 y-pass, then level 8 y-pass: row (z, x)
        int z, x;
        corner_yrow(t, z, x);
        lds_ld<16>(A, z * SP_ + x, RP_, r);
        row<16, true, WT, R>(r);
        if (t < 64) row<8, true, WT, R>(r);
        lds_st<16>(A, z * SP_ + x, RP_, r);
    }
    if (t < 64) {
        wave_lds_order();
        {  // level 8 z-pass: row (y, x) in [0,8)^2
            const int y = t >> 3, x = t & 7;
            lds_ld<8>(A, y * RP_ + x, SP_, r);
            row<8, true, WT, R>(r);
            lds_st<8>(A, y * RP_ + x, SP_, r);
        }
        wave_lds_order();
        {  // level 8 x-pass: row (z, y)
            const int z = t >> 3, y = t & 7;
            lds_ld<8>(A, z * SP_ + y * RP_, 1, r);
            row<8, true, WT, R>(r);
            lds_st<8>(A, z * SP_ + y * RP_, 1, r);
        }
    }
    __syncthreads();
}

// Inverse passes of all levels except the final level-32 x-pass, which the
// caller fuses with the store: level 8 (x, z, y), level 16 (y fused, x, z),
// level 32 (z fused, y).  The first wave runs the level-8 x- and z-passes and
// then its level-16 y-rows; the other y-rows lie outside [0,8)^3 and start at
// once.
template <int WT, int RP_, int SP_, typename R>
CZW_DI void iwt16_8(R *A, int t)
{
    {
        R r[16];
        if (t < 64) {
            {  // level 8 x-pass
                const int z = t >> 3, y = t & 7;
                lds_ld<8>(A, z * SP_ + y * RP_, 1, r);
                row<8, false, WT, R>(r);
                lds_st<8>(A, z * SP_ + y * RP_, 1, r);
            }
            wave_lds_order();
            {  // level 8 z-pass
                const int y = t >> 3, x = t & 7;
                lds_ld<8>(A, y * RP_ + x, SP_, r);
                row<8, false, WT, R>(r);
                lds_st<8>(A, y * RP_ + x, SP_, r);
            }
            wave_lds_order();
        }
        if (t < 256) {  // level 8 y-pass, then level 16 y-pass
            int z, x;
            corner_yrow(t, z, x);
            lds_ld<16>(A, z * SP_ + x, RP_, r);
            if (t < 64) row<8, false, WT, R>(r);
            row<16, false, WT, R>(r);
            lds_st<16>(A, z * SP_ + x, RP_, r);
        }
        __syncthreads();
        if (t < 256) {  // level 16 x-pass: lanes step through y, or through z
                        // where rows are 16 banks apart (RP_ = 16 mod 32)
            const int z = (RP_ % 32 == 16) ? (t & 15) : (t >> 4), y = (RP_ % 32 == 16) ? (t >> 4) : (t & 15);
            lds_ld<16>(A, z * SP_ + y * RP_, 1, r);
            row<16, false, WT, R>(r);
            lds_st<16>(A, z * SP_ + y * RP_, 1, r);
        }
        __syncthreads();
    }
}

// Stream positions [8p, 8p+8) of the [0,16)^3 corner, p < 512 (see
// for_each_run): two runs of 4 for p < 64 (base 4^3 and level-8 details),
// one run of 8 otherwise (level-16 details).
template <int RP_, int SP_, typename F>
CZW_DI void corner_piece(int p, F &&f)
{
    if (p < 64) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int s = 8 * p + 4 * q;
            const int code = s >> 6, k = s & 63;
            const int a0 = 4 * ((code >> 2) & 1) + (k >> 4);
            const int a1 = 4 * ((code >> 1) & 1) + ((k >> 2) & 3);
            const int o2 = 4 * (code & 1);
            f(q, 4, a0 * SP_ + a1 * RP_ + o2, 1, a1 + 32 * a0, o2);
        }
    } else {
        const int s = 8 * p;
        const int code = s >> 9, k = s & 511;
        const int a0 = 8 * ((code >> 2) & 1) + (k >> 6);
        const int a1 = 8 * ((code >> 1) & 1) + ((k >> 3) & 7);
        const int o2 = 8 * (code & 1);
        f(0, 8, a1 * SP_ + o2 * RP_ + a0, RP_, a1 + 32 * a0, o2);
    }
}

// ---------------------------------------------------------------------------
// host helpers

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
// given block size (the persistent kernels use one ~140 KB LDS block per CU).
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

static bool field_ok(int nx, int ny, int nz)
{
    return nx > 0 && ny > 0 && nz > 0 && !(nx & 31) && !(ny & 31) && !(nz & 31) &&
           (size_t)(nx / 32) * (ny / 32) * (nz / 32) < (size_t)INT32_MAX;
}

// Exclusive scan of record sizes 4 + nbytes[b] into offsets[0..n] (one
// workgroup); defined with the binary32 kernels.
int scan_records(const uint32_t *nbytes, int n, uint64_t *offsets, hipStream_t s);

// Per-block size words after the slots of a compressor workspace: the payload
// size, then the region counts the packer needs (binary32: the corner count;
// binary64: the corner and level-32 code counts).
inline size_t size_words(int prec) { return prec == 8 ? 9 : 2; }

// Half-block kernels (czw_half.hip) for Real of prec bytes; layout is
// LAYOUT_BLOCKS, LAYOUT_FIELD(_WIDE) or LAYOUT_BATCH (then src/dst come from
// the field table); groups = 0 launches every resident workgroup; queue =
// nullptr keeps the static block order.  The compressor and packer here
// serve binary64 (binary32 compresses in czw_kernels.hip, except batches
// launched with the half-block tile); the decoder takes both precisions.
int decompress_half(int prec, int layout, const void *records, size_t rec_bytes, const uint64_t *offsets,
                    int nblocks, FieldGeom g, int wtype, int shuffle, void *dst, uint32_t *err,
                    const czw_field_t *fields, int nf, int groups, uint32_t *queue, hipStream_t s);
int compress_slots_f64(int layout, const void *src, int nblocks, FieldGeom g, float eps, int wtype, int zerobits,
                       uint8_t *slots, uint32_t *nbytes, const czw_field_t *fields, int nf, int groups,
                       uint32_t *queue, hipStream_t s);
int pack_records_f64(const uint8_t *slots, const uint32_t *nbytes, int nblocks, int shuffle, uint8_t *records,
                     uint64_t *offsets, hipStream_t s);
// Binary32 batches with the half-block tile (nbytes: 9 words per block).
int compress_slots_h32_batch(const czw_field_t *fields, int nf, int nblocks, int wtype, int zerobits, uint8_t *slots,
                             uint32_t *nbytes, int groups, uint32_t *queue, hipStream_t s);
int pack_records_h32(const uint8_t *slots, const uint32_t *nbytes, int nblocks, int shuffle, uint8_t *records,
                     uint64_t *offsets, hipStream_t s);
// Resident workgroups (device-wide) of the batch-layout half-block kernels.
int resident_compress_f64(int layout, int wtype);
int resident_compress_h32(int wtype);
int resident_decompress(int prec, int layout, int wtype, int shuffle);

}  // namespace czw
