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
#include <atomic>
#include <mutex>
#include <type_traits>

#include "czw_common.hpp"

namespace czw {

constexpr int RP = 33;
constexpr int SP = 32 * 33 + 1;
constexpr int LDS_WORDS = 32 * SP;
constexpr size_t SLOT32 = MASK_BYTES + 4 * (size_t)NB3;
// The compressor's workgroup: NTH threads, CR = 1024/NTH level-32 rows per
// thread in every pass.  512 threads (8 wave64, two per SIMD) hold two rows
// each in 256 VGPRs; 1024 threads (four waves per SIMD) one row each in 128.
template <int NTH>
struct CTile {
    static_assert(NTH == 512 || NTH == 1024, "512 or 1024 threads");
    static constexpr int CR = 1024 / NTH;
};
constexpr int CT = 512;  // the default tile

// ---------------------------------------------------------------------------
// The x-rows of thread t.  Block layout: rows q = t + NTH*k, (z, y) =
// (q >> 5, q & 31), 16-byte loads along x.  HDF5 field layouts, two rows per
// thread: (y, z0) and (y, z0 + 1), y = t >> 4, z0 = 2*(t & 15), so that one
// 8-byte load reads both rows' values at one x; one row per thread: (y, z) =
// (t >> 5, t & 31).  Either way each load instruction of a wave reads whole
// 128-byte runs of consecutive Z.
template <int LAYOUT, int NTH>
CZW_DI int xrow_of(int t, int k)
{
    if constexpr (LAYOUT == LAYOUT_BLOCKS || NTH == 1024) return t + NTH * k;
    else return ((t >> 4) << 5) | (2 * (t & 15) + k);
}

template <int LAYOUT>
CZW_DI int xrow_lds_base(int q)
{
    if constexpr (LAYOUT == LAYOUT_BLOCKS) return (q >> 5) * SP + (q & 31) * RP;
    else return (q & 31) * SP + (q >> 5) * RP;
}

// x in [X0, X1) of thread t's x-rows into r[k][X0 .. X1).
template <int LAYOUT, int NTH, int X0, int X1>
CZW_DI void load_xrows_f32(const float *__restrict__ src, const FieldGeom &g, int b, int t,
                           float (&r)[CTile<NTH>::CR][32])
{
    constexpr int CR = CTile<NTH>::CR;
    if constexpr (LAYOUT == LAYOUT_BLOCKS) {
#pragma unroll
        for (int k = 0; k < CR; ++k) {
            const czw_f32x4 *p = reinterpret_cast<const czw_f32x4 *>(src + (size_t)b * NB3 + (size_t)(t + NTH * k) * 32);
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
        const int so = (int)xs * 4;
        if constexpr (CR == 2) {
            const int vo = ((t >> 4) * g.nz + 2 * (t & 15)) * 4;
#pragma unroll
            for (int x = X0; x < X1; ++x) {
                const czw_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, vo, x * so, CZW_FIELD_CPOL);
                r[0][x] = __uint_as_float(v.x);
                r[1][x] = __uint_as_float(v.y);
            }
        } else {
            const int vo = ((t >> 5) * g.nz + (t & 31)) * 4;
#pragma unroll
            for (int x = X0; x < X1; ++x)
                r[0][x] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, vo, x * so, CZW_FIELD_CPOL));
        }
    } else {
        size_t xs;
        if constexpr (CR == 2) {
            const float *p = src + field_row(g, b, t >> 4, 2 * (t & 15), &xs);
#pragma unroll
            for (int x = X0; x < X1; ++x) {
                const czw_f32x2 v = __builtin_nontemporal_load(reinterpret_cast<const czw_f32x2 *>(p + x * xs));
                r[0][x] = v.x;
                r[1][x] = v.y;
            }
        } else {
            const float *p = src + field_row(g, b, t >> 5, t & 31, &xs);
#pragma unroll
            for (int x = X0; x < X1; ++x) r[0][x] = __builtin_nontemporal_load(p + x * xs);
        }
    }
}

// ---------------------------------------------------------------------------
// Record offsets by decoupled look-back.  Block b's status word (zero before
// the launch) becomes ST_A | record bytes as soon as its size is known, and
// ST_P | (offset + record bytes), its inclusive prefix, once its offset is
// known; ST_E marks a block that gave up waiting.  The words are read and
// written with device-scope atomics, which gfx950 serves coherently across
// its XCDs' L2 caches.
constexpr uint64_t ST_A = 1ull << 62, ST_P = 2ull << 62, ST_E = 3ull << 62, ST_VAL = (1ull << 62) - 1;

// The stores are device-scope atomics (global-memory instructions: a flat
// access would also hold up the LDS waits).  The reads are plain buffer loads
// with the device-coherent cache policy (sc1, as the device-scope atomic load
// is), so that the compiler need not wait for each before issuing the next
// (a wait on vector memory also waits for every load issued before, here the
// next block's prefetch); a re-read is kept apart from the read before it by
// a compiler barrier.
typedef __attribute__((address_space(1))) uint64_t g_u64;
constexpr int ST_CPOL = 16;  // sc1
CZW_DI void st_store(uint64_t *p, uint64_t v) { __hip_atomic_store((g_u64 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Record offset of block b (the sum of its predecessors' record bytes), found
// by one wave from its predecessors' status words, nearest first, down to the
// nearest inclusive prefix.  A read covers LB_WIN windows of 64 predecessors
// (one memory round trip); s0 holds the first read, issued by the caller
// ahead of time (look_back_issue).  A window whose nearest prefix lies behind
// a not-ready word is read again, at most `limit` times; returns false then,
// or when a predecessor gave up.
constexpr int LB_WIN = 4;

struct LookBack {
    uint64_t s[LB_WIN];
};

CZW_DI void look_back_issue(const uint64_t *status, int j, LookBack &lb)
{
    const int lane = threadIdx.x & 63;
    const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(status);
    asm volatile("" ::: "memory");
#pragma unroll
    for (int w = 0; w < LB_WIN; ++w) {  // unconditional loads (word 0 exists), then a select
        const int i = j - lane - 64 * w;
        const czw_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, (i >= 0 ? i : 0) * 8, 0, ST_CPOL);
        lb.s[w] = i >= 0 ? (((uint64_t)v.y << 32) | v.x) : ST_P;  // before block 0: prefix 0
    }
}

CZW_DI bool look_back(const uint64_t *status, int b, uint32_t limit, LookBack &lb, uint64_t *prefix)
{
    const int lane = threadIdx.x & 63;
    uint64_t acc = 0;
    uint32_t spins = 0;
    int j = b - 1;  // nearest predecessor not yet summed
    while (j >= 0) {
        bool retry = false;
#pragma unroll
        for (int w = 0; w < LB_WIN; ++w) {
            const uint64_t s = lb.s[w];
            const uint32_t flag = (uint32_t)(s >> 62);
            const uint64_t done = __ballot(flag >= 2u);
            const uint64_t wait = __ballot(flag == 0u);
            const int lp = done ? __builtin_ctzll(done) : 64;
            const uint64_t upto = lp >= 63 ? ~0ull : (2ull << lp) - 1ull;  // lanes 0 .. lp
            if (wait & upto) {
                retry = true;
                break;
            }
            if (lp < 64 && ((__ballot(flag == 3u) >> lp) & 1ull)) return false;
            uint64_t v = lane <= lp ? (s & ST_VAL) : 0ull;
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
            acc += v;
            if (lp < 64) {
                *prefix = acc;
                return true;
            }
            j -= 64;
        }
        if (retry) {
            if (++spins > limit) return false;
            __builtin_amdgcn_s_sleep(2);
        }
        if (j >= 0) look_back_issue(status, j, lb);
    }
    *prefix = acc;
    return true;
}

// ---------------------------------------------------------------------------
// Compressor, binary32.  Mirrors WaveletCompressorGeneric::compress
// (WaveletCompressor.cpp:153-182): fwt, threshold + mask, zerobits, shuffle,
// and, with REC, the record framing of the writer
// (SerializerIO_WaveletCompression_MPI_Simple.h:237-240).
//
// A persistent grid of 1024-thread workgroups walks the blocks; a workgroup
// owns one 32^3 block in LDS at a time while the next block's x-rows are
// loaded into registers, so the HBM reads of one block overlap the
// arithmetic of the previous one.  A block's output is staged in LDS (mask in
// M; survivors contiguous in stream order in A from STAGE32_AT - Sc) and
// written at the start of the workgroup's next block, once its record
// offset is known: the look-back runs on the last wave while the other waves
// compute the next block's level-32 x-pass.
//
// REC: blocks come from a work queue (hdr[0], zero at launch), so a block is
// only ever waited for by blocks issued after it to running workgroups; the
// output is the record stream with offsets[nblocks + 1].  Slot mode (!REC):
// static block order, block b's payload at slot b in the region layout read
// by k_pack_records, its sizes in nbytes[b] and nbytes[nblocks + b].
constexpr int CORNER_MAX = 4096;
constexpr int NRUN32 = 1792;  // level-32 stream runs: 7 codes x 16 x 16 runs of 16
constexpr int STAGE32_AT = 16 * SP;  // the z >= 16 half of A
constexpr int STAGE32 = ((LDS_WORDS - STAGE32_AT - 64) / 4) * 4;  // staged survivors (64 dummy words follow)
constexpr int OVF_WORDS = 7 * 4096;  // level-32 survivors of a block at most
// REC: a block's mask and staged survivors are parked in an LDS buffer of
// this many words at the start of the next block, and written one block
// later still, once the next block has found its record offset in the
// background (the LDS the slot path leaves free)
constexpr int SB_WORDS = 3904;
constexpr int MAX_GROUPS = 256;      // compressor workgroups per launch at most (overflow areas)
static_assert(STAGE32_AT >= CORNER_MAX, "corner survivors are staged below the level-32 ones");

struct CompressOut {
    uint8_t *dst;       // REC: the record stream; slot mode: the slots
    uint64_t *offsets;  // REC: nblocks + 1 record offsets
    uint64_t *status;   // REC: nblocks look-back words, zero at launch
    uint32_t *ovf;      // REC: MAX_GROUPS x OVF_WORDS, level-32 survivors of blocks too dense to stage in LDS
    uint32_t *nbytes;   // slot mode: 2 * nblocks sizes
    uint32_t *hdr;      // [0] block queue (zero at launch; REC, or slot mode with queue), [1] error bits
    uint32_t limit;     // look-back re-reads before a block gives up
    uint32_t zkeep;     // survivor bits kept by zerobits
    int shuffle;
    int queue;          // slot mode: take blocks from the queue in hdr[0] instead of a static stride
};

// Block b's record offset e (ok) or its give-up: status word, offsets, error bit.
CZW_DI void publish_offset(const CompressOut &o, int b, int nblocks, bool ok, uint64_t e, uint32_t n)
{
    if (ok) {
        if (b > 0) st_store(o.status + b, ST_P | (e + 4u + n));
        o.offsets[b] = e;
        if (b == nblocks - 1) o.offsets[nblocks] = e + 4u + n;
    } else {
        st_store(o.status + b, ST_E);
        atomicOr(o.hdr + 1, (uint32_t)CZW_ERR_LOOKBACK);
        o.offsets[b] = CZW_OFFSET_INVALID;
        if (b == nblocks - 1) o.offsets[nblocks] = CZW_OFFSET_INVALID;
    }
}

template <int WT, int LAYOUT, bool REC, int NTH>
__global__ __launch_bounds__(NTH) void k_compress_f32(const float *__restrict__ src, FieldGeom g, int nblocks,
                                                      float eps, CompressOut out,
                                                      const czw_field_t *__restrict__ fields, int nf)
{
    constexpr int RL = row_layout<LAYOUT>();
    constexpr int CR = CTile<NTH>::CR;
    __shared__ __align__(16) float A[LDS_WORDS];
    __shared__ uint32_t M[1024];
    __shared__ uint32_t cnt[2048];  // level-32 run counts, then their exclusive offsets
    __shared__ uint32_t wsum[NTH / 64];
    __shared__ uint32_t s32_total;
    __shared__ int tk[3];  // REC: block tickets, one iteration ahead of their use
    __shared__ uint64_t e_sh;  // record offset of the pending block
    __shared__ __align__(16) uint32_t SB[REC ? SB_WORDS : 1];  // REC: a parked block's mask and survivors
    __shared__ uint64_t e_q;   // REC: the parked block's record offset
    __shared__ czw_field_t T[LAYOUT == LAYOUT_BATCH ? BATCH_MAX_FIELDS : 1];
    const int t = threadIdx.x;
    const int G = gridDim.x;
    load_table<LAYOUT>(T, fields, nf);
    load_one<LAYOUT>(T, src, nullptr, g, eps);
    constexpr int TL = table_layout<LAYOUT>();
    const int ntab = one_field<LAYOUT>() ? 1 : nf;
    const bool queue = REC || out.queue;  // REC always queues: a block waits only for blocks issued before it
    // slot mode reports no device errors: clear the error word of the header (if given)
    if (!REC && out.hdr && blockIdx.x == 0 && t == 0) out.hdr[1] = 0u;
    if (queue) {
        if (t == 0) {
            tk[0] = (int)atomicAdd(out.hdr, 1u);
            tk[1] = (int)atomicAdd(out.hdr, 1u);
        }
        __syncthreads();
    }
    int b = queue ? tk[0] : (int)blockIdx.x;
    float nx[CR][32];
    float eps_b = eps;
    if (b < nblocks) {
        const BlockRef<float> br = locate<TL, float>(T, ntab, b, src, nullptr, g, eps);
        eps_b = br.eps;
        load_xrows_f32<RL, NTH, 0, 32>(br.src, br.g, br.lb, t, nx);
    }
    // The pending block bp: its payload staged in LDS (mask in M, survivors
    // from STAGE32_AT - pSc in A), written at the start of the workgroup's
    // next iteration.  REC: its record offset is found first, by look-back
    // on the last wave; slot mode: straight to its slot.
    int bp = -1;
    uint32_t pSc = 0, pS32 = 0;
    int qb = -1;  // REC: the block parked in SB, framed at the start of the next iteration
    uint32_t qSc = 0, qS32 = 0;
    for (int it = 0;; ++it) {
        const bool have = b < nblocks;
        if (!have && bp < 0) break;
        // the ticket of the block after next: taken now, stored in LDS at the
        // end of this block, so that no wave waits for the atomic's return
        int tkv = 0;
        if (queue && have && t == 0) tkv = (int)atomicAdd(out.hdr, 1u);
        if constexpr (REC) {
            const uint32_t zk = out.zkeep;
            // block bk's record at offset e: size word, mask mk, survivors
            // sv[0 .. SL) and the level-32 overflow area for the rest
            auto frame = [&](uint64_t e, uint32_t S, uint32_t SL, uint32_t Sc, const uint32_t *mk,
                             const uint32_t *sv) {
                const uint32_t n = MASK_BYTES + 4u * S;
                const uint32_t *ov = out.ovf + (size_t)blockIdx.x * OVF_WORDS;
                uint32_t *rec = reinterpret_cast<uint32_t *>(out.dst + e);
                if (t == 0) rec[0] = n;
#pragma unroll
                for (int k = 0; k < CR; ++k) rec[1 + t + NTH * k] = mk[t + NTH * k];
                if (!out.shuffle) {
                    for (uint32_t i = t; i < SL; i += NTH) rec[1 + 1024 + i] = sv[i] & zk;
                    for (uint32_t i = SL + t; i < S; i += NTH) rec[1 + 1024 + i] = ov[i - Sc] & zk;
                } else {  // byte p of survivor i at p*S + i (WaveletCompressor.cpp:56-73)
                    uint8_t *p8 = out.dst + e + 4 + MASK_BYTES;
                    auto put = [&](uint32_t i, uint32_t w) {
                        p8[i] = (uint8_t)w;
                        p8[S + i] = (uint8_t)(w >> 8);
                        p8[2 * S + i] = (uint8_t)(w >> 16);
                        p8[3 * S + i] = (uint8_t)(w >> 24);
                    };
                    for (uint32_t i = t; i < SL; i += NTH) put(i, sv[i] & zk);
                    for (uint32_t i = SL + t; i < S; i += NTH) put(i, ov[i - Sc] & zk);
                }
            };
            // block qb, parked in SB one block ago: its offset was found
            // during that block (e_q), so no wave waits for it here
            if (qb >= 0) {
                const uint64_t e = e_q;
                if (e != CZW_OFFSET_INVALID) frame(e, qSc + qS32, qSc + qS32, qSc, SB, SB + 1024);
                qb = -1;
            }
            __syncthreads();  // SB is read; bp's corner survivors were staged after its last barrier
            if (bp >= 0) {
                const uint32_t S = pSc + pS32;
                const uint32_t *st = reinterpret_cast<const uint32_t *>(A) + (STAGE32_AT - pSc);
                if (have && pS32 <= (uint32_t)STAGE32 && S <= (uint32_t)(SB_WORDS - 1024)) {
                    // park bp in SB: its offset is found during the next block
#pragma unroll
                    for (int k = 0; k < CR; ++k) SB[t + NTH * k] = M[t + NTH * k];
                    for (uint32_t i = t; i < S; i += NTH) SB[1024 + i] = st[i];
                    qb = bp;
                    qSc = pSc;
                    qS32 = pS32;
                } else {  // the last block, or too dense to park: look back now
                    const uint32_t n = MASK_BYTES + 4u * S;
                    if (t >= NTH - 64) {
                        LookBack lb;
                        look_back_issue(out.status, bp - 1, lb);
                        uint64_t e = 0;
                        const bool ok = look_back(out.status, bp, out.limit, lb, &e);
                        if (t == NTH - 64) {
                            publish_offset(out, bp, nblocks, ok, e, n);
                            e_sh = ok ? e : CZW_OFFSET_INVALID;
                        }
                    }
                    __syncthreads();
                    const uint64_t e = e_sh;
                    if (e != CZW_OFFSET_INVALID) frame(e, S, pS32 <= (uint32_t)STAGE32 ? S : pSc, pSc, M, st);
                }
                __syncthreads();  // A, M, SB and e_sh are free
                bp = -1;
            }
        } else if (bp >= 0) {
            const bool staged = pS32 <= (uint32_t)STAGE32;
            const uint32_t *st = reinterpret_cast<const uint32_t *>(A) + (STAGE32_AT - pSc);
            const uint32_t zk = out.zkeep;
            // the corner survivors were staged after the block's last barrier
            __syncthreads();
            uint32_t *slot = reinterpret_cast<uint32_t *>(out.dst + (size_t)bp * SLOT32);
#pragma unroll
            for (int k = 0; k < CR; ++k) slot[t + NTH * k] = M[t + NTH * k];
            for (uint32_t i = t; i < pSc; i += NTH) slot[1024 + i] = st[i] & zk;
            if (staged)
                for (uint32_t i = t; i < pS32; i += NTH) slot[1024 + CORNER_MAX + i] = st[pSc + i] & zk;
            __syncthreads();  // A and M are free
            bp = -1;
        }
        if (!have) break;
#pragma unroll
        for (int k = 0; k < CR; ++k) {  // level 32, x-pass, from the prefetched registers
            float r[32];
#pragma unroll
            for (int i = 0; i < 32; ++i) r[i] = nx[k][i];
            row<32, true, WT>(r);
            lds_st<32>(A, xrow_lds_base<RL>(xrow_of<RL, NTH>(t, k)), 1, r);
        }
        if (t < 2048 - NRUN32) cnt[NRUN32 + t] = 0u;
        __syncthreads();
        // Prefetch the next block's x-rows in four quarters of x, issued in
        // four phases: a CU keeps only some 48 KB of loads in flight, and a
        // wave that issues past that waits, so 32 KB at a time keeps the
        // waves computing while the loads run.  The last block reloads
        // itself, so nx is always written and is dead between its use and
        // its reload.
        const int bnext = queue ? tk[(it + 1) % 3] : b + G;
        const int bn = min(bnext, nblocks - 1);
        const float eps_c = eps_b;
        const BlockRef<float> brn = locate<TL, float>(T, ntab, bn, src, nullptr, g, eps);  // found once per block
        auto prefetch = [&](auto x0c) {
            constexpr int X0 = decltype(x0c)::value;
            if constexpr (X0 == 0) eps_b = brn.eps;
            load_xrows_f32<RL, NTH, X0, X0 + 8>(brn.src, brn.g, brn.lb, t, nx);
        };
        prefetch(std::integral_constant<int, 0>());
        {  // level 32, y-pass: rows (z, x) = (q >> 5, q & 31); every row is
           // loaded before the first is written back, so that the loads of
           // the second overlap the arithmetic of the first
            float r[CR][32];
#pragma unroll
            for (int k = 0; k < CR; ++k) lds_ld<32>(A, ((t + NTH * k) >> 5) * SP + ((t + NTH * k) & 31), RP, r[k]);
#pragma unroll
            for (int k = 0; k < CR; ++k) {
                row<32, true, WT>(r[k]);
                lds_st<32>(A, ((t + NTH * k) >> 5) * SP + ((t + NTH * k) & 31), RP, r[k]);
            }
        }
        __syncthreads();
        prefetch(std::integral_constant<int, 8>());
        // level 32, z-pass: row (y, x) = (q >> 5, q & 31) = physical mask word
        // x + 32*y, bit z (level-32 layout (a0,a1,a2) = (y,x,z)).  Outside the
        // 16^2 corner of (y,x) all 32 results are final; inside it only
        // z >= 16 is, and the level-16 z-pass continues on z < 16.  Keep bits
        // are taken here; the values stay in registers until they are staged.
        float r[CR][32];
        uint32_t kb[CR];
#pragma unroll
        for (int k = 0; k < CR; ++k) lds_ld<32>(A, ((t + NTH * k) >> 5) * RP + ((t + NTH * k) & 31), SP, r[k]);
#pragma unroll
        for (int k = 0; k < CR; ++k) {
            const int q = t + NTH * k, y = q >> 5, x = q & 31;
            const bool corner = (y < 16) && (x < 16);
            row<32, true, WT>(r[k]);
            uint32_t bits = 0;
#pragma unroll
            for (int z = 0; z < 32; ++z) bits |= (uint32_t)(fabsf(r[k][z]) > eps_c) << z;
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
            // stream run (code, i0 = y&15, i1 = x&15) of each z-half; code
            // bits: z >= 16, x >= 16, y >= 16 (FullWaveletTransform.h:68-72)
            const int c0 = 2 * (x >> 4) + 4 * (y >> 4);
            const int rbase = (y & 15) * 16 + (x & 15);
            if (!corner) cnt[(c0 - 1) * 256 + rbase] = __popc(bits & 0xffffu);
            cnt[c0 * 256 + rbase] = __popc(bits >> 16);
        }
        __syncthreads();
        {  // exclusive scan of the run counts in stream order
            constexpr int PER = 2048 / NTH;
            uint32_t c4[PER], sum = 0;
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                c4[i] = cnt[PER * t + i];
                sum += c4[i];
            }
            uint32_t S32;
            uint32_t run = block_excl_scan<NTH>(sum, wsum, &S32);
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                cnt[PER * t + i] = run;
                run += c4[i];
            }
            if (t == 0) s32_total = S32;
        }
        __syncthreads();
        const uint32_t S32 = s32_total;
        // Level-32 survivors, from registers.  The z >= 16 half of A is dead
        // after the z-pass (the remaining passes only touch the [0,16)^3
        // corner), so survivors are compacted there branch-free (unkept
        // values go to a per-lane dummy word).  A block with more survivors
        // than that half holds sends them to its overflow area (REC) or
        // straight to its slot.  Zerobits are applied on the way out.
        const bool staged = S32 <= (uint32_t)STAGE32;
#pragma unroll
        for (int k = 0; k < CR; ++k) {
            const int q = t + NTH * k, y = q >> 5, x = q & 31;
            const int c0 = 2 * (x >> 4) + 4 * (y >> 4);
            const int rbase = (y & 15) * 16 + (x & 15);
            uint32_t i0 = (c0 > 0) ? cnt[(c0 - 1) * 256 + rbase] : 0u;
            uint32_t i1 = cnt[c0 * 256 + rbase];
            if (staged) {
                // Byte addresses into A, each run walked from its last
                // survivor down: with m = 0 or -1 the keep bit, the target is
                // (m & a) | (~m & dummy) (v_bfi_b32) and the next a is a + 4*m
                // (v_lshl_add_u32), three VALU per position with the bit
                // extract.  Both are written as instructions, as the compiler
                // otherwise rebuilds the addresses from a running index.
                // Groups of eight positions z that no row of the wave keeps
                // are skipped on a scalar test of the wave's OR of the keep
                // bits (most of them at the densities the threshold leaves);
                // a kept group runs branch-free (a test per position made
                // two taken branches per kept position: 2.7% slower).
                char *Ab = reinterpret_cast<char *>(A);
                const uint32_t n0 = __popc(kb[k] & 0xffffu), n1 = __popc(kb[k] >> 16);
                uint32_t a0 = (STAGE32_AT + i0 + n0 - 1) * 4u, a1 = (STAGE32_AT + i1 + n1 - 1) * 4u;
                const uint32_t ad = (STAGE32_AT + STAGE32 + (t & 63)) * 4u;
                const uint32_t wk = wave_or(kb[k]);
                constexpr int C = 8;
#pragma unroll
                for (int c = 32 / C - 1; c >= 0; --c) {
                    if (!((wk >> (C * c)) & ((1u << C) - 1u))) continue;
#pragma unroll
                    for (int z = C * c + C - 1; z >= C * c; --z) {
                        const uint32_t m = (uint32_t)((int32_t)(kb[k] << (31 - z)) >> 31);
                        uint32_t &a = z < 16 ? a0 : a1;
                        uint32_t at;
                        asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(at) : "v"(m), "v"(a), "v"(ad));
                        *reinterpret_cast<float *>(Ab + at) = r[k][z];
                        asm("v_lshl_add_u32 %0, %1, 2, %2" : "=v"(a) : "v"(m), "v"(a));
                    }
                }
            } else {
                uint32_t *d;
                uint32_t zk;
                if constexpr (REC) {
                    d = out.ovf + (size_t)blockIdx.x * OVF_WORDS;
                    zk = 0xffffffffu;
                } else {
                    d = reinterpret_cast<uint32_t *>(out.dst + (size_t)b * SLOT32) + 1024 + CORNER_MAX;
                    zk = out.zkeep;
                }
#pragma unroll
                for (int z = 0; z < 32; ++z) {
                    uint32_t &i = z < 16 ? i0 : i1;
                    if ((kb[k] >> z) & 1u) d[i++] = __float_as_uint(r[k][z]) & zk;
                }
            }
        }
        prefetch(std::integral_constant<int, 16>());
        __syncthreads();
        // Levels 16 and 8 on the [0,16)^3 corner keep 256, then 64 threads
        // busy; every thread issues its last prefetch quarter in a pass where
        // it has no rows.  REC: the last wave meanwhile finds the parked
        // block's record offset (its status reads issued here, evaluated
        // after the level-16 passes).
        {
            float c[16];
            LookBack lbq;
            if constexpr (REC)
                if (qb >= 0 && t >= NTH - 64) look_back_issue(out.status, qb - 1, lbq);
            if (t < 256) {  // level 16 x-pass: row (z, y) in [0,16)^2
                const int z = t >> 4, yy = t & 15;
                lds_ld<16>(A, z * SP + yy * RP, 1, c);
                row<16, true, WT>(c);
                lds_st<16>(A, z * SP + yy * RP, 1, c);
            } else {
                prefetch(std::integral_constant<int, 24>());
            }
            __syncthreads();
            if (t < 256) {  // level 16 y-pass, then level 8 y-pass: row (z, x)
                int z, xx;
                corner_yrow(t, z, xx);
                lds_ld<16>(A, z * SP + xx, RP, c);
                row<16, true, WT>(c);
                if (t < 64) row<8, true, WT>(c);
                lds_st<16>(A, z * SP + xx, RP, c);
            }
            if (t < 64) {  // the level-8 passes stay in the first wave: no workgroup barrier
                wave_lds_order();
                {  // level 8 z-pass: row (y, x) in [0,8)^2
                    const int yy = t >> 3, xx = t & 7;
                    lds_ld<8>(A, yy * RP + xx, SP, c);
                    row<8, true, WT>(c);
                    lds_st<8>(A, yy * RP + xx, SP, c);
                }
                wave_lds_order();
                {  // level 8 x-pass: row (z, y)
                    const int z = t >> 3, yy = t & 7;
                    lds_ld<8>(A, z * SP + yy * RP, 1, c);
                    row<8, true, WT>(c);
                    lds_st<8>(A, z * SP + yy * RP, 1, c);
                }
            } else if (t < 256) {
                prefetch(std::integral_constant<int, 24>());
            }
            if constexpr (REC) {
                if (qb >= 0 && t >= NTH - 64) {
                    uint64_t e = 0;
                    const bool ok = look_back(out.status, qb, out.limit, lbq, &e);
                    if (t == NTH - 64) {
                        publish_offset(out, qb, nblocks, ok, e, MASK_BYTES + 4u * (qSc + qS32));
                        e_q = ok ? e : CZW_OFFSET_INVALID;
                    }
                }
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
                    const float cv = A[addr + i * stride];
                    v[q * h + i] = cv;
                    bits |= (uint32_t)((t < 8) || (fabsf(cv) > eps_c)) << i;
                }
            }
            keep |= bits << (q * h);
            atomicOr(&M[word], bits << shift);
        });
        uint32_t Sc;
        uint32_t idx = block_excl_scan<NTH>(__popc(keep), wsum, &Sc);
        {  // corner survivors just below the level-32 ones (the corner's values
           // are all read: the scan's barrier is behind every read)
            float *sc = A + (STAGE32_AT - Sc);
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if ((keep >> j) & 1u) sc[idx++] = v[j];
        }
        if (t == 0) {
            const uint32_t n = MASK_BYTES + 4u * (Sc + S32);
            if constexpr (REC) {
                st_store(out.status + b, (b == 0 ? ST_P : ST_A) | (uint64_t)(4u + n));
            } else {
                out.nbytes[b] = n;
                out.nbytes[nblocks + b] = Sc;
            }
            if (queue) tk[(it + 2) % 3] = tkv;
        }
        bp = b;
        pSc = Sc;
        pS32 = S32;
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
constexpr int PK_T = 512, PK_W = 8, PK_B = 2, PK_FUSED_MAX = 4096;
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
        // the sizes of at most PK_FUSED_MAX records sum to less than 2^32
        const uint32_t wpart = wave_incl_scan((uint32_t)part, t & 63);
        if ((t & 63) == 63) wsum[t >> 6] = wpart;
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

// Compressor workspace.  Both precisions start with a 256-byte header: word 0
// the compressor's block queue, word 1 its device error bits (CZW_ERR_*),
// word 2 the batch decoder's block queue.  Then the per-block slots and size
// words of the two-stage path (compress into slots, frame into records); or,
// for the binary32 record path of czw_compress_field_lowmem, the look-back
// status words (8 bytes per block) and an overflow area for each of up to
// MAX_GROUPS workgroups.
constexpr size_t WS_HDR = 256;

static size_t ovf_groups(int nblocks) { return (size_t)std::min(nblocks, MAX_GROUPS); }

static size_t slot_bytes_of(int prec) { return MASK_BYTES + (size_t)prec * NB3; }

static size_t workspace_bytes(int nblocks, int prec)
{
    if (nblocks <= 0 || (prec != 4 && prec != 8)) return 0;
    return WS_HDR + (slot_bytes_of(prec) + 4 * size_words(prec)) * (size_t)nblocks;
}

static size_t lowmem_workspace_bytes(int nblocks, int prec)
{
    if (nblocks <= 0 || prec != 4) return 0;
    return WS_HDR + 8 * (size_t)nblocks + ovf_groups(nblocks) * OVF_WORDS * 4;
}

static std::atomic<uint32_t> g_lookback_limit{1u << 20};

// Clears n words from p (a fill kernel rather than a memset, so that a
// captured graph replays it).
__global__ void k_fill_u32(uint32_t *__restrict__ p, int n, uint32_t v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

static int fill_u32(uint32_t *p, int n, uint32_t v, hipStream_t s)
{
    const int grid = std::max(1, std::min(256, (n + 1023) / 1024));
    hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, s, p, n, v);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

static uint32_t zkeep_of(int zerobits) { return zerobits ? ~((1u << zerobits) - 1u) : 0xffffffffu; }

template <int LAYOUT, bool REC, int NTH = CT>
static const void *compress_f32_fn(int wt)
{
    return wt == 1 ? (const void *)k_compress_f32<1, LAYOUT, REC, NTH>
         : wt == 2 ? (const void *)k_compress_f32<2, LAYOUT, REC, NTH>
                   : (const void *)k_compress_f32<3, LAYOUT, REC, NTH>;
}

// Launches k_compress_f32 for wavelet index wt on min(nblocks, resident
// workgroups[, MAX_GROUPS for REC][, groups]) workgroups: never more than
// are resident, so that no second persistent wave waits behind the first.
template <int LAYOUT, bool REC, int NTH = CT>
static int launch_compress_f32(int wt, const float *src, FieldGeom g, int nblocks, float eps, const CompressOut &o,
                               const czw_field_t *fields, int nf, int groups, hipStream_t s)
{
    int G = std::min(nblocks, resident_groups(compress_f32_fn<LAYOUT, REC, NTH>(wt), NTH));
    if (REC) G = std::min(G, MAX_GROUPS);
    if (groups > 0) G = std::min(G, groups);
    const dim3 grid(G), blk(NTH);
    if (wt == 1) hipLaunchKernelGGL((k_compress_f32<1, LAYOUT, REC, NTH>), grid, blk, 0, s, src, g, nblocks, eps, o, fields, nf);
    if (wt == 2) hipLaunchKernelGGL((k_compress_f32<2, LAYOUT, REC, NTH>), grid, blk, 0, s, src, g, nblocks, eps, o, fields, nf);
    if (wt == 3) hipLaunchKernelGGL((k_compress_f32<3, LAYOUT, REC, NTH>), grid, blk, 0, s, src, g, nblocks, eps, o, fields, nf);
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

// Stage 1 of the binary32 path: payloads into slots, sizes into nb.  With
// qhdr, blocks come from the queue in qhdr[0] (the caller clears it);
// otherwise the workspace header hdr (if given) gets its error word cleared.
template <int LAYOUT>
static int compress_slots_f32(const float *src, int nblocks, FieldGeom g, float eps, int wtype, int zerobits,
                              uint8_t *slots, uint32_t *nb, const czw_field_t *fields, int nf, int groups,
                              uint32_t *qhdr, uint32_t *hdr, hipStream_t s)
{
    const int wt = wt_index(wtype);
    if (wt < 0 || !zb_ok(zerobits) || nblocks <= 0 || !slots || !nb) return CZW_EINVAL;
    if (LAYOUT != LAYOUT_BATCH && !src) return CZW_EINVAL;
    const CompressOut o{slots, nullptr, nullptr, nullptr, nb, qhdr ? qhdr : hdr, 0u, zkeep_of(zerobits), 0,
                        qhdr ? 1 : 0};
    return launch_compress_f32<LAYOUT, false>(wt, src, g, nblocks, eps, o, fields, nf, groups, s);
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

// Both precisions, two stages: clear the header, compress into slots, frame
// the slots into records.  prec 8 uses the half-block compressor.
template <int LAYOUT>
static int compress_2stage(int prec, const void *src, int nblocks, FieldGeom g, float eps, int wtype, int zerobits,
                           int shuffle, void *records, uint64_t *offsets, void *workspace,
                           const czw_field_t *fields, int nf, const czw_shape_t &shape, hipStream_t s)
{
    if (!records || !offsets || !workspace || nblocks <= 0 || (prec != 4 && prec != 8)) return CZW_EINVAL;
    uint32_t *hdr = static_cast<uint32_t *>(workspace);
    // the queue and the error word start at zero (binary32 without a queue:
    // the compressor clears the error word itself, no extra launch)
    if ((prec == 8 || shape.queue) && fill_u32(hdr, (int)(WS_HDR / 4), 0u, s) != CZW_OK) return CZW_EHIP;
    uint8_t *slots = static_cast<uint8_t *>(workspace) + WS_HDR;
    uint32_t *nb = reinterpret_cast<uint32_t *>(slots + slot_bytes_of(prec) * (size_t)nblocks);
    uint32_t *q = shape.queue ? hdr : nullptr;
    int rc;
    if (prec == 4)
        rc = compress_slots_f32<LAYOUT>(static_cast<const float *>(src), nblocks, g, eps, wtype, zerobits, slots, nb,
                                        fields, nf, shape.groups, q, hdr, s);
    else
        rc = compress_slots_f64(LAYOUT, src, nblocks, g, eps, wtype, zerobits, slots, nb, fields, nf, shape.groups, q,
                                s);
    if (rc != CZW_OK) return rc;
    if (prec == 4) return pack_records(slots, SLOT32, nb, nblocks, shuffle, static_cast<uint8_t *>(records), offsets, s);
    return pack_records_f64(slots, nb, nblocks, shuffle, static_cast<uint8_t *>(records), offsets, s);
}

// Binary32 compress straight to records with a small workspace: one fill of
// the header and status words, one compressor launch (record offsets by
// look-back).  `scratch` is lowmem_workspace_bytes(nblocks, 4) bytes.
template <int LAYOUT>
static int compress_rec_f32(const float *src, int nblocks, FieldGeom g, float eps, int wtype, int zerobits,
                            int shuffle, void *records, uint64_t *offsets, void *scratch, hipStream_t s)
{
    const int wt = wt_index(wtype);
    if (wt < 0 || !zb_ok(zerobits) || nblocks <= 0 || !records || !offsets || !scratch || !src) return CZW_EINVAL;
    uint8_t *ws = static_cast<uint8_t *>(scratch);
    uint32_t *hdr = reinterpret_cast<uint32_t *>(ws);
    uint64_t *status = reinterpret_cast<uint64_t *>(ws + WS_HDR);
    uint32_t *ovf = reinterpret_cast<uint32_t *>(ws + WS_HDR + 8 * (size_t)nblocks);
    // the header's words and the status words are contiguous 32-bit words
    if (fill_u32(hdr, (int)(WS_HDR / 4) + 2 * nblocks, 0u, s) != CZW_OK) return CZW_EHIP;
    const CompressOut o{static_cast<uint8_t *>(records), offsets, status, ovf, nullptr, hdr,
                        g_lookback_limit.load(std::memory_order_relaxed), zkeep_of(zerobits), shuffle ? 1 : 0, 1};
    return launch_compress_f32<LAYOUT, true>(wt, src, g, nblocks, eps, o, nullptr, 0, 0, s);
}

// ---------------------------------------------------------------------------
// Allocation-free variant of czw_compress_field_lowmem: its scratch is a
// file-scope device array of fixed size, one copy per GPU, whose address is
// looked up once per device (hipGetSymbolAddress) and used from then on.
// Every call on a GPU shares that GPU's array, so calls that overlap on one
// GPU race; the fill at the start of each call clears what must start at
// zero (header and status words).
constexpr int NOALLOC_MAX_BLOCKS = 32768;  // 1024^3 binary32 values
constexpr size_t NOALLOC_BYTES = WS_HDR + 8 * (size_t)NOALLOC_MAX_BLOCKS + (size_t)MAX_GROUPS * OVF_WORDS * 4;
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

static void *scratch_ptr()  // this device's copy of scratch, nullptr if it cannot be found
{
    constexpr int MAXDEV = 64;
    static void *ptr[MAXDEV];
    static std::mutex mu;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAXDEV) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        ptr[dev] = p;
    }
    return ptr[dev];
}

// Every layout of a single field, both paths.
static int compress_any(int layout, const void *src, int nblocks, FieldGeom g, int prec, float eps, int wtype,
                        int zerobits, int shuffle, void *records, uint64_t *offsets, void *workspace, bool lowmem,
                        hipStream_t s)
{
    if (lowmem) {
        if (prec != 4) return CZW_EUNSUPPORTED;
        const float *f = static_cast<const float *>(src);
        switch (layout) {
        case LAYOUT_BLOCKS:
            return compress_rec_f32<LAYOUT_BLOCKS>(f, nblocks, g, eps, wtype, zerobits, shuffle, records, offsets,
                                                   workspace, s);
        case LAYOUT_FIELD:
            return compress_rec_f32<LAYOUT_FIELD>(f, nblocks, g, eps, wtype, zerobits, shuffle, records, offsets,
                                                  workspace, s);
        case LAYOUT_FIELD_WIDE:
            return compress_rec_f32<LAYOUT_FIELD_WIDE>(f, nblocks, g, eps, wtype, zerobits, shuffle, records,
                                                       offsets, workspace, s);
        default:
            return CZW_EINVAL;
        }
    }
    const czw_shape_t sh{0, 0};
    switch (layout) {
    case LAYOUT_BLOCKS:
        return compress_2stage<LAYOUT_BLOCKS>(prec, src, nblocks, g, eps, wtype, zerobits, shuffle, records, offsets,
                                              workspace, nullptr, 0, sh, s);
    case LAYOUT_FIELD:
        return compress_2stage<LAYOUT_FIELD>(prec, src, nblocks, g, eps, wtype, zerobits, shuffle, records, offsets,
                                             workspace, nullptr, 0, sh, s);
    case LAYOUT_FIELD_WIDE:
        return compress_2stage<LAYOUT_FIELD_WIDE>(prec, src, nblocks, g, eps, wtype, zerobits, shuffle, records,
                                                  offsets, workspace, nullptr, 0, sh, s);
    default:
        return CZW_EINVAL;
    }
}

// ---------------------------------------------------------------------------
// batched mode

static_assert(sizeof(czw_field_t) == 40, "czw_field_t layout");
static_assert(BATCH_MAX_FIELDS == CZW_BATCH_MAX_FIELDS, "field table size");

static bool batch_ok(const czw_batch_t *B)
{
    return B && B->fields && B->nfields > 0 && B->nfields <= BATCH_MAX_FIELDS && B->nblocks > 0 && B->workspace &&
           B->records && B->offsets && B->compress.groups >= 0 && B->decompress.groups >= 0 &&
           (B->compress.tile == CZW_TILE_BLOCK || B->compress.tile == CZW_TILE_HALF) && wt_index(B->wtype) > 0 &&
           (B->prec == 4 || B->prec == 8);
}

// A batch's workspace: the header, the slots, then size_words of either
// tile per block (binary32: 9 for the half-block tile's region counts).
static size_t batch_workspace_bytes(int nblocks, int prec)
{
    if (nblocks <= 0 || (prec != 4 && prec != 8)) return 0;
    return WS_HDR + (slot_bytes_of(prec) + 4 * size_words(8)) * (size_t)nblocks;
}

static int batch_compress(const czw_batch_t *B, hipStream_t s)
{
    if (!batch_ok(B) || !zb_ok(B->zerobits)) return CZW_EINVAL;
    if (B->prec == 4 && B->compress.tile == CZW_TILE_HALF) {
        uint32_t *hdr = static_cast<uint32_t *>(B->workspace);
        if (fill_u32(hdr, (int)(WS_HDR / 4), 0u, s) != CZW_OK) return CZW_EHIP;
        uint8_t *slots = static_cast<uint8_t *>(B->workspace) + WS_HDR;
        uint32_t *nb = reinterpret_cast<uint32_t *>(slots + SLOT32 * (size_t)B->nblocks);
        const int rc = compress_slots_h32_batch(B->fields, B->nfields, B->nblocks, B->wtype, B->zerobits, slots, nb,
                                                B->compress.groups, B->compress.queue ? hdr : nullptr, s);
        if (rc != CZW_OK) return rc;
        return pack_records_h32(slots, nb, B->nblocks, B->shuffle, static_cast<uint8_t *>(B->records), B->offsets,
                                s);
    }
    return compress_2stage<LAYOUT_BATCH>(B->prec, nullptr, B->nblocks, FieldGeom{0, 0, 0}, 0.0f, B->wtype,
                                         B->zerobits, B->shuffle, B->records, B->offsets, B->workspace, B->fields,
                                         B->nfields, B->compress, s);
}

static int batch_decompress(const czw_batch_t *B, hipStream_t s)
{
    if (!batch_ok(B) || !B->err) return CZW_EINVAL;
    uint32_t *q = B->decompress.queue ? static_cast<uint32_t *>(B->workspace) + 2 : nullptr;
    if (fill_u32(B->err, 1, 0u, s) != CZW_OK) return CZW_EHIP;
    if (q && fill_u32(q, 1, 0u, s) != CZW_OK) return CZW_EHIP;
    return decompress_half(B->prec, LAYOUT_BATCH, B->records, B->records_bytes, B->offsets, B->nblocks,
                           FieldGeom{0, 0, 0}, B->wtype, B->shuffle, nullptr, B->err, B->fields, B->nfields,
                           B->decompress.groups, q, s);
}

// Resident workgroups of the kernel a batch stage launches.
static int batch_resident(const czw_batch_t *B, int stage, int tile)
{
    const int wt = wt_index(B->wtype);
    if (stage == 1 && B->prec == 4 && tile == CZW_TILE_HALF) return resident_compress_h32(B->wtype);
    if (stage == 1 && B->prec == 4) return resident_groups(compress_f32_fn<LAYOUT_BATCH, false>(wt), CT);
    return stage == 1 ? resident_compress_f64(LAYOUT_BATCH, B->wtype)
                      : resident_decompress(B->prec, LAYOUT_BATCH, B->wtype, B->shuffle);
}

} // namespace czw

using namespace czw;

extern "C" {

int czw_version(void) { return 200; }

size_t czw_compress_workspace_bytes(int nblocks, int prec) { return workspace_bytes(nblocks, prec); }

size_t czw_records_bound(int nblocks, int prec)
{
    if (nblocks <= 0 || (prec != 4 && prec != 8)) return 0;
    return (4 + MASK_BYTES + (size_t)prec * NB3) * (size_t)nblocks;
}

int czw_compress_status(const void *workspace, hipStream_t stream)
{
    if (!workspace) return CZW_EINVAL;
    uint32_t e = 0;
    if (hipMemcpyAsync(&e, static_cast<const uint8_t *>(workspace) + 4, 4, hipMemcpyDeviceToHost, stream) !=
            hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return CZW_EHIP;
    return e ? CZW_EDEVICE : CZW_OK;
}

uint32_t czw_set_lookback_limit(uint32_t rereads)
{
    return g_lookback_limit.exchange(rereads ? rereads : (1u << 20));
}

int czw_compress_field(const void *field, int nx, int ny, int nz, int prec, float eps, int wtype, int zerobits,
                       int shuffle, void *records, uint64_t *offsets, void *workspace, hipStream_t stream)
{
    if (!field_ok(nx, ny, nz) || !field) return CZW_EINVAL;
    const int nb = (nx / 32) * (ny / 32) * (nz / 32);
    const int layout = field_fits_buffer(ny, nz, prec) ? LAYOUT_FIELD : LAYOUT_FIELD_WIDE;
    return compress_any(layout, field, nb, FieldGeom{nx, ny, nz}, prec, eps, wtype, zerobits, shuffle, records,
                        offsets, workspace, false, stream);
}

size_t czw_compress_lowmem_bytes(int nblocks, int prec) { return lowmem_workspace_bytes(nblocks, prec); }

int czw_compress_field_noalloc(const void *field, int nx, int ny, int nz, int prec, float eps, int wtype, int zerobits,
                               int shuffle, void *records, uint64_t *offsets, hipStream_t stream)
{
    if (!field_ok(nx, ny, nz) || !field) return CZW_EINVAL;
    if (prec != 4) return CZW_EUNSUPPORTED;
    const int nb = (nx / 32) * (ny / 32) * (nz / 32);
    if (nb > NOALLOC_MAX_BLOCKS) return CZW_ETOOBIG;
    void *p = scratch_ptr();
    if (!p) return CZW_EHIP;
    const int layout = field_fits_buffer(ny, nz, prec) ? LAYOUT_FIELD : LAYOUT_FIELD_WIDE;
    return compress_any(layout, field, nb, FieldGeom{nx, ny, nz}, prec, eps, wtype, zerobits, shuffle, records,
                        offsets, p, true, stream);
}

int czw_noalloc_errors(uint32_t *bits, hipStream_t stream)
{
    void *p = scratch_ptr();
    if (!p || !bits) return p ? CZW_EINVAL : CZW_EHIP;
    if (hipMemcpyAsync(bits, static_cast<uint8_t *>(p) + 4, 4, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return CZW_EHIP;
    return CZW_OK;
}

int czw_compress_field_lowmem(const void *field, int nx, int ny, int nz, int prec, float eps, int wtype, int zerobits,
                              int shuffle, void *records, uint64_t *offsets, void *scratch, hipStream_t stream)
{
    if (!field_ok(nx, ny, nz) || !field) return CZW_EINVAL;
    const int nb = (nx / 32) * (ny / 32) * (nz / 32);
    const int layout = field_fits_buffer(ny, nz, prec) ? LAYOUT_FIELD : LAYOUT_FIELD_WIDE;
    return compress_any(layout, field, nb, FieldGeom{nx, ny, nz}, prec, eps, wtype, zerobits, shuffle, records,
                        offsets, scratch, true, stream);
}

int czw_compress_field_range(const void *field, int nx, int ny, int nz, int prec, float eps, int wtype,
                             int zerobits, int shuffle, int block0, int nblocks, void *records, uint64_t *offsets,
                             void *workspace, hipStream_t stream)
{
    if (!field_ok(nx, ny, nz) || !field) return CZW_EINVAL;
    const int total = (nx / 32) * (ny / 32) * (nz / 32);
    if (block0 < 0 || nblocks <= 0 || block0 > total - nblocks) return CZW_EINVAL;
    const int layout = field_fits_buffer(ny, nz, prec) ? LAYOUT_FIELD : LAYOUT_FIELD_WIDE;
    return compress_any(layout, field, nblocks, FieldGeom{nx, ny, nz, block0}, prec, eps, wtype, zerobits, shuffle,
                        records, offsets, workspace, false, stream);
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
                                                      static_cast<uint8_t *>(slots), nbytes, nullptr, 0, 0, nullptr,
                                                      nullptr, stream)
                   : compress_slots_f32<LAYOUT_FIELD_WIDE>(static_cast<const float *>(field), nb, g, eps, wtype,
                                                           zerobits, static_cast<uint8_t *>(slots), nbytes, nullptr, 0,
                                                           0, nullptr, nullptr, stream);
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
    if (!blocks) return CZW_EINVAL;
    return compress_any(LAYOUT_BLOCKS, blocks, nblocks, FieldGeom{0, 0, 0}, prec, eps, wtype, zerobits, shuffle,
                        records, offsets, workspace, false, stream);
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

size_t czw_batch_workspace_bytes(int nblocks, int prec) { return batch_workspace_bytes(nblocks, prec); }

int czw_batch_compress(const czw_batch_t *batch, hipStream_t stream) { return batch_compress(batch, stream); }

int czw_batch_decompress(const czw_batch_t *batch, hipStream_t stream) { return batch_decompress(batch, stream); }

int czw_batch_autotune(czw_batch_t *batch, int stage, hipStream_t stream, float *best_ms)
{
    if (!batch_ok(batch) || (stage != 1 && stage != 2)) return CZW_EINVAL;
    // every resident workgroup of the stage's kernel, or half of them; with
    // the block queue or in static order; binary32 compress: either tile
    int shapes[8][3];
    int nshapes = 0;
    const int ntiles = (stage == 1 && batch->prec == 4) ? 2 : 1;
    for (int tile = 0; tile < ntiles; ++tile) {
        const int R = std::max(1, batch_resident(batch, stage, tile));
        const int g[4][2] = {{R, 1}, {std::max(1, R / 2), 1}, {R, 0}, {std::max(1, R / 2), 0}};
        for (int k = 0; k < 4; ++k) {
            shapes[nshapes][0] = g[k][0];
            shapes[nshapes][1] = g[k][1];
            shapes[nshapes][2] = tile;
            ++nshapes;
        }
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return CZW_EHIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return CZW_EHIP;
    }
    czw_batch_t trial = *batch;
    czw_shape_t &sh = stage == 1 ? trial.compress : trial.decompress;
    float best = 1e30f;
    czw_shape_t win = sh;
    int rc = CZW_OK;
    for (int k = 0; k < nshapes && rc == CZW_OK; ++k) {
        sh.groups = shapes[k][0];
        sh.queue = shapes[k][1];
        if (stage == 1) sh.tile = shapes[k][2];
        auto run = [&]() { return stage == 1 ? batch_compress(&trial, stream) : batch_decompress(&trial, stream); };
        if ((rc = run()) != CZW_OK) break;  // warm-up
        (void)hipEventRecord(e0, stream);
        for (int r = 0; r < 3 && rc == CZW_OK; ++r) rc = run();
        (void)hipEventRecord(e1, stream);
        if (hipEventSynchronize(e1) != hipSuccess) rc = CZW_EHIP;
        float ms = 0.0f;
        if (rc == CZW_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms / 3 < best) {
            best = ms / 3;
            win = sh;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != CZW_OK) return rc;
    (stage == 1 ? batch->compress : batch->decompress) = win;  // the other stage's shape is left as it is
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
