/*
 * czw.h -- C ABI of the MI355X (gfx950) wavelet block compressor.
 *
 * This is the device-side replacement for CubismZ's per-block compressor
 * WaveletCompressorGeneric<32,Real> (Compressor/source/WaveletCompressor.h:35-94)
 * and for the per-block loops that drive it: the writer loop
 * SerializerIO_WaveletCompression_MPI_SimpleBlocking::_compress
 * (Compressor/source/SerializerIO_WaveletCompression_MPI_Simple.h:187-345) and
 * the reader Reader_WaveletCompression::load_block2
 * (Compressor/source/Reader_WaveletCompression.h:635-754).  Instead of one
 * 32^3 block per call, every entry point handles a batch of blocks on one HIP
 * stream.
 *
 * Conventions
 *  - All data pointers are DEVICE pointers owned by the caller; the library
 *    allocates nothing in these calls.
 *  - prec is sizeof(Real): 4 (the reference's float build) or 8 (double build).
 *  - wtype follows the reference: 1 = InterpWavelet4thOrder,
 *    2 = LiftedInterpWavelet4thOrder, 3 = AverageInterpWavelet3rdOrder,
 *    0 = "None" (computes like 3, FullWaveletTransform.h:28-35).
 *  - eps is a float as in WaveletCompressorGeneric::compress(const float ...)
 *    (WaveletCompressor.h:61); the double build compares against (double)eps.
 *  - zerobits in {0,4,8,12,16} clears the low bits of the first 32-bit word
 *    of every survivor (WaveletCompressor.cpp:34-43,166-173); shuffle != 0
 *    applies the byte-plane shuffle (WaveletCompressor.cpp:56-93).
 *  - Calls are asynchronous on `stream`; they return a status code and never
 *    abort.  Concurrent calls on different streams are safe if they do not
 *    share buffers.
 *
 * Record stream ("records")
 *   For block b, at byte offsets[b]:  int32 nbytes, then the payload of
 *   nbytes bytes = [mask 4096 B, bit p at byte p>>3 bit p&7][S survivors].
 *   offsets has nblocks+1 entries (offsets[nblocks] = total bytes).  This is
 *   byte-for-byte the per-thread chunk buffer of the reference writer
 *   (SerializerIO_WaveletCompression_MPI_Simple.h:237-240), so any run of
 *   consecutive records can be handed to zlib as one .cz chunk.
 *
 * Field layout ("field" entry points)
 *   An HDF5-ordered field f[NX][NY][NZ] (NZ fastest), each dimension a
 *   multiple of 32.  Block (ix,iy,iz) holds block[z][y][x] =
 *   f[32ix+x][32iy+y][32iz+z] (Tools/Test_IO_Compressed.h:200-212) and blocks
 *   are numbered b = ix + bx*(iy + by*iz) (Cubism Grid.h:68).
 * Block layout ("blocks" entry points)
 *   blocks[nblocks][32][32][32] in the reference's uncompressed_data() order
 *   [z][y][x], x fastest (WaveletCompressor.h:57).
 */
#ifndef CZW_H
#define CZW_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CZW_OK = 0,
    CZW_EINVAL = -1,      /* bad argument (sizes, wtype, zerobits, null) */
    CZW_EHIP = -2,        /* a HIP launch failed */
    CZW_EUNSUPPORTED = -3, /* combination not built */
    CZW_EDEVICE = -4,      /* the device reported an error (czw_compress_status) */
    CZW_ETOOBIG = -5       /* more blocks than a fixed-size scratch holds */
};

/* Device error bits (OR-ed): the decompressors write them into *err, the
 * binary32 compressors into the 32-bit word at byte 4 of their workspace. */
enum {
    CZW_ERR_NBYTES = 1,   /* record size/offset invalid or outside the records buffer */
    CZW_ERR_POPCOUNT = 2, /* mask popcount differs from the survivor count */
    CZW_ERR_LOOKBACK = 4  /* a block gave up waiting for its predecessors' record sizes */
};

/* offsets[] value of a block whose record was not written (CZW_ERR_LOOKBACK). */
#define CZW_OFFSET_INVALID UINT64_MAX

/* Library version (major*100 + minor). */
int czw_version(void);

/* Bytes of scratch the compressors need for nblocks blocks: a 256-byte
 * header (its 32-bit word 1, byte 4, holds the device error bits of the last
 * compress that used it; 0 = every record written), then one fixed-size
 * slot per block (czw_slot_bytes) and the per-block size words. */
size_t czw_compress_workspace_bytes(int nblocks, int prec);

/* Upper bound of the record stream size for nblocks blocks. */
size_t czw_records_bound(int nblocks, int prec);

/* Replaces the loop body of SerializerIO_..._MPI_Simple.h:212-240 (gather a
 * block, WaveletCompressorGeneric::compress at WaveletCompressor.cpp:153-182,
 * frame the record) for every block of an HDF5-ordered field.  Two stages on
 * `stream`: the compressor writes each block's payload into its slot of the
 * workspace, then a framing kernel concatenates the slots into the record
 * stream and writes the offsets. */
int czw_compress_field(const void *field, int nx, int ny, int nz, int prec, float eps, int wtype,
                       int zerobits, int shuffle, void *records, uint64_t *offsets, void *workspace,
                       hipStream_t stream);

/* czw_compress_field with a small workspace: scratch of
 * czw_compress_lowmem_bytes(nblocks, prec) bytes, 256 + 8*nblocks +
 * min(nblocks, 256) * 114688 bytes, instead of one slot per block (binary32
 * only; CZW_EUNSUPPORTED for prec 8).  One compressor launch writes the
 * records itself: each block's record offset, the sum of its predecessors'
 * record sizes, is found on the device by decoupled look-back, and a block
 * waits only for blocks taken from the block queue before it.  A block that
 * waits longer than the look-back limit gives up: its record is not written,
 * offsets[b] = CZW_OFFSET_INVALID and the scratch's error word gets
 * CZW_ERR_LOOKBACK (read it with czw_compress_status).  That takes another
 * kernel holding the device's compute units as long, e.g. one that waits on
 * this call.  Same records and offsets as czw_compress_field. */
size_t czw_compress_lowmem_bytes(int nblocks, int prec);
int czw_compress_field_lowmem(const void *field, int nx, int ny, int nz, int prec, float eps, int wtype,
                              int zerobits, int shuffle, void *records, uint64_t *offsets, void *scratch,
                              hipStream_t stream);

/* czw_compress_field_lowmem without the scratch argument; no counterpart in
 * the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * the call allocates nothing and uses a fixed device array of the library,
 * about 28 MiB, one per GPU, held until the process ends.  At most 32768
 * blocks (a 1024^3 binary32 field); more returns CZW_ETOOBIG; prec 8
 * returns CZW_EUNSUPPORTED.  Every call on one GPU shares that array, so
 * calls that overlap on one GPU (different streams, or host threads) race:
 * serialize them.  czw_noalloc_errors waits for `stream` and returns the
 * array's error bits (CZW_ERR_LOOKBACK) from the last call on this GPU. */
int czw_compress_field_noalloc(const void *field, int nx, int ny, int nz, int prec, float eps, int wtype,
                               int zerobits, int shuffle, void *records, uint64_t *offsets, hipStream_t stream);
int czw_noalloc_errors(uint32_t *bits, hipStream_t stream);

/* czw_compress_field for the field's blocks [block0, block0 + nblocks) only
 * (block numbering as in "Field layout"); records/offsets/workspace are sized
 * for nblocks.  A writer cuts a field into such ranges to overlap the host
 * zlib stage of one range with the device work of the next. */
int czw_compress_field_range(const void *field, int nx, int ny, int nz, int prec, float eps, int wtype,
                             int zerobits, int shuffle, int block0, int nblocks, void *records, uint64_t *offsets,
                             void *workspace, hipStream_t stream);

/* The two stages of czw_compress_field, for callers that time the stages
 * apart or keep fixed-size per-block slots.  Stage 1 writes block b's mask and
 * survivors (zerobits applied, not yet shuffled) into slots +
 * b*czw_slot_bytes(prec) in an internal region layout, and sets nbytes[b] to
 * the payload size followed by internal region counts: nbytes has 2*nblocks
 * entries for prec 4 and 9*nblocks for prec 8.  Stage 2 frames the payloads
 * into the contiguous record stream, applying the byte shuffle when
 * shuffle != 0; slot_bytes must be czw_slot_bytes(prec). */
size_t czw_slot_bytes(int prec);
int czw_compress_field_slots(const void *field, int nx, int ny, int nz, int prec, float eps, int wtype,
                             int zerobits, void *slots, uint32_t *nbytes, hipStream_t stream);
int czw_pack_records(const void *slots, size_t slot_bytes, const uint32_t *nbytes, int nblocks, int shuffle,
                     void *records, uint64_t *offsets, hipStream_t stream);

/* Waits for `stream` and returns CZW_EDEVICE if the last compress on this
 * workspace (or lowmem scratch) reported an error (its error word is
 * non-zero), else CZW_OK. */
int czw_compress_status(const void *workspace, hipStream_t stream);

/* Re-reads of a not-ready predecessor's status before a block of
 * czw_compress_field_lowmem gives up (0 restores the default, 2^20, about a
 * second).  Returns the previous value.  Applies to calls made after it,
 * process-wide. */
uint32_t czw_set_lookback_limit(uint32_t rereads);

/* Same as czw_compress_field for blocks already laid out [nb][z][y][x]. */
int czw_compress_blocks(const void *blocks, int nblocks, int prec, float eps, int wtype, int zerobits,
                        int shuffle, void *records, uint64_t *offsets, void *workspace,
                        hipStream_t stream);

/* Replaces WaveletCompressorGeneric::decompress (WaveletCompressor.cpp:232-259)
 * plus copy_to (WaveletCompressor.h:73-82) for every record, writing an
 * HDF5-ordered field.  offsets[b] points at record b's int32 size word; the
 * records need not be contiguous, but every record must lie inside
 * [records, records + records_bytes).  *err (device) is OR-ed with CZW_ERR_*
 * for corrupt or out-of-range records, whose blocks are left unwritten. */
int czw_decompress_field(const void *records, size_t records_bytes, const uint64_t *offsets, int nx,
                         int ny, int nz, int prec, int wtype, int shuffle, void *field, uint32_t *err,
                         hipStream_t stream);

/* Same as czw_decompress_field writing blocks [nb][z][y][x]. */
int czw_decompress_blocks(const void *records, size_t records_bytes, const uint64_t *offsets,
                          int nblocks, int prec, int wtype, int shuffle, void *blocks, uint32_t *err,
                          hipStream_t stream);

/* ---------------------------------------------------------------------
 * Batched mode: many fields of different shapes and thresholds in one
 * persistent launch per stage.  No counterpart call exists in the reference,
 * whose writer compresses one field per process (hdf2cz); a batch is the
 * concatenation of what per-field calls would produce.
 *
 * Blocks are numbered batch-wide: field f owns [block0, block0 + nblocks),
 * each field in its own block order (see "Field layout").  The records of a
 * batch are the per-field record streams back to back, so field f's record
 * stream is records[offsets[block0] .. offsets[block0 + nblocks]) with
 * offsets rebased by offsets[block0].
 */
typedef struct czw_field {
    const void *src; /* compress: the field, HDF5 order [nx][ny][nz], device */
    void *dst;       /* decompress: output field, same layout, device */
    int32_t nx, ny, nz; /* multiples of 32 */
    float eps;       /* this field's threshold */
    int32_t block0;  /* set by czw_batch_prepare */
    int32_t nblocks; /* set by czw_batch_prepare */
} czw_field_t;

#define CZW_BATCH_MAX_FIELDS 64

/* Launch shape of one stage of a batch. */
typedef struct czw_shape {
    int groups; /* persistent workgroups; 0 = as many as are resident (never more) */
    int queue;  /* 1: dynamic block queue; 0: static stride */
    int tile;   /* binary32 compress stage: CZW_TILE_BLOCK or CZW_TILE_HALF;
                   ignored elsewhere (binary64 compresses, and every stage
                   decompresses, with the half-block tile) */
} czw_shape_t;

/* LDS tile and sweep schedule of the binary32 batch compressor.
 * CZW_TILE_BLOCK: a whole 32^3 block in LDS per workgroup (one workgroup per
 *   CU), all level-32 passes on it, the next block's rows prefetched into
 *   registers meanwhile.
 * CZW_TILE_HALF: one x-half of a block in LDS at a time (two workgroups per
 *   CU), the x >= 16 half's level-32 passes and survivors first, then the
 *   x < 16 half's and the [0,16)^3 corner's; no prefetch, the other
 *   workgroup on the CU fills the wait for the loads.
 * Both give the same records. */
#define CZW_TILE_BLOCK 0
#define CZW_TILE_HALF 1

typedef struct czw_batch {
    const czw_field_t *fields; /* DEVICE copy of the prepared table */
    int nfields;               /* 1 .. CZW_BATCH_MAX_FIELDS */
    int nblocks;               /* total, from czw_batch_prepare */
    int prec, wtype, zerobits, shuffle;
    void *workspace;           /* czw_batch_workspace_bytes(nblocks, prec) */
    void *records;             /* czw_records_bound(nblocks, prec) bytes */
    uint64_t *offsets;         /* nblocks + 1 */
    size_t records_bytes;      /* size of the records buffer (decompress bound) */
    uint32_t *err;             /* decompress error bits; cleared by each decompress */
    czw_shape_t compress;      /* launch shape of the compress stage */
    czw_shape_t decompress;    /* launch shape of the decompress stage */
} czw_batch_t;

/* Host: validate a table for Real of prec bytes, fill block0/nblocks, return
 * the total in *nblocks.  Batch fields need 32*ny*nz*prec < 2^31 (the
 * per-field entry points have no such limit). */
int czw_batch_prepare(czw_field_t *fields, int nfields, int prec, int *nblocks);

/* Scratch of a batch: the compressor workspace of its nblocks (header word 0
 * the compress queue, word 1 the error bits, word 2 the decompress queue),
 * with room for either tile's per-block size words. */
size_t czw_batch_workspace_bytes(int nblocks, int prec);

/* Compress every field (fwt, thresholds, survivor streams, records). */
int czw_batch_compress(const czw_batch_t *batch, hipStream_t stream);

/* Decompress every field from batch->records into each field's dst. */
int czw_batch_decompress(const czw_batch_t *batch, hipStream_t stream);

/* Time the launch shapes of one stage (1 compress, 2 decompress) on this
 * batch -- all or half of the stage kernel's resident workgroups x block
 * queue or static order, and for the binary32 compress stage x both tiles
 * (CZW_TILE_BLOCK, CZW_TILE_HALF) -- and store the fastest in that stage's
 * shape only; *best_ms receives its time.  Runs the stage several times on
 * `stream`. */
int czw_batch_autotune(czw_batch_t *batch, int stage, hipStream_t stream, float *best_ms);

/* Capture the stages (bit 0 compress, bit 1 decompress) of a batch into an
 * executable HIP graph; launch it with czw_graph_launch. */
int czw_batch_graph(const czw_batch_t *batch, int stages, void **graph);
int czw_graph_launch(void *graph, hipStream_t stream);
void czw_graph_destroy(void *graph);

/* Whole-block multilevel transform, in place, on nblocks blocks
 * [nb][32][32][32] kept in their original orientation [z][y][x] (the
 * coefficients are not permuted into the reference's stored layout).
 * op 0: forward (FullTransformEngine::fwt, FullWaveletTransform.h:44-49),
 * op 1: inverse (iwt, :51-56), op 2: adjoint (transpose) of the inverse, the
 * backward of the differentiable inverse.  Binary32 ops 0/1 are bit-identical
 * to the reference; op 2 is the exact transpose of the linear inverse.
 * No counterpart call exists in the reference for op 2. */
int czw_transform_blocks(void *blocks, int nblocks, int prec, int wtype, int op, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CZW_H */
