/*
 * cz_host.h -- C ABI of the host-side .cz container library (libczhost.so).
 *
 * Replaces the reference's serializer and reader around the block compressor:
 *  - czh_write: the chunking/zlib/LUT/header/metadata part of
 *    SerializerIO_WaveletCompression_MPI_SimpleBlocking::_write
 *    (Compressor/source/SerializerIO_WaveletCompression_MPI_Simple.h:101-184,
 *    489-591) and _to_file (:347-466) for one writer;
 *  - czh_open/czh_info: Reader_WaveletCompression::load_file
 *    (Compressor/source/Reader_WaveletCompression.h:255-560);
 *  - czh_gather: the inflate + record lookup of load_block2 (:635-698), done
 *    once per chunk for the whole file instead of once per block.
 * All buffers here are HOST memory (czh_compress_device_field excepted: it
 * takes a device field).  The device side is include/czw.h.
 */
#ifndef CZ_HOST_H
#define CZ_HOST_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { CZH_OK = 0, CZH_EINVAL = -1, CZH_EIO = -2, CZH_EZLIB = -3, CZH_ECORRUPT = -4 };
enum { CZH_ENCODER_NONE = 0, CZH_ENCODER_ZLIB = 1, CZH_ENCODER_LZ4 = 2 };

typedef struct czh_file czh_file;

/* Blocks per chunk of one writer thread: 15 for prec 4, 7 for prec 8. */
int czh_entries_per_chunk(int prec);

/* Write a single-subdomain .cz file from a host record stream in block order
 * (block b = ix + bpd[0]*(iy + bpd[1]*iz)).  wtype < 0 records "Wavelets: none"
 * (raw blocks).  threshold is printed as the Real of the given precision.
 * nthreads <= 0 keeps the OpenMP default for the per-chunk deflate.
 * records holds records_bytes bytes; CZH_EINVAL unless every record lies in
 * it (offsets non-decreasing, offsets[nblocks] <= records_bytes) and its size
 * word fits between its offset and the next. */
int czh_write(const char *path, int prec, const int bpd[3], int wtype, double threshold, int encoder,
              const uint8_t *records, size_t records_bytes, const uint64_t *offsets, int nblocks, int nthreads);

/* The file that nranks = prod(bpd[d] / sub_bpd[d]) writer ranks, each holding
 * a subdomain of sub_bpd blocks, write together (the MPI writer of
 * SerializerIO_WaveletCompression_MPI_Simple.h:347-466 with MPI_Exscan'd
 * offsets; Tools/Test_IO_Compressed.h:236-291 with -nprocx/y/z).  The records
 * come in block order of the whole grid, as for czh_write; ranks are placed
 * in MPI's Cartesian order (last coordinate fastest) and each one chunks its
 * own blocks, x fastest.  sub_bpd == bpd is czh_write. */
int czh_write_subdomains(const char *path, int prec, const int bpd[3], const int sub_bpd[3], int wtype,
                         double threshold, int encoder, const uint8_t *records, size_t records_bytes,
                         const uint64_t *offsets, int nblocks, int nthreads);

/* Read and parse a .cz file (any number of subdomains).  Returns NULL on
 * error with a message in errbuf. */
czh_file *czh_open(const char *path, char *errbuf, size_t errlen);

/* czh_open with flags: CZH_OPEN_SWAP reads a file written with the other byte
 * order, as the reference tools' -swap (Reader_WaveletCompression.h:133-225):
 * header displacement, metadata, LUTs, record size words and the records'
 * 4-byte units from byte 4096 on are byte-reversed. */
#define CZH_OPEN_SWAP 1
czh_file *czh_open_ex(const char *path, int flags, char *errbuf, size_t errlen);
void czh_close(czh_file *f);
int czh_info(const czh_file *f, int *prec, int bpd[3], int sub_bpd[3], int *wtype, double *threshold, int *encoder);
uint64_t czh_file_bytes(const czh_file *f);

/* Inflate every chunk once; *records (malloc'ed, free with czh_free) holds
 * the records of all blocks and (*offsets)[b] the position of block b's int32
 * size word; *offsets has nblocks+1 entries. */
int czh_gather(czh_file *f, int nthreads, uint8_t **records, uint64_t *records_bytes, uint64_t **offsets);
void czh_free(void *p);

/* The writer's per-chunk deflate (zlib level Z_DEFAULT_COMPRESSION, Z_FINISH). */
int czh_deflate(const uint8_t *in, uint64_t n, uint8_t *out, uint64_t cap, uint64_t *produced);

/* The writer's LZ4 stage on one chunk: the greedy parse of the LZ4_compress
 * the reference vendors (ThirdParty/lz4/lz4.c, called at
 * CompressionEncoders.h:230-231), and the block decoder its reader uses
 * (LZ4_uncompress_unknownOutputSize, CompressionEncoders.h:75-76). */
int czh_lz4_compress(const uint8_t *in, uint64_t n, uint8_t *out, uint64_t cap, uint64_t *produced);
int czh_lz4_decompress(const uint8_t *in, uint64_t n, uint8_t *out, uint64_t cap, uint64_t *produced);

/* Compress an HDF5-ordered DEVICE field [nx][ny][nz] of Real (prec bytes)
 * straight to a single-writer .cz file (path may be NULL: only the size is
 * reported).  The field is cut into slabs of whole chunks of about
 * slab_blocks blocks (<= 0: 1024).  With overlap != 0 the device compresses
 * slab k+1 (czw_compress_field_range) while nthreads host threads run the
 * second stage (encoder: zlib, LZ4 or none) on slab k; with overlap == 0 the
 * two alternate.  The bytes equal czh_write of the whole field's records.
 * times (may be NULL) receives 5 values in seconds: wall, device compress,
 * host second stage, device-to-host copies, and writing the file (streamed
 * chunk by chunk; with a NULL path only its size is summed).  Allocates its
 * own device and pinned host buffers. */
int czh_compress_device_field(const char *path, const void *field, int nx, int ny, int nz, int prec, float eps,
                              int wtype, int zerobits, int shuffle, int encoder, int nthreads, int slab_blocks,
                              int overlap, hipStream_t stream, uint64_t *file_bytes, double *times);

#ifdef __cplusplus
}
#endif

#endif /* CZ_HOST_H */
