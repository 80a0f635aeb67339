// cz_file.hpp -- host side of the .cz container: chunked zlib stage, ASCII
// header, block metadata and chunk look-up table.
//
// File layout (SerializerIO_WaveletCompression_MPI_Simple.h:347-466, 489-591):
//   @0    uint64 header displacement
//   @8    "\n==============START-BINARY-OCEAN==============\n"      (48 bytes)
//   @56   per subdomain (writer rank), in rank order: its chunks, then
//         uint64 lut[nchunks] holding each chunk's start relative to the
//         subdomain's first byte
//   @hd   ASCII header (SerializerIO_..._MPI_Simple.h:515-553)
//         BlockMetadata[NBLOCKS], subdomain by subdomain          (20 bytes each)
//         "\n==============START-BINARY-LUT==============\n"       (46 bytes)
//         HeaderLUT[SUBDOMAINS]                                  (12 bytes each)
// A chunk is zlib (RFC 1950, level 6) of consecutive block records
// [int32 nbytes][payload]; one writer thread puts ENTRIES consecutive blocks
// in a chunk, ENTRIES = 4 MiB / (sizeof(WaveletCompressor) + 4) = 15 (float)
// or 7 (double) (SerializerIO_..._MPI_Simple.h:68-76, 334-339).
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace cz {

#pragma pack(push, 1)
// WaveletSerializationTypes.h:16-18 (packed on-disk records)
struct BlockMetadata {
    int32_t idcompression, subid, ix, iy, iz;
};
struct HeaderLUT {
    uint64_t aggregate_bytes;
    int32_t nchunks;
};
struct CompressedBlock {
    uint64_t start, extent;
    int32_t subid;
};
#pragma pack(pop)
static_assert(sizeof(BlockMetadata) == 20, "BlockMetadata must be 20 bytes");
static_assert(sizeof(HeaderLUT) == 12, "HeaderLUT must be 12 bytes");
static_assert(sizeof(CompressedBlock) == 20, "CompressedBlock must be 20 bytes");

extern const char *const OCEAN_TITLE;  // 48 bytes
extern const char *const LUT_TITLE;    // 46 bytes

enum Encoder { ENC_NONE = 0, ENC_ZLIB = 1, ENC_LZ4 = 2 };

// Name recorded for a wavelet type (FullWaveletTransform.h:28-35); wtype < 0
// means no wavelet stage (raw blocks, the reference's "default" build).
const char *wavelet_name(int wtype);
int wavelet_type(const std::string &name);  // inverse of wavelet_name, -1 for "none"

// Blocks per chunk of one writer thread for sizeof(Real) = prec.
int entries_per_chunk(int prec);

struct Params {
    int prec = 4;                  // sizeof(Real)
    int bpd[3] = {1, 1, 1};        // blocks per dimension (global)
    int sub_bpd[3] = {1, 1, 1};    // blocks per dimension of one subdomain
    int wtype = 3;                 // < 0: raw blocks ("Wavelets: none")
    double threshold = 0.0;        // printed as Real
    int encoder = ENC_ZLIB;
    bool halffloat = false;
};

// The ASCII header text exactly as the reference writes it.
std::string make_header(const Params &p);

// Deflate one chunk the way the reference writer does (deflateInit with
// Z_DEFAULT_COMPRESSION, Z_FINISH).  Returns false on a zlib error.
bool deflate_chunk(const uint8_t *in, size_t n, std::vector<uint8_t> &out);
// Inflate a whole chunk; returns false unless inflate ends with Z_STREAM_END.
bool inflate_chunk(const uint8_t *in, size_t n, std::vector<uint8_t> &out);
// The LZ4 stage (CompressionEncoders.h:192-238 with the vendored LZ4 of
// ThirdParty/lz4): the same greedy parse as its LZ4_compress, and a block
// decoder bounded by max_out (cz_lz4.cpp).
void lz4_compress(const uint8_t *src, size_t n, std::vector<uint8_t> &dst);
bool lz4_decompress(const uint8_t *src, size_t n, std::vector<uint8_t> &dst, size_t max_out);
// Bytes of one writer chunk's records at most (ENTRIES records of a full block).
size_t chunk_bound(int prec);

// bpd and sub_bpd are positive and each sub_bpd[d] divides bpd[d].
bool subdomains_valid(const Params &p);
// The file order of the blocks when p.bpd is cut into subdomains of
// p.sub_bpd blocks, one writer rank each: order[q] is the block id
// (ix + bx*(iy + by*iz)) of the q-th block in the file.  Ranks follow MPI's
// Cartesian order (last coordinate fastest); each rank's blocks are x
// fastest.  The identity for a single subdomain.
void file_block_order(const Params &p, std::vector<int> &order);

// Serialize a record stream in block order (x fastest over the whole grid)
// into a complete .cz image, cut into subdomains as p.sub_bpd says (one
// writer rank each).  records/offsets as in czw.h.
int build_cz(const Params &p, const uint8_t *records, const uint64_t *offsets, int nblocks, int nthreads,
             std::vector<uint8_t> &file);
// The two halves of build_cz: the chunk stage (the second stage per chunk of
// ENTRIES records, OpenMP over chunks) and the file image around the finished
// chunks.  encode_chunks takes the blocks in file order: with order == nullptr
// the records are already in that order (offsets relative to records);
// otherwise block q of the file is records[offsets[order[q]] ...].  Each run
// of per_subdomain blocks (0: all) is one subdomain, whose last chunk may be
// short.  assemble_cz takes the chunks subdomain by subdomain.
int encode_chunks(const Params &p, const uint8_t *records, const uint64_t *offsets, int nblocks, int nthreads,
                  std::vector<std::vector<uint8_t>> &chunks, const int *order = nullptr, int per_subdomain = 0);
int assemble_cz(const Params &p, const std::vector<std::vector<uint8_t>> &chunks, int nblocks,
                std::vector<uint8_t> &file);

// Where a CzWriter puts the file: a memory image, an open FILE (binary,
// positioned at its start), or nowhere (only the size is counted).
class CzSink {
  public:
    CzSink() = default;
    explicit CzSink(std::vector<uint8_t> *mem) : mem_(mem) {}
    explicit CzSink(FILE *f) : file_(f) {}
    bool put(const void *d, size_t n);
    bool patch_head(uint64_t v);  // bytes [0, 8)
    uint64_t bytes() const { return bytes_; }

  private:
    std::vector<uint8_t> *mem_ = nullptr;
    FILE *file_ = nullptr;
    uint64_t bytes_ = 0;
};

// A .cz file written front to back as its chunks become available, so that
// only the chunks not yet written need to be held: begin(), then chunk() for
// every chunk in file order (subdomain by subdomain, as encode_chunks cuts
// them), then finish().  Each returns 0, -1 on misuse or bad parameters, or
// -3 when the sink fails.
class CzWriter {
  public:
    CzWriter(const Params &p, int nblocks, CzSink &sink);
    bool valid() const { return ok_; }
    int nchunks() const { return ok_ ? nsub_ * ncs_ : 0; }
    int begin();
    int chunk(const uint8_t *d, size_t n);
    int finish();

  private:
    Params p_;
    int nblocks_, E_ = 1, nsub_ = 0, ncs_ = 0, nc_ = 0;
    bool ok_ = false, started_ = false;
    CzSink &sink_;
    std::vector<uint64_t> lut_;
    std::vector<HeaderLUT> hls_;
    uint64_t sub_bytes_ = 0, ocean_bytes_ = 0;
};

struct FileInfo {
    Params p;
    int nblocks = 0;
    uint64_t header_displacement = 0;
    std::vector<CompressedBlock> idx2chunk;  // per block id ix + bx*(iy + by*iz)
    std::vector<uint64_t> chunk_start;       // absolute file offsets
    std::vector<uint64_t> chunk_size;
    bool swap = false;                       // the file's binary fields are byte-swapped
};

// Parse header, metadata and look-up tables (Reader_WaveletCompression.h:255-560).
// Returns 0 or a negative error; *msg receives a description.  swap: the
// file was written with the other byte order (the tools' -swap, Reader_
// WaveletCompression.h:133-225,280,421,449,478): the header displacement,
// metadata, HeaderLUTs and chunk offsets are byte-reversed on reading, and
// gather_records reverses each record's size word and, in 4-byte units, its
// bytes from 4096 on (:605-620; raw blocks included, as the reference does).
int parse_cz(const std::vector<uint8_t> &file, FileInfo &info, std::string *msg, bool swap = false);

// Inflate every chunk once and gather all block records in block order:
// records receives the concatenated inflated chunks and offsets[b] the
// position of block b's int32 size word (offsets has nblocks+1 entries, the
// last = records.size()).
int gather_records(const std::vector<uint8_t> &file, const FileInfo &info, int nthreads, std::vector<uint8_t> &records,
                   std::vector<uint64_t> &offsets, std::string *msg);

bool read_file(const std::string &path, std::vector<uint8_t> &data);
bool write_file(const std::string &path, const std::vector<uint8_t> &data);

}  // namespace cz
