// cz_file.cpp -- .cz container writer/reader (host side).  See cz_file.hpp.
#include "cz_file.hpp"

#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace cz {

const char *const OCEAN_TITLE = "\n==============START-BINARY-OCEAN==============\n";
const char *const LUT_TITLE = "\n==============START-BINARY-LUT==============\n";
static const char *const HEADER_TITLE = "\n==============START-ASCI-HEADER==============\n";
static const char *const META_TITLE = "==============START-BINARY-METABLOCKS==============\n";

const char *wavelet_name(int wtype)
{
    switch (wtype) {
    case 0: return "None";
    case 1: return "InterpWavelet4thOrder";
    case 2: return "LiftedInterpWavelet4thOrder";
    case 3: return "AverageInterpWavelet3rdOrder";
    default: return wtype < 0 ? "none" : "None";
    }
}

int wavelet_type(const std::string &name)
{
    for (int w = 0; w <= 3; ++w)
        if (name == wavelet_name(w)) return w;
    return -1;
}

size_t chunk_bound(int prec)
{
    return (size_t)entries_per_chunk(prec) * (4 + 4096 + (size_t)prec * 32768);
}

int entries_per_chunk(int prec)
{
    // sizeof(WaveletCompressorGeneric<32,Real>) by layout arithmetic
    // (WaveletCompressor.h:35-55): vtable pointer (8), FullTransform<32> =
    // its 1-byte engine base padded to alignof(Real) + the 32^3 work array,
    // the BUFMAXSIZE byte buffer, then the 8-byte bufsize at an 8-aligned
    // offset: 266264 bytes for float, 528408 for double.  ENTRIES =
    // 4 MiB / (sizeof + sizeof(int)) (SerializerIO_..._MPI_Simple.h:68-76).
    const long nb3 = 32768;
    const long before = 8 + (prec + prec * nb3) + (4096 + prec * nb3);
    const long size = ((before + 7) / 8) * 8 + 8;
    const long e = (4L * 1024 * 1024) / (size + 4);
    return e > 0 ? (int)e : 1;
}

std::string make_header(const Params &p)
{
    std::stringstream ss;
    ss << HEADER_TITLE;
    ss << "Endianess: " << "little" << "\n";
    ss << "sizeofReal: " << p.prec << "\n";
    ss << "sizeofsize_t: " << sizeof(size_t) << "\n";
    ss << "sizeofBlockMetadata: " << sizeof(BlockMetadata) << "\n";
    ss << "sizeofHeaderLUT: " << sizeof(HeaderLUT) << "\n";
    ss << "sizeofCompressedBlock: " << sizeof(CompressedBlock) << "\n";
    ss << "Blocksize: " << 32 << "\n";
    ss << "Blocks: " << p.bpd[0] << " x " << p.bpd[1] << " x " << p.bpd[2] << "\n";
    // Extent = h * blocks * 32 with h = 1 / max(blocks * 32) (Cubism GridMPI.h:82-85)
    const double h = 1.0 / (double)std::max(p.bpd[0] * 32, std::max(p.bpd[1] * 32, p.bpd[2] * 32));
    ss << "Extent: " << h * p.bpd[0] * 32 << " " << h * p.bpd[1] * 32 << " " << h * p.bpd[2] * 32 << "\n";
    ss << "SubdomainBlocks: " << p.sub_bpd[0] << " x " << p.sub_bpd[1] << " x " << p.sub_bpd[2] << "\n";
    ss << "HalfFloat: " << (p.halffloat ? "yes" : "no") << "\n";
    ss << "Wavelets: " << wavelet_name(p.wtype) << "\n";
    // the threshold member has type Real
    if (p.prec == 4) ss << "WaveletThreshold: " << (float)p.threshold << "\n";
    else ss << "WaveletThreshold: " << p.threshold << "\n";
    ss << "Encoder: " << (p.encoder == ENC_ZLIB ? "zlib" : p.encoder == ENC_LZ4 ? "lz4" : "none") << "\n";
    ss << META_TITLE;
    return ss.str();
}

bool deflate_chunk(const uint8_t *in, size_t n, std::vector<uint8_t> &out)
{
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (deflateInit(&zs, Z_DEFAULT_COMPRESSION) != Z_OK) return false;
    out.resize(deflateBound(&zs, (uLong)n) + 16);
    zs.next_in = const_cast<Bytef *>(in);
    zs.avail_in = (uInt)n;
    zs.next_out = out.data();
    zs.avail_out = (uInt)out.size();
    const int rc = deflate(&zs, Z_FINISH);
    const size_t produced = zs.total_out;
    deflateEnd(&zs);
    if (rc != Z_STREAM_END) return false;
    out.resize(produced);
    return true;
}

bool inflate_chunk(const uint8_t *in, size_t n, std::vector<uint8_t> &out)
{
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return false;
    zs.next_in = const_cast<Bytef *>(in);
    zs.avail_in = (uInt)n;
    size_t cap = std::max<size_t>(n * 4, 1 << 16);
    out.resize(cap);
    int rc = Z_OK;
    while (true) {
        zs.next_out = out.data() + zs.total_out;
        zs.avail_out = (uInt)(out.size() - zs.total_out);
        rc = inflate(&zs, Z_FINISH);
        if (rc == Z_STREAM_END) break;
        if (rc != Z_BUF_ERROR && rc != Z_OK) break;
        if (zs.avail_out != 0 && zs.avail_in == 0) {  // truncated stream
            rc = Z_DATA_ERROR;
            break;
        }
        out.resize(out.size() * 2);
    }
    const size_t produced = zs.total_out;
    inflateEnd(&zs);
    if (rc != Z_STREAM_END) return false;
    out.resize(produced);
    return true;
}

bool subdomains_valid(const Params &p)
{
    for (int d = 0; d < 3; ++d)
        if (p.bpd[d] <= 0 || p.sub_bpd[d] <= 0 || p.bpd[d] % p.sub_bpd[d]) return false;
    return true;
}

void file_block_order(const Params &p, std::vector<int> &order)
{
    const int sx = p.sub_bpd[0], sy = p.sub_bpd[1], sz = p.sub_bpd[2];
    const int py = p.bpd[1] / sy, pz = p.bpd[2] / sz;
    const int bps = sx * sy * sz, nb = p.bpd[0] * p.bpd[1] * p.bpd[2];
    order.resize(nb);
    for (int q = 0; q < nb; ++q) {
        // rank q / bps at Cartesian coordinates (cx, cy, cz), the last one
        // fastest (MPI_Cart_create, GridMPI.h:73-75); its blocks x fastest
        // (Grid.h:80-85), offset by coordinates * subdomain blocks (GridMPI.h:111-114)
        const int r = q / bps, i = q % bps;
        const int cz = r % pz, cy = (r / pz) % py, cx = r / (py * pz);
        const int gx = cx * sx + i % sx, gy = cy * sy + (i / sx) % sy, gz = cz * sz + i / (sx * sy);
        order[q] = gx + p.bpd[0] * (gy + p.bpd[1] * gz);
    }
}

int encode_chunks(const Params &p, const uint8_t *records, const uint64_t *offsets, int nblocks, int nthreads,
                  std::vector<std::vector<uint8_t>> &chunks, const int *order, int per_subdomain)
{
    if (nblocks <= 0 || (p.encoder != ENC_NONE && p.encoder != ENC_ZLIB && p.encoder != ENC_LZ4)) return -1;
    const int E = entries_per_chunk(p.prec);
    const int bps = per_subdomain > 0 ? per_subdomain : nblocks;
    if (nblocks % bps) return -1;
    const int ncs = (bps + E - 1) / E;  // chunks per subdomain
    const int nchunks = ncs * (nblocks / bps);
    chunks.assign(nchunks, std::vector<uint8_t>());
    int bad = 0;
#ifdef _OPENMP
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad) num_threads(nt)
#endif
    for (int c = 0; c < nchunks; ++c) {
        // file positions [q0, q1) of this chunk's blocks
        const int q0 = (c / ncs) * bps + (c % ncs) * E;
        const int q1 = (c / ncs) * bps + std::min((c % ncs + 1) * E, bps);
        const uint8_t *src;
        size_t len;
        std::vector<uint8_t> gathered;
        if (!order) {
            src = records + offsets[q0];
            len = offsets[q1] - offsets[q0];
        } else {  // the records of one chunk are not adjacent in block order
            for (int q = q0; q < q1; ++q)
                gathered.insert(gathered.end(), records + offsets[order[q]], records + offsets[order[q] + 1]);
            src = gathered.data();
            len = gathered.size();
        }
        if (p.encoder == ENC_ZLIB) {
            if (!deflate_chunk(src, len, chunks[c])) bad = 1;
        } else if (p.encoder == ENC_LZ4) {
            lz4_compress(src, len, chunks[c]);
        } else {
            chunks[c].assign(src, src + len);
        }
    }
#ifndef _OPENMP
    (void)nthreads;
#endif
    return bad ? -2 : 0;
}

bool CzSink::put(const void *d, size_t n)
{
    bytes_ += n;
    if (mem_) {
        const uint8_t *q = static_cast<const uint8_t *>(d);
        mem_->insert(mem_->end(), q, q + n);
    } else if (file_) {
        return std::fwrite(d, 1, n, file_) == n;
    }
    return true;
}

bool CzSink::patch_head(uint64_t v)
{
    if (mem_) {
        if (mem_->size() < sizeof(v)) return false;
        std::memcpy(mem_->data(), &v, sizeof(v));
    } else if (file_) {
        return std::fseek(file_, 0, SEEK_SET) == 0 && std::fwrite(&v, 1, sizeof(v), file_) == sizeof(v) &&
               std::fseek(file_, 0, SEEK_END) == 0;
    }
    return true;
}

// The layout: uint64 header displacement, the ocean title; per subdomain its
// chunks, then uint64 lut[ncs] of chunk starts relative to its first byte,
// so that HeaderLUT = {bytes incl. the lut, ncs}
// (SerializerIO_..._MPI_Simple.h:563-575), subdomains in rank order
// (MPI_Exscan of the byte counts, :385-401); then the ASCII header, the block
// metadata, the LUT title and the HeaderLUTs.  The displacement is only known
// at the end, so a zero goes first and finish() patches it.
CzWriter::CzWriter(const Params &p, int nblocks, CzSink &sink) : p_(p), nblocks_(nblocks), sink_(sink)
{
    E_ = entries_per_chunk(p.prec);
    ok_ = subdomains_valid(p) && nblocks > 0 && p.bpd[0] * p.bpd[1] * p.bpd[2] == nblocks;
    if (!ok_) return;
    const int bps = p.sub_bpd[0] * p.sub_bpd[1] * p.sub_bpd[2];
    nsub_ = nblocks / bps;
    ncs_ = (bps + E_ - 1) / E_;
    lut_.resize(ncs_);
}

int CzWriter::begin()
{
    if (!ok_ || started_) return -1;
    started_ = true;
    const uint64_t zero = 0;
    return sink_.put(&zero, sizeof(zero)) && sink_.put(OCEAN_TITLE, std::strlen(OCEAN_TITLE)) ? 0 : -3;
}

int CzWriter::chunk(const uint8_t *d, size_t n)
{
    if (!ok_ || !started_ || (int)hls_.size() == nsub_) return -1;
    lut_[nc_] = sub_bytes_;
    sub_bytes_ += n;
    if (!sink_.put(d, n)) return -3;
    if (++nc_ == ncs_) {  // the subdomain's last chunk: its lut follows
        if (!sink_.put(lut_.data(), sizeof(uint64_t) * ncs_)) return -3;
        hls_.push_back(HeaderLUT{sub_bytes_ + sizeof(uint64_t) * ncs_, ncs_});
        ocean_bytes_ += hls_.back().aggregate_bytes;
        nc_ = 0;
        sub_bytes_ = 0;
    }
    return 0;
}

int CzWriter::finish()
{
    if (!ok_ || !started_ || (int)hls_.size() != nsub_) return -1;
    const std::string header = make_header(p_);
    // per subdomain, BlockMetadata {chunk, index in chunk, global ix, iy, iz}
    // of its blocks in its own order (SerializerIO_..._MPI_Simple.h:169-177,
    // 328-332; written at rank * NBLOCKS, :431-441)
    std::vector<int> order;
    file_block_order(p_, order);
    const int bps = nblocks_ / nsub_, bx = p_.bpd[0], by = p_.bpd[1];
    std::vector<BlockMetadata> meta(nblocks_);
    for (int q = 0; q < nblocks_; ++q) {
        const int i = q % bps, g = order[q];
        meta[q] = BlockMetadata{i / E_, i % E_, g % bx, (g / bx) % by, g / (bx * by)};
    }
    const uint64_t header_disp = sizeof(uint64_t) + std::strlen(OCEAN_TITLE) + ocean_bytes_;
    const bool ok = sink_.put(header.data(), header.size()) &&
                    sink_.put(meta.data(), sizeof(BlockMetadata) * (size_t)nblocks_) &&
                    sink_.put(LUT_TITLE, std::strlen(LUT_TITLE)) &&
                    sink_.put(hls_.data(), sizeof(HeaderLUT) * (size_t)nsub_) && sink_.patch_head(header_disp);
    return ok ? 0 : -3;
}

int assemble_cz(const Params &p, const std::vector<std::vector<uint8_t>> &chunks, int nblocks,
                std::vector<uint8_t> &file)
{
    file.clear();
    CzSink sink(&file);
    CzWriter w(p, nblocks, sink);
    if (!w.valid() || (int)chunks.size() != w.nchunks()) return -1;
    int rc = w.begin();
    for (size_t c = 0; rc == 0 && c < chunks.size(); ++c) rc = w.chunk(chunks[c].data(), chunks[c].size());
    return rc ? rc : w.finish();
}

int build_cz(const Params &p, const uint8_t *records, const uint64_t *offsets, int nblocks, int nthreads,
             std::vector<uint8_t> &file)
{
    if (!subdomains_valid(p) || nblocks <= 0 || p.bpd[0] * p.bpd[1] * p.bpd[2] != nblocks) return -1;
    const int bps = p.sub_bpd[0] * p.sub_bpd[1] * p.sub_bpd[2];
    std::vector<int> order;
    if (bps != nblocks) file_block_order(p, order);
    std::vector<std::vector<uint8_t>> chunks;
    const int rc = encode_chunks(p, records, offsets, nblocks, nthreads, chunks, order.empty() ? nullptr : order.data(),
                                 bps);
    if (rc) return rc;
    return assemble_cz(p, chunks, nblocks, file);
}

static bool parse_kv(const std::string &line, const char *key, std::string &val)
{
    const size_t k = std::strlen(key);
    if (line.compare(0, k, key) != 0) return false;
    size_t i = k;
    while (i < line.size() && line[i] == ' ') ++i;
    val = line.substr(i);
    return true;
}

static uint32_t bswap_u32(uint32_t v) { return __builtin_bswap32(v); }
static uint64_t bswap_u64(uint64_t v) { return __builtin_bswap64(v); }

int parse_cz(const std::vector<uint8_t> &file, FileInfo &info, std::string *msg, bool swap)
{
    auto fail = [&](const char *m) {
        if (msg) *msg = m;
        return -1;
    };
    auto fix32 = [&](int32_t v) { return swap ? (int32_t)bswap_u32((uint32_t)v) : v; };
    auto fix64 = [&](uint64_t v) { return swap ? bswap_u64(v) : v; };
    info.swap = swap;
    const size_t ocean = sizeof(uint64_t) + std::strlen(OCEAN_TITLE);
    if (file.size() < ocean) return fail("file too short");
    uint64_t hd;
    std::memcpy(&hd, file.data(), sizeof(hd));
    hd = fix64(hd);
    if (hd < ocean || hd >= file.size()) return fail("bad header displacement");
    if (std::memcmp(file.data() + 8, OCEAN_TITLE, std::strlen(OCEAN_TITLE)) != 0) return fail("missing ocean title");
    info.header_displacement = hd;
    const std::string tail(reinterpret_cast<const char *>(file.data() + hd), file.size() - hd);
    if (tail.compare(0, std::strlen(HEADER_TITLE), HEADER_TITLE) != 0) return fail("missing ASCII header title");
    const size_t meta = tail.find(META_TITLE);
    if (meta == std::string::npos) return fail("missing metablocks title");
    std::istringstream hs(tail.substr(std::strlen(HEADER_TITLE), meta - std::strlen(HEADER_TITLE)));
    std::string line, v;
    Params &p = info.p;
    int sizeof_st = -1, sizeof_bm = -1, sizeof_hl = -1, sizeof_cb = -1, bs = -1;
    std::string wav, enc;
    bool have_blocks = false, have_sub = false;
    while (std::getline(hs, line)) {
        if (parse_kv(line, "Endianess:", v)) {
            if (v != "little" && !swap) return fail("big-endian file: read it with -swap");
        } else if (parse_kv(line, "sizeofReal:", v)) p.prec = std::atoi(v.c_str());
        else if (parse_kv(line, "sizeofsize_t:", v)) sizeof_st = std::atoi(v.c_str());
        else if (parse_kv(line, "sizeofBlockMetadata:", v)) sizeof_bm = std::atoi(v.c_str());
        else if (parse_kv(line, "sizeofHeaderLUT:", v)) sizeof_hl = std::atoi(v.c_str());
        else if (parse_kv(line, "sizeofCompressedBlock:", v)) sizeof_cb = std::atoi(v.c_str());
        else if (parse_kv(line, "Blocksize:", v)) bs = std::atoi(v.c_str());
        else if (parse_kv(line, "Blocks:", v))
            have_blocks = std::sscanf(v.c_str(), "%d x %d x %d", &p.bpd[0], &p.bpd[1], &p.bpd[2]) == 3;
        else if (parse_kv(line, "SubdomainBlocks:", v))
            have_sub = std::sscanf(v.c_str(), "%d x %d x %d", &p.sub_bpd[0], &p.sub_bpd[1], &p.sub_bpd[2]) == 3;
        else if (parse_kv(line, "HalfFloat:", v)) p.halffloat = (v == "yes");
        else if (parse_kv(line, "Wavelets:", v)) wav = v;
        else if (parse_kv(line, "WaveletThreshold:", v)) p.threshold = std::atof(v.c_str());
        else if (parse_kv(line, "Encoder:", v)) enc = v;
    }
    if (p.prec != 4 && p.prec != 8) return fail("sizeofReal must be 4 or 8");
    if (sizeof_st != 8 || sizeof_bm != 20 || sizeof_hl != 12 || sizeof_cb != 20) return fail("record sizes differ");
    if (bs != 32) return fail("Blocksize must be 32");
    if (!have_blocks || !have_sub) return fail("missing block counts");
    for (int d = 0; d < 3; ++d)
        if (p.bpd[d] <= 0 || p.sub_bpd[d] <= 0) return fail("bad block counts");
    p.wtype = (wav == "none") ? -1 : wavelet_type(wav);
    if (wav != "none" && p.wtype < 0) return fail("unknown wavelet type");
    if (enc == "zlib") p.encoder = ENC_ZLIB;
    else if (enc == "lz4") p.encoder = ENC_LZ4;
    else if (enc == "none") p.encoder = ENC_NONE;
    else return fail("unsupported encoder (zlib, lz4 and none are known)");

    const long long nb = (long long)p.bpd[0] * p.bpd[1] * p.bpd[2];
    const long long bps = (long long)p.sub_bpd[0] * p.sub_bpd[1] * p.sub_bpd[2];
    if (nb > (1LL << 31) - 1 || nb % bps) return fail("block counts inconsistent");
    info.nblocks = (int)nb;
    const size_t mstart = hd + meta + std::strlen(META_TITLE);
    const size_t lstart = mstart + (size_t)nb * sizeof(BlockMetadata);
    if (lstart + std::strlen(LUT_TITLE) > file.size()) return fail("truncated metadata");
    if (std::memcmp(file.data() + lstart, LUT_TITLE, std::strlen(LUT_TITLE)) != 0) return fail("missing LUT title");
    const int nsub = (int)(nb / bps);
    const size_t hstart = lstart + std::strlen(LUT_TITLE);
    if (hstart + (size_t)nsub * sizeof(HeaderLUT) > file.size()) return fail("truncated HeaderLUT");

    std::vector<BlockMetadata> mb(nb);
    std::memcpy(mb.data(), file.data() + mstart, (size_t)nb * sizeof(BlockMetadata));
    if (swap)
        for (auto &m : mb) {
            m.idcompression = fix32(m.idcompression);
            m.subid = fix32(m.subid);
            m.ix = fix32(m.ix);
            m.iy = fix32(m.iy);
            m.iz = fix32(m.iz);
        }
    // per-subdomain look-up tables (Reader_WaveletCompression.h:439-522)
    uint64_t base = ocean;
    info.chunk_start.clear();
    info.chunk_size.clear();
    for (int s = 0, cur = 0; s < nsub; ++s) {
        HeaderLUT hl;
        std::memcpy(&hl, file.data() + hstart + (size_t)s * sizeof(HeaderLUT), sizeof(hl));
        hl.aggregate_bytes = fix64(hl.aggregate_bytes);
        hl.nchunks = fix32(hl.nchunks);
        const int nc = hl.nchunks;
        // the subdomain lies inside the binary ocean before any arithmetic on its size
        if (nc <= 0 || base > hd || hl.aggregate_bytes > hd - base || hl.aggregate_bytes < (uint64_t)nc * 8)
            return fail("bad HeaderLUT");
        const uint64_t payload = hl.aggregate_bytes - (uint64_t)nc * 8;
        const uint64_t lut_at = base + payload;
        if (lut_at + (uint64_t)nc * 8 > hd) return fail("LUT outside the binary ocean");
        std::vector<uint64_t> lut(nc);
        std::memcpy(lut.data(), file.data() + lut_at, (size_t)nc * 8);
        for (auto &l : lut) l = fix64(l);
        const size_t g0 = info.chunk_start.size();
        for (int i = 0; i < nc; ++i) {
            const uint64_t end = (i + 1 < nc) ? lut[i + 1] : payload;
            if (lut[i] >= end) return fail("chunk offsets not increasing");
            info.chunk_start.push_back(base + lut[i]);
            info.chunk_size.push_back(end - lut[i]);
        }
        for (long long i = 0; i < bps; ++i, ++cur) mb[cur].idcompression += (int)g0;
        base += hl.aggregate_bytes;
    }
    if (base != hd) return fail("subdomain sizes do not add up to the header displacement");
    info.idx2chunk.assign(nb, CompressedBlock{0, 0, -1});
    for (long long i = 0; i < nb; ++i) {
        const BlockMetadata &m = mb[i];
        if (m.idcompression < 0 || (size_t)m.idcompression >= info.chunk_start.size() || m.subid < 0 || m.ix < 0 ||
            m.ix >= p.bpd[0] || m.iy < 0 || m.iy >= p.bpd[1] || m.iz < 0 || m.iz >= p.bpd[2])
            return fail("metadata entry out of range");
        const size_t id = (size_t)m.ix + (size_t)p.bpd[0] * ((size_t)m.iy + (size_t)p.bpd[1] * m.iz);
        info.idx2chunk[id] = CompressedBlock{info.chunk_start[m.idcompression], info.chunk_size[m.idcompression],
                                             m.subid};
    }
    for (auto &cb : info.idx2chunk)
        if (cb.subid < 0) return fail("a block has no metadata entry");
    return 0;
}

int gather_records(const std::vector<uint8_t> &file, const FileInfo &info, int nthreads,
                   std::vector<uint8_t> &records, std::vector<uint64_t> &offsets, std::string *msg)
{
    const size_t nc = info.chunk_start.size();
    std::vector<std::vector<uint8_t>> raw(nc);
    int bad = 0;
#ifdef _OPENMP
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad) num_threads(nt)
#endif
    for (long long c = 0; c < (long long)nc; ++c) {
        const uint8_t *src = file.data() + info.chunk_start[c];
        if (info.p.encoder == ENC_ZLIB) {
            if (!inflate_chunk(src, info.chunk_size[c], raw[c])) bad = 1;
        } else if (info.p.encoder == ENC_LZ4) {
            if (!lz4_decompress(src, info.chunk_size[c], raw[c], chunk_bound(info.p.prec))) bad = 1;
        } else {
            raw[c].assign(src, src + info.chunk_size[c]);
        }
    }
#ifndef _OPENMP
    (void)nthreads;
#endif
    if (bad) {
        if (msg) *msg = "a chunk failed to inflate";
        return -2;
    }
    std::vector<uint64_t> cbase(nc + 1, 0);
    for (size_t c = 0; c < nc; ++c) cbase[c + 1] = cbase[c] + raw[c].size();
    records.resize(cbase[nc]);
    for (size_t c = 0; c < nc; ++c)
        if (!raw[c].empty()) std::memcpy(records.data() + cbase[c], raw[c].data(), raw[c].size());
    // record starts inside each chunk (walk the int32 size words once per chunk)
    std::vector<std::vector<uint64_t>> starts(nc);
    for (size_t c = 0; c < nc; ++c) {
        uint64_t o = 0;
        while (o + 4 <= raw[c].size()) {
            int32_t n;
            std::memcpy(&n, raw[c].data() + o, 4);
            if (info.swap) {  // the size word, then the record's 4-byte units from byte 4096 on
                n = (int32_t)bswap_u32((uint32_t)n);
                if (n >= 0 && o + 4 + (uint64_t)n <= raw[c].size()) {
                    uint8_t *r = records.data() + cbase[c] + o;
                    std::memcpy(r, &n, 4);
                    for (uint64_t i = 4096; i + 4 <= (uint64_t)n; i += 4) {
                        uint32_t w;
                        std::memcpy(&w, r + 4 + i, 4);
                        w = bswap_u32(w);
                        std::memcpy(r + 4 + i, &w, 4);
                    }
                }
            }
            if (n < 0 || o + 4 + (uint64_t)n > raw[c].size()) {
                if (msg) *msg = "record size word runs past its chunk";
                return -3;
            }
            starts[c].push_back(cbase[c] + o);
            o += 4 + (uint64_t)n;
        }
    }
    std::vector<size_t> chunk_of(nc);
    offsets.assign((size_t)info.nblocks + 1, 0);
    // map each block's (chunk start, subid) to its record
    std::vector<std::pair<uint64_t, size_t>> order(nc);
    for (size_t c = 0; c < nc; ++c) order[c] = {info.chunk_start[c], c};
    std::sort(order.begin(), order.end());
    for (int b = 0; b < info.nblocks; ++b) {
        const CompressedBlock &cb = info.idx2chunk[b];
        auto it = std::lower_bound(order.begin(), order.end(), std::make_pair(cb.start, (size_t)0));
        if (it == order.end() || it->first != cb.start) {
            if (msg) *msg = "block refers to an unknown chunk";
            return -3;
        }
        const size_t c = it->second;
        if ((size_t)cb.subid >= starts[c].size()) {
            if (msg) *msg = "block index within chunk out of range";
            return -3;
        }
        offsets[b] = starts[c][cb.subid];
    }
    offsets[info.nblocks] = records.size();
    return 0;
}

bool read_file(const std::string &path, std::vector<uint8_t> &data)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    data.resize(n > 0 ? (size_t)n : 0);
    const bool ok = n >= 0 && std::fread(data.data(), 1, data.size(), f) == data.size();
    std::fclose(f);
    return ok;
}

bool write_file(const std::string &path, const std::vector<uint8_t> &data)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data.data(), 1, data.size(), f) == data.size();
    return std::fclose(f) == 0 && ok;
}

}  // namespace cz

// ---------------------------------------------------------------------------
// C ABI for language bindings (declared in include/cz_host.h)
#include "../../include/cz_host.h"

struct czh_file {
    std::vector<uint8_t> bytes;
    cz::FileInfo info;
    std::string err;
};

extern "C" {

int czh_entries_per_chunk(int prec) { return cz::entries_per_chunk(prec); }

int czh_write(const char *path, int prec, const int bpd[3], int wtype, double threshold, int encoder,
              const uint8_t *records, size_t records_bytes, const uint64_t *offsets, int nblocks, int nthreads)
{
    return czh_write_subdomains(path, prec, bpd, bpd, wtype, threshold, encoder, records, records_bytes, offsets,
                                nblocks, nthreads);
}

int czh_write_subdomains(const char *path, int prec, const int bpd[3], const int sub_bpd[3], int wtype,
                         double threshold, int encoder, const uint8_t *records, size_t records_bytes,
                         const uint64_t *offsets, int nblocks, int nthreads)
{
    if (!path || !bpd || !sub_bpd || !records || !offsets || (prec != 4 && prec != 8) || nblocks <= 0)
        return CZH_EINVAL;
    // every record inside the buffer: [int32 n][n bytes] between its offset and the next
    for (int b = 0; b < nblocks; ++b) {
        if (offsets[b + 1] < offsets[b] || offsets[b + 1] - offsets[b] < 4 || offsets[b + 1] > records_bytes)
            return CZH_EINVAL;
        int32_t n;
        std::memcpy(&n, records + offsets[b], 4);
        if (n < 0 || 4 + (uint64_t)n > offsets[b + 1] - offsets[b]) return CZH_EINVAL;
    }
    cz::Params p;
    p.prec = prec;
    for (int d = 0; d < 3; ++d) {
        p.bpd[d] = bpd[d];
        p.sub_bpd[d] = sub_bpd[d];
    }
    p.wtype = wtype;
    p.threshold = threshold;
    p.encoder = encoder;
    std::vector<uint8_t> file;
    const int rc = cz::build_cz(p, records, offsets, nblocks, nthreads, file);
    if (rc) return rc == -1 ? CZH_EINVAL : CZH_EZLIB;
    return cz::write_file(path, file) ? CZH_OK : CZH_EIO;
}

czh_file *czh_open(const char *path, char *errbuf, size_t errlen) { return czh_open_ex(path, 0, errbuf, errlen); }

czh_file *czh_open_ex(const char *path, int flags, char *errbuf, size_t errlen)
{
    czh_file *h = new czh_file;
    std::string msg;
    if (!path || !cz::read_file(path, h->bytes)) msg = "cannot read file";
    else if (cz::parse_cz(h->bytes, h->info, &msg, (flags & CZH_OPEN_SWAP) != 0) != 0) {
    } else return h;
    if (errbuf && errlen) std::snprintf(errbuf, errlen, "%s", msg.c_str());
    delete h;
    return nullptr;
}

void czh_close(czh_file *h) { delete h; }

int czh_info(const czh_file *h, int *prec, int bpd[3], int sub_bpd[3], int *wtype, double *threshold, int *encoder)
{
    if (!h) return CZH_EINVAL;
    const cz::Params &p = h->info.p;
    if (prec) *prec = p.prec;
    for (int d = 0; d < 3; ++d) {
        if (bpd) bpd[d] = p.bpd[d];
        if (sub_bpd) sub_bpd[d] = p.sub_bpd[d];
    }
    if (wtype) *wtype = p.wtype;
    if (threshold) *threshold = p.threshold;
    if (encoder) *encoder = p.encoder;
    return CZH_OK;
}

uint64_t czh_file_bytes(const czh_file *h) { return h ? h->bytes.size() : 0; }

int czh_gather(czh_file *h, int nthreads, uint8_t **records, uint64_t *records_bytes, uint64_t **offsets)
{
    if (!h || !records || !records_bytes || !offsets) return CZH_EINVAL;
    std::vector<uint8_t> rec;
    std::vector<uint64_t> off;
    const int rc = cz::gather_records(h->bytes, h->info, nthreads, rec, off, &h->err);
    if (rc) return rc == -2 ? CZH_EZLIB : CZH_ECORRUPT;
    *records = static_cast<uint8_t *>(std::malloc(rec.size() ? rec.size() : 1));
    *offsets = static_cast<uint64_t *>(std::malloc(off.size() * sizeof(uint64_t)));
    if (!*records || !*offsets) {
        std::free(*records);
        std::free(*offsets);
        return CZH_EIO;
    }
    if (!rec.empty()) std::memcpy(*records, rec.data(), rec.size());
    std::memcpy(*offsets, off.data(), off.size() * sizeof(uint64_t));
    *records_bytes = rec.size();
    return CZH_OK;
}

void czh_free(void *p) { std::free(p); }

int czh_lz4_compress(const uint8_t *in, uint64_t n, uint8_t *out, uint64_t cap, uint64_t *produced)
{
    if ((!in && n) || !out || !produced || n > 0x7fffffffull) return CZH_EINVAL;
    std::vector<uint8_t> o;
    cz::lz4_compress(in, (size_t)n, o);
    if (o.size() > cap) return CZH_EINVAL;
    std::memcpy(out, o.data(), o.size());
    *produced = o.size();
    return CZH_OK;
}

int czh_lz4_decompress(const uint8_t *in, uint64_t n, uint8_t *out, uint64_t cap, uint64_t *produced)
{
    if ((!in && n) || !produced) return CZH_EINVAL;
    std::vector<uint8_t> o;
    if (!cz::lz4_decompress(in, (size_t)n, o, (size_t)cap)) return CZH_ECORRUPT;
    if (!o.empty()) std::memcpy(out, o.data(), o.size());
    *produced = o.size();
    return CZH_OK;
}

int czh_deflate(const uint8_t *in, uint64_t n, uint8_t *out, uint64_t cap, uint64_t *produced)
{
    std::vector<uint8_t> o;
    if (!cz::deflate_chunk(in, n, o)) return CZH_EZLIB;
    if (o.size() > cap) return CZH_EINVAL;
    std::memcpy(out, o.data(), o.size());
    *produced = o.size();
    return CZH_OK;
}

}  // extern "C"
