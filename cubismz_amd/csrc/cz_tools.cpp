// cz_tools.cpp -- the hdf2cz / cz2hdf / cz2diff command-line tools on top of
// the gfx950 block compressor (include/czw.h) and the .cz container
// (cz_file.hpp).  One source; CZ_TOOL selects the program:
//   1 hdf2cz   field -> .cz         (Tools/hdf2cz.cpp, Tools/Test_IO_Compressed.h)
//   2 cz2hdf   .cz -> field (+xmf)  (Tools/cz2hdf.cpp)
//   3 cz2diff  two .cz -> RES line  (Tools/cz2diff.cpp)
// Flags follow the reference tools ("-key value"; a bare flag means true).
// Fields are HDF5 datasets (when built with HDF5) or raw files in HDF5 order
// f[X][Y][Z] (Z fastest) given with -rawfile.
#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/cz_host.h"
#include "../../include/czw.h"
#include "cz_file.hpp"

#ifdef CZ_HAVE_HDF5
#include <hdf5.h>
#endif

namespace {

// -key value pairs; values may be negative numbers (Cubism ArgumentParser.h:147-196)
struct Args {
    std::map<std::string, std::string> kv;
    Args(int argc, char **argv)
    {
        for (int i = 1; i < argc; ++i) {
            std::string k = argv[i];
            if (k.size() < 2 || k[0] != '-') continue;
            std::string v = "true";
            if (i + 1 < argc) {
                const char *n = argv[i + 1];
                const bool is_value = n[0] != '-' || (n[1] && (std::isdigit((unsigned char)n[1]) || n[1] == '.'));
                if (is_value) {
                    v = n;
                    ++i;
                }
            }
            kv[k] = v;
        }
    }
    bool has(const char *k) const { return kv.count(k) != 0; }
    std::string str(const char *k, const char *d) const { return has(k) ? kv.at(k) : d; }
    int i(const char *k, int d) const { return has(k) ? std::atoi(kv.at(k).c_str()) : d; }
    double f(const char *k, double d) const { return has(k) ? std::atof(kv.at(k).c_str()) : d; }
};

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

#define HIPCHECK(x)                                                                          \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) {                                                              \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            std::exit(EXIT_FAILURE);                                                         \
        }                                                                                    \
    } while (0)

// A field in HDF5 order, stored as bytes of Real (prec 4 or 8).
struct Field {
    int n[3] = {0, 0, 0};
    int prec = 4;
    std::vector<uint8_t> data;
    size_t count() const { return (size_t)n[0] * n[1] * n[2]; }
    double at(size_t i) const
    {
        if (prec == 4) {
            float v;
            std::memcpy(&v, data.data() + 4 * i, 4);
            return v;
        }
        double v;
        std::memcpy(&v, data.data() + 8 * i, 8);
        return v;
    }
};

bool read_raw(const std::string &path, Field &f)
{
    std::vector<uint8_t> b;
    if (!cz::read_file(path, b) || b.size() != f.count() * f.prec) {
        std::fprintf(stderr, "raw file %s: expected %zu bytes\n", path.c_str(), f.count() * f.prec);
        return false;
    }
    f.data.swap(b);
    return true;
}

#ifdef CZ_HAVE_HDF5
// Read channel `channel` of dataset `dset` with dims (NX,NY,NZ[,NC]) as Real
// (Tools/Test_IO_Compressed.h:137-212).
bool read_h5(const std::string &path, const std::string &dset, int channel, Field &f)
{
    hid_t file = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (file < 0) return false;
    hid_t ds = H5Dopen2(file, dset.c_str(), H5P_DEFAULT);
    if (ds < 0) {
        H5Fclose(file);
        return false;
    }
    hid_t sp = H5Dget_space(ds);
    const int nd = H5Sget_simple_extent_ndims(sp);
    hsize_t dims[4] = {1, 1, 1, 1};
    H5Sget_simple_extent_dims(sp, dims, NULL);
    const hsize_t nc = nd == 4 ? dims[3] : 1;
    for (int d = 0; d < 3; ++d) {
        if (f.n[d] && (hsize_t)f.n[d] != dims[d]) {
            std::fprintf(stderr, "The dataset size does not match the one specified by the user. Aborting...\n");
            return false;
        }
        f.n[d] = (int)dims[d];
    }
    std::vector<uint8_t> all(f.count() * nc * f.prec);
    const hid_t mt = f.prec == 4 ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE;
    const bool ok = H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, all.data()) >= 0;
    H5Sclose(sp);
    H5Dclose(ds);
    H5Fclose(file);
    if (!ok) return false;
    f.data.resize(f.count() * f.prec);
    for (size_t i = 0; i < f.count(); ++i)
        std::memcpy(f.data.data() + i * f.prec, all.data() + (i * nc + channel) * f.prec, f.prec);
    return true;
}

// Write dataset "data" with dims (NX,NY,NZ,1) (Tools/cz2hdf.cpp:158-199).
// /data as [NX][NY][NZ][nch] (the field's data holds nch values per point).
bool write_h5(const std::string &path, const Field &f, int nch = 1)
{
    hid_t file = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) return false;
    hsize_t dims[4] = {(hsize_t)f.n[0], (hsize_t)f.n[1], (hsize_t)f.n[2], (hsize_t)nch};
    hid_t sp = H5Screate_simple(4, dims, NULL);
    const hid_t mt = f.prec == 4 ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE;
    hid_t ds = H5Dcreate2(file, "data", mt, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const bool ok = ds >= 0 && H5Dwrite(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, f.data.data()) >= 0;
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    H5Fclose(file);
    return ok;
}
#endif

[[maybe_unused]] bool load_field(const Args &a, Field &f)
{
    if (a.has("-rawfile")) return read_raw(a.str("-rawfile", ""), f);
#ifdef CZ_HAVE_HDF5
    if (a.has("-h5file")) return read_h5(a.str("-h5file", ""), a.str("-dataset", "/data"), a.i("-channel", 0), f);
#endif
    std::fprintf(stderr, "no input: give -rawfile (or -h5file in an HDF5 build)\n");
    return false;
}

// Decode a whole .cz into a field: every chunk inflated once (czh gather),
// all wavelet blocks decoded by one device launch.  The file does not record
// the byte shuffle (the reference selects it at compile time, _USE_SHUFFLE3_,
// for its writer and reader alike), so the caller says whether it was used.
// swap reads a file of the other byte order (the tools' -swap).
bool decode_cz(const std::string &path, int expect_wtype, Field &f, bool verbose, bool shuffle, bool swap)
{
    std::vector<uint8_t> bytes;
    cz::FileInfo info;
    std::string msg;
    if (!cz::read_file(path, bytes)) {
        std::fprintf(stderr, "cannot read %s\n", path.c_str());
        return false;
    }
    if (cz::parse_cz(bytes, info, &msg, swap)) {
        std::fprintf(stderr, "%s: %s\n", path.c_str(), msg.c_str());
        return false;
    }
    const cz::Params &p = info.p;
    if (expect_wtype >= 0 && p.wtype >= 0 && cz::wavelet_type(cz::wavelet_name(expect_wtype)) != p.wtype) {
        std::fprintf(stderr, "\nATTENZIONE:\nWavelets in the file is %s and i have %s\n", cz::wavelet_name(p.wtype),
                     cz::wavelet_name(expect_wtype));
        return false;
    }
    if (verbose)
        std::printf("%s: Blocks: %d x %d x %d, sizeofReal %d, Wavelets: %s, WaveletThreshold: %g, Encoder: %s\n",
                    path.c_str(), p.bpd[0], p.bpd[1], p.bpd[2], p.prec, cz::wavelet_name(p.wtype), p.threshold,
                    p.encoder == cz::ENC_ZLIB ? "zlib" : p.encoder == cz::ENC_LZ4 ? "lz4" : "none");
    std::vector<uint8_t> rec;
    std::vector<uint64_t> off;
    if (cz::gather_records(bytes, info, 0, rec, off, &msg)) {
        std::fprintf(stderr, "%s: %s\n", path.c_str(), msg.c_str());
        return false;
    }
    f.prec = p.prec;
    for (int d = 0; d < 3; ++d) f.n[d] = p.bpd[d] * 32;
    f.data.assign(f.count() * f.prec, 0);
    const int nb = info.nblocks;
    if (p.wtype < 0) {  // raw blocks [z][y][x] (the reference's uncompressed build)
        for (int b = 0; b < nb; ++b) {
            int32_t n;
            std::memcpy(&n, rec.data() + off[b], 4);
            if (n != 32768 * f.prec) {
                std::fprintf(stderr, "%s: raw block %d has %d bytes\n", path.c_str(), b, n);
                return false;
            }
            const uint8_t *blk = rec.data() + off[b] + 4;
            const int ix = b % p.bpd[0], iy = (b / p.bpd[0]) % p.bpd[1], iz = b / (p.bpd[0] * p.bpd[1]);
            for (int z = 0; z < 32; ++z)
                for (int y = 0; y < 32; ++y)
                    for (int x = 0; x < 32; ++x) {
                        const size_t gi = ((size_t)(32 * ix + x) * f.n[1] + (32 * iy + y)) * f.n[2] + 32 * iz + z;
                        std::memcpy(f.data.data() + gi * f.prec, blk + ((size_t)(z * 32 + y) * 32 + x) * f.prec,
                                    f.prec);
                    }
        }
        return true;
    }
    void *drec, *dfield;
    uint64_t *doff;
    uint32_t *derr;
    HIPCHECK(hipMalloc(&drec, std::max<size_t>(rec.size(), 4)));
    HIPCHECK(hipMalloc(&doff, off.size() * 8));
    HIPCHECK(hipMalloc(&dfield, f.data.size()));
    HIPCHECK(hipMalloc(&derr, 4));
    HIPCHECK(hipMemcpy(drec, rec.data(), rec.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(doff, off.data(), off.size() * 8, hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(derr, 0, 4));
    const int rc = czw_decompress_field(drec, rec.size(), doff, f.n[0], f.n[1], f.n[2], f.prec, p.wtype, shuffle ? 1 : 0,
                                        dfield, derr, 0);
    if (rc != CZW_OK) {
        std::fprintf(stderr, "czw_decompress_field failed (%d)\n", rc);
        return false;
    }
    uint32_t err = 0;
    HIPCHECK(hipMemcpy(&err, derr, 4, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(f.data.data(), dfield, f.data.size(), hipMemcpyDeviceToHost));
    (void)hipFree(drec);
    (void)hipFree(doff);
    (void)hipFree(dfield);
    (void)hipFree(derr);
    if (err) {
        std::fprintf(stderr, "%s: corrupt block records (error bits %#x)\n", path.c_str(), err);
        return false;
    }
    return true;
}

#if CZ_TOOL == 1
// hdf2cz (Tools/hdf2cz.cpp, Test_IO_Compressed.h:249-342).  As there,
// -nprocx/y/z set the writer ranks and -bpdx/y/z the blocks of each rank, so
// the field holds nproc * bpd blocks per dimension; one process writes the
// file all those ranks would write together.  Extra flags: -precision
// float|double (raw input), -zerobits, -shuffle, -encoder zlib|lz4|none,
// -nowavelets (raw blocks, like the reference's default build).
int tool_main(int argc, char **argv)
{
    Args a(argc, argv);
    const std::string out = a.str("-czfile", "none");
    if (a.has("-help") || out == "none" || (!a.has("-h5file") && !a.has("-rawfile"))) {
        std::printf("Usage: %s -h5file <hdf5 file> | -rawfile <raw file> -czfile <cz file> -threshold <e> [-wtype <wt>] "
                    "[-bpdx <nbx>] [-bpdy <nby>] [-bpdz <nbz>] [-nprocx <px>] [-nprocy <py>] [-nprocz <pz>] "
                    "[-precision float|double] [-zerobits <z>] "
                    "[-shuffle] [-encoder zlib|lz4|none] [-nowavelets]\n", "hdf2cz");
        return 1;
    }
    const int nproc[3] = {a.i("-nprocx", 1), a.i("-nprocy", 1), a.i("-nprocz", 1)};
    const int sub[3] = {a.i("-bpdx", 1), a.i("-bpdy", 1), a.i("-bpdz", 1)};
    for (int d = 0; d < 3; ++d)
        if (nproc[d] < 1 || sub[d] < 1) {
            std::fprintf(stderr, "hdf2cz: -nproc* and -bpd* must be positive\n");
            return 1;
        }
    Field f;
    f.prec = a.str("-precision", "float") == "double" ? 8 : 4;
    for (int d = 0; d < 3; ++d) f.n[d] = nproc[d] * sub[d] * 32;
    if (!load_field(a, f)) return 1;
    const int wtype = a.has("-nowavelets") ? -1 : a.i("-wtype", 3);
    const double threshold = a.f("-threshold", 0.0);
    const int zerobits = a.i("-zerobits", 0), shuffle = a.has("-shuffle") && a.str("-shuffle", "1") != "0";
    const std::string path = a.str("-fpath", ".") + "/" + out;
    std::printf("setting threshold to %f\n", threshold);

    const double t0 = now();
    cz::Params p;
    p.prec = f.prec;
    for (int d = 0; d < 3; ++d) {
        p.bpd[d] = f.n[d] / 32;
        p.sub_bpd[d] = sub[d];
    }
    p.wtype = wtype;
    p.threshold = threshold;
    {
        const std::string e = a.str("-encoder", "zlib");
        if (e != "zlib" && e != "lz4" && e != "none") {
            std::fprintf(stderr, "hdf2cz: unknown -encoder %s (zlib, lz4 or none)\n", e.c_str());
            return 1;
        }
        p.encoder = e == "none" ? cz::ENC_NONE : e == "lz4" ? cz::ENC_LZ4 : cz::ENC_ZLIB;
    }
    const int nb = p.bpd[0] * p.bpd[1] * p.bpd[2];
    size_t file_size = 0;
    if (wtype < 0) {  // raw blocks [z][y][x]
        std::vector<uint8_t> rec;
        std::vector<uint64_t> off(nb + 1);
        const int rb = 4 + 32768 * f.prec;
        rec.resize((size_t)nb * rb);
        for (int b = 0; b < nb; ++b) {
            off[b] = (uint64_t)b * rb;
            const int32_t n = 32768 * f.prec;
            std::memcpy(rec.data() + off[b], &n, 4);
            const int ix = b % p.bpd[0], iy = (b / p.bpd[0]) % p.bpd[1], iz = b / (p.bpd[0] * p.bpd[1]);
            for (int z = 0; z < 32; ++z)
                for (int y = 0; y < 32; ++y)
                    for (int x = 0; x < 32; ++x) {
                        const size_t gi = ((size_t)(32 * ix + x) * f.n[1] + (32 * iy + y)) * f.n[2] + 32 * iz + z;
                        std::memcpy(rec.data() + off[b] + 4 + ((size_t)(z * 32 + y) * 32 + x) * f.prec,
                                    f.data.data() + gi * f.prec, f.prec);
                    }
        }
        off[nb] = rec.size();
        std::vector<uint8_t> file;
        if (cz::build_cz(p, rec.data(), off.data(), nb, 0, file) || !cz::write_file(path, file)) {
            std::fprintf(stderr, "hdf2cz: writing %s failed\n", path.c_str());
            return 1;
        }
        file_size = file.size();
    } else if (nproc[0] * nproc[1] * nproc[2] > 1) {
        // several subdomains: the whole field on the device at once, then the
        // records regrouped rank by rank on the host
        void *dfield, *drec, *dws;
        uint64_t *doff;
        HIPCHECK(hipMalloc(&dfield, f.data.size()));
        HIPCHECK(hipMalloc(&drec, czw_records_bound(nb, f.prec)));
        HIPCHECK(hipMalloc(&dws, czw_compress_workspace_bytes(nb, f.prec)));
        HIPCHECK(hipMalloc(reinterpret_cast<void **>(&doff), sizeof(uint64_t) * (nb + 1)));
        HIPCHECK(hipMemcpy(dfield, f.data.data(), f.data.size(), hipMemcpyHostToDevice));
        const int rc = czw_compress_field(dfield, f.n[0], f.n[1], f.n[2], f.prec, (float)threshold, wtype, zerobits,
                                          shuffle, drec, doff, dws, nullptr);
        std::vector<uint64_t> off(nb + 1);
        std::vector<uint8_t> rec;
        if (rc == CZW_OK && czw_compress_status(dws, nullptr) == CZW_OK) {
            HIPCHECK(hipMemcpy(off.data(), doff, off.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
            rec.resize(off[nb]);
            HIPCHECK(hipMemcpy(rec.data(), drec, off[nb], hipMemcpyDeviceToHost));
        }
        (void)hipFree(dfield);
        (void)hipFree(drec);
        (void)hipFree(dws);
        (void)hipFree(doff);
        std::vector<uint8_t> file;
        if (rc != CZW_OK || rec.size() != off[nb] || off[nb] == 0 || cz::build_cz(p, rec.data(), off.data(), nb, 0, file) ||
            !cz::write_file(path, file)) {
            std::fprintf(stderr, "hdf2cz: compressing to %s failed\n", path.c_str());
            return 1;
        }
        file_size = file.size();
    } else {
        // device compressor, zlib on host threads overlapped slab by slab
        void *dfield;
        HIPCHECK(hipMalloc(&dfield, f.data.size()));
        HIPCHECK(hipMemcpy(dfield, f.data.data(), f.data.size(), hipMemcpyHostToDevice));
        uint64_t fb = 0;
        const int rc = czh_compress_device_field(path.c_str(), dfield, f.n[0], f.n[1], f.n[2], f.prec, (float)threshold,
                                                 wtype, zerobits, shuffle, p.encoder, 0, 0, 1, 0, &fb, nullptr);
        (void)hipFree(dfield);
        if (rc != CZH_OK) {
            std::fprintf(stderr, "hdf2cz: compressing to %s failed (%d)\n", path.c_str(), rc);
            return 1;
        }
        file_size = fb;
    }
    const double t1 = now();
    const int nsub = nproc[0] * nproc[1] * nproc[2];
    const size_t written = file_size - (cz::make_header(p).size() + nb * sizeof(cz::BlockMetadata) +
                                          std::strlen(cz::LUT_TITLE) + nsub * sizeof(cz::HeaderLUT) + 56);
    std::printf("Channel %d: %.2f kB, wavelet-threshold: %.1e, compr. rate: %.2f\n", 0, written / 1024., threshold,
                32768.0 * f.prec * nb / (float)written);
    std::printf("done\nElapsed time: %g seconds\n", t1 - t0);
    return 0;
}
#elif CZ_TOOL == 2
// cz2hdf (Tools/cz2hdf.cpp): decode up to three .cz files (-czfile,
// -czfile2, -czfile3: the channels of /data, cz2hdf.cpp:51-70) over a region
// of interest of whole blocks (-xs/-xe/-ys/-ye/-zs/-ze, block indices, -1 =
// the whole extent, :72-153) to <name>.h5 + <name>.xmf (/data is
// [NX][NY][NZ][channels], the reference's _TRANSPOSE_DATA_ layout), or to a
// raw file of the same layout with -rawfile.
int tool_main(int argc, char **argv)
{
    Args a(argc, argv);
    const std::string in[3] = {a.str("-czfile", "none"), a.str("-czfile2", "none"), a.str("-czfile3", "none")};
    if (a.has("-help") || in[0] == "none" || (!a.has("-h5file") && !a.has("-rawfile"))) {
        std::printf("Usage: %s -czfile <cz file> -h5file <h5 basefilename> | -rawfile <raw file> [-wtype <wt>] "
                    "[-czfile2 <cz file>] [-czfile3 <cz file>] [-xs/-xe/-ys/-ye/-zs/-ze <block>] [-shuffle] [-swap]\n",
                    argv[0]);
        return 1;
    }
    const bool swap = a.has("-swap");  // cz2hdf.cpp:79
    const double t0 = now();
    const bool shuffle = a.has("-shuffle") && a.str("-shuffle", "1") != "0";
    Field ch[3];
    int nch = 0;
    for (int c = 0; c < 3; ++c) {
        if (in[c] == "none") continue;
        if (!decode_cz(in[c], a.i("-wtype", -1), ch[nch], true, shuffle, swap)) return 1;
        if (nch && (std::memcmp(ch[nch].n, ch[0].n, sizeof(ch[0].n)) || ch[nch].prec != ch[0].prec)) {
            std::fprintf(stderr, "cz2hdf: %s differs in extent or precision from %s\n", in[c].c_str(), in[0].c_str());
            return 1;
        }
        ++nch;
    }
    const Field &f0 = ch[0];
    const int nbk[3] = {f0.n[0] / 32, f0.n[1] / 32, f0.n[2] / 32};
    std::printf("I found in total %dx%dx%d blocks.\n", nbk[0], nbk[1], nbk[2]);
    const char *ks[3] = {"-xs", "-ys", "-zs"}, *ke[3] = {"-xe", "-ye", "-ze"};
    int st[3], en[3];
    for (int d = 0; d < 3; ++d) {
        const int s0 = a.i(ks[d], -1), e0 = a.i(ke[d], -1);
        st[d] = s0 == -1 ? 0 : s0;
        en[d] = e0 == -1 ? nbk[d] - 1 : e0;
        if (st[d] < 0 || en[d] >= nbk[d] || st[d] > en[d]) {
            std::fprintf(stderr, "cz2hdf: empty or out-of-range region in %c\n", "xyz"[d]);
            return 1;
        }
    }
    std::printf("ROI = [%d,%d]x[%d,%d]x[%d,%d]\n", st[0], en[0], st[1], en[1], st[2], en[2]);
    Field out;
    out.prec = f0.prec;
    for (int d = 0; d < 3; ++d) out.n[d] = (en[d] - st[d] + 1) * 32;
    out.data.resize(out.count() * nch * out.prec);
    const size_t p = out.prec;
    for (int x = 0; x < out.n[0]; ++x)
        for (int y = 0; y < out.n[1]; ++y)
            for (int c = 0; c < nch; ++c) {
                const size_t src = ((size_t)(x + 32 * st[0]) * f0.n[1] + (y + 32 * st[1])) * f0.n[2] + 32 * st[2];
                uint8_t *dst = out.data.data() + (((size_t)x * out.n[1] + y) * out.n[2] * nch + c) * p;
                const uint8_t *s = ch[c].data.data() + src * p;
                for (int z = 0; z < out.n[2]; ++z) std::memcpy(dst + (size_t)z * nch * p, s + (size_t)z * p, p);
            }
    if (a.has("-rawfile")) {
        if (!cz::write_file(a.str("-rawfile", ""), out.data)) return 1;
    } else {
#ifdef CZ_HAVE_HDF5
        const std::string base = a.str("-h5file", "out");
        const std::string h5 = base + ".h5";
        if (!write_h5(h5, out, nch)) {
            std::fprintf(stderr, "cz2hdf: writing %s failed\n", h5.c_str());
            return 1;
        }
        FILE *xmf = std::fopen((base + ".xmf").c_str(), "w");  // cz2hdf.cpp:318-374
        if (xmf) {
            const int mx = std::max(out.n[0], std::max(out.n[1], out.n[2]));
            const char *nt = out.prec == 4 ? "Float\" Precision=\"4" : "Double\" Precision=\"8";
            std::string name = h5.substr(h5.find_last_of('/') == std::string::npos ? 0 : h5.find_last_of('/') + 1);
            std::fprintf(xmf, "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n<Xdmf Version=\"2.0\">\n");
            std::fprintf(xmf, " <Domain>\n   <Grid GridType=\"Uniform\">\n     <Time Value=\"%05d\"/>\n", 0);
            std::fprintf(xmf, "     <Topology TopologyType=\"3DCORECTMesh\" Dimensions=\"%d %d %d\"/>\n", out.n[0], out.n[1],
                         out.n[2]);
            std::fprintf(xmf, "     <Geometry GeometryType=\"ORIGIN_DXDYDZ\">\n");
            std::fprintf(xmf, "       <DataItem Name=\"Origin\" Dimensions=\"3\" NumberType=\"%s\" Format=\"XML\">\n", nt);
            std::fprintf(xmf, "        %e %e %e\n       </DataItem>\n", 0., 0., 0.);
            std::fprintf(xmf, "       <DataItem Name=\"Spacing\" Dimensions=\"3\" NumberType=\"%s\" Format=\"XML\">\n", nt);
            const double h = out.prec == 4 ? (double)(1.f / (float)mx) : 1.0 / (double)mx;
            std::fprintf(xmf, "        %e %e %e\n       </DataItem>\n     </Geometry>\n", h, h, h);
            std::fprintf(xmf, "     <Attribute Name=\"data\" AttributeType=\"%s\" Center=\"Node\">\n",
                         nch == 1 ? "Scalar" : "Vector");
            std::fprintf(xmf, "       <DataItem Dimensions=\"%d %d %d %d\" NumberType=\"Float\" Precision=\"%d\" Format=\"HDF\">\n",
                         out.n[0], out.n[1], out.n[2], nch, out.prec);
            std::fprintf(xmf, "        %s:/data\n       </DataItem>\n     </Attribute>\n   </Grid>\n </Domain>\n</Xdmf>\n",
                         name.c_str());
            std::fclose(xmf);
        }
#else
        std::fprintf(stderr, "cz2hdf: built without HDF5; use -rawfile\n");
        return 1;
#endif
    }
    std::printf("Elapsed time = %.3lf seconds\n", now() - t0);
    return 0;
}
#elif CZ_TOOL == 3
// cz2diff (Tools/cz2diff.cpp:29-268): error norms of czfile1 against czfile2,
// accumulated in the reference's block and element order, and the RES line.
int tool_main(int argc, char **argv)
{
    Args a(argc, argv);
    const std::string f1n = a.str("-czfile1", "none"), f2n = a.str("-czfile2", "none");
    if (a.has("-help") || f1n == "none" || (f2n == "none" && !a.has("-rawfile2"))) {
        std::printf("Usage: %s -czfile1 <cz file1> [-wtype <wt>] -czfile2 <cz reference file2> [-shuffle] [-swap]\n",
                    argv[0]);
        return 1;
    }
    const double init_t0 = now();
    Field A, B;
    const bool shuffle = a.has("-shuffle") && a.str("-shuffle", "1") != "0";  // both files, as in one build
    const bool swap = a.has("-swap");  // both readers, cz2diff.cpp:70-74
    if (!decode_cz(f1n, a.i("-wtype", 3), A, false, shuffle, swap)) return 1;
    if (a.has("-rawfile2")) {
        B.prec = A.prec;
        std::memcpy(B.n, A.n, sizeof(B.n));
        if (!read_raw(a.str("-rawfile2", ""), B)) return 1;
    } else if (!decode_cz(f2n, -1, B, false, shuffle, swap)) {
        return 1;
    }
    const double init_t1 = now();
    std::printf("[czfile1] I found in total %dx%dx%d blocks.\n", A.n[0] / 32, A.n[1] / 32, A.n[2] / 32);
    std::printf("[czfile2] I found in total %dx%dx%d blocks.\n", B.n[0] / 32, B.n[1] / 32, B.n[2] / 32);
    if (A.n[0] != B.n[0] || A.n[1] != B.n[1] || A.n[2] != B.n[2]) {
        std::printf("Dimensions differ, exiting..\n");
        return 1;
    }
    const double t0 = now();
    const int NBX = A.n[0] / 32, NBY = A.n[1] / 32, NBZ = A.n[2] / 32;
    long n = 0;
    double e_inf = 0, e_1 = 0, e_2 = 0, n_inf = 0, n_1 = 0, n_2 = 0;
    double maxdata = -DBL_MAX, mindata = DBL_MAX;
    for (int b = 0; b < NBX * NBY * NBZ; ++b) {
        const int z = b / (NBY * NBX), y = (b / NBX) % NBY, x = b % NBX;
        for (int zb = 0; zb < 32; zb++)
            for (int yb = 0; yb < 32; yb++)
                for (int xb = 0; xb < 32; xb++) {
                    const size_t gi = ((size_t)(32 * x + xb) * A.n[1] + (32 * y + yb)) * A.n[2] + 32 * z + zb;
                    const double f1 = A.at(gi), f2 = B.at(gi);
                    if (f2 > maxdata) maxdata = f2;
                    if (f2 < mindata) mindata = f2;
                    const double v = std::fabs(f2);
                    if (v > n_inf) n_inf = v;
                    n_1 += v;
                    n_2 += v * v;
                    const double err = std::fabs(f1 - f2);
                    if (err > e_inf) e_inf = err;
                    e_1 += err;
                    e_2 += err * err;
                    n++;
                }
    }
    const double t1 = now();
    std::printf("Init time = %.3lf seconds\n", init_t1 - init_t0);
    std::printf("Elapsed time = %.3lf seconds\n", t1 - t0);
    const double linf = e_inf;
    const long nall = n;
    const double l1 = e_1 / nall;
    const long double mse = e_2 / nall;
    const double uncompressed = (double)A.prec * nall;
    double compressed = uncompressed;
    struct stat st;
    if (stat(f1n.c_str(), &st) == 0) compressed = (double)st.st_size;
    std::printf("compression-rate: %.2f rel-linf-error: %e rel-mean-error: %e\n", uncompressed / compressed, linf, l1);
    const double psnr = 20 * std::log10((maxdata - mindata) / (2 * std::sqrt(mse)));
    std::printf("RES: %12s %12s %12s %12s %12s %12s %12s %12s\n", "CR", "rel(e_inf)", "rel(e_1)", "mean(e_1)",
                "rel(e_2)", "mean(e_2)", "BPS", "PSNR");
    std::printf("RES: %12.2f %.6e %.6e %.6e %.6e %.6e %12.04f %12.04f\n", uncompressed / compressed, e_inf / n_inf,
                e_1 / n_1, e_1 / n, std::sqrt(e_2) / std::sqrt(n_2), std::sqrt(e_2) / n, compressed * 8. / nall, psnr);
    return 0;
}
#endif

}  // namespace

int main(int argc, char **argv) { return tool_main(argc, argv); }
