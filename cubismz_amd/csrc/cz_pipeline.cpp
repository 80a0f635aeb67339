// cz_pipeline.cpp -- a device field straight to a .cz file, with the host
// second stage (zlib, LZ4 or none) overlapped with the device compressor.
//
// The reference writer compresses a block, appends its record to the
// thread's chunk buffer and deflates the buffer every ENTRIES blocks
// (SerializerIO_WaveletCompression_MPI_Simple.h:187-345, 101-184).  Here the
// field is cut into slabs of whole chunks (a multiple of ENTRIES blocks).  The
// device compresses slab k+1 (czw_compress_field_range) while host threads
// encode the chunks of slab k, and each slab's chunks go to the file as soon
// as they are encoded (cz::CzWriter), so host memory holds two slabs of
// records and one of chunks.  The chunks, and so the file, are the same as
// those of one writer compressing the whole field in block order.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/cz_host.h"
#include "../../include/czw.h"
#include "cz_file.hpp"

namespace {

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

struct DeviceSlab {
    void *records = nullptr, *workspace = nullptr;
    uint64_t *offsets = nullptr;
    ~DeviceSlab()
    {
        if (records) (void)hipFree(records);
        if (workspace) (void)hipFree(workspace);
        if (offsets) (void)hipFree(offsets);
    }
};

struct Events {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    bool create()
    {
        for (auto &x : e)
            if (hipEventCreate(&x) != hipSuccess) return false;
        return true;
    }
    ~Events()
    {
        for (auto x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

struct OutFile {
    FILE *f = nullptr;
    const char *path = nullptr;
    bool keep = false;
    ~OutFile()
    {
        if (f) std::fclose(f);
        if (path && !keep) std::remove(path);  // no partial file is left behind
    }
};

struct HostBuf {
    uint8_t *p = nullptr;
    size_t cap = 0;
    ~HostBuf()
    {
        if (p) (void)hipHostFree(p);
    }
    bool reserve(size_t n)
    {
        if (n <= cap) return true;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        if (hipHostMalloc(reinterpret_cast<void **>(&p), n, hipHostMallocDefault) != hipSuccess) return false;
        cap = n;
        return true;
    }
};

}  // namespace

extern "C" int czh_compress_device_field(const char *path, const void *field, int nx, int ny, int nz, int prec,
                                         float eps, int wtype, int zerobits, int shuffle, int encoder, int nthreads,
                                         int slab_blocks, int overlap, hipStream_t stream, uint64_t *file_bytes,
                                         double *times)
{
    if (!field || (prec != 4 && prec != 8) || nx <= 0 || ny <= 0 || nz <= 0 || (nx | ny | nz) & 31 ||
        (encoder != CZH_ENCODER_NONE && encoder != CZH_ENCODER_ZLIB && encoder != CZH_ENCODER_LZ4))
        return CZH_EINVAL;
    const auto t0 = std::chrono::steady_clock::now();
    const int bpd[3] = {nx / 32, ny / 32, nz / 32};
    const int nb = bpd[0] * bpd[1] * bpd[2];
    const int E = cz::entries_per_chunk(prec);
    int slab = slab_blocks > 0 ? slab_blocks : 1024;
    slab = std::max(E, slab / E * E);
    slab = std::min(slab, (nb + E - 1) / E * E);
    const int nslabs = (nb + slab - 1) / slab;

    DeviceSlab d;
    const size_t rbound = czw_records_bound(slab, prec);
    if (hipMalloc(&d.records, rbound) != hipSuccess || hipMalloc(&d.workspace, czw_compress_workspace_bytes(slab, prec)) !=
            hipSuccess || hipMalloc(reinterpret_cast<void **>(&d.offsets), sizeof(uint64_t) * (slab + 1)) != hipSuccess)
        return CZH_EIO;
    HostBuf rec_h[2], off_h[2];
    for (auto &o : off_h)
        if (!o.reserve(sizeof(uint64_t) * (slab + 1))) return CZH_EIO;
    Events ev;
    if (!ev.create()) return CZH_EIO;
    hipEvent_t ev_c0 = ev.e[0], ev_c1 = ev.e[1], ev_off = ev.e[2];

    cz::Params p;
    p.prec = prec;
    for (int k = 0; k < 3; ++k) p.bpd[k] = p.sub_bpd[k] = bpd[k];
    p.wtype = wtype;
    p.threshold = eps;
    p.encoder = encoder;
    OutFile out;
    if (path) {
        out.f = std::fopen(path, "wb");
        if (!out.f) return CZH_EIO;
        out.path = path;
    }
    cz::CzSink sink = path ? cz::CzSink(out.f) : cz::CzSink();
    cz::CzWriter writer(p, nb, sink);
    if (!writer.valid()) return CZH_EINVAL;
    double t_gpu = 0, t_host = 0, t_copy = 0, t_write = 0;
    int rc = CZH_OK;
    {
        const auto tw = std::chrono::steady_clock::now();
        if (writer.begin()) return CZH_EIO;
        t_write += seconds_since(tw);
    }

    // Device side of slab k: compress, then copy its offsets and records to
    // host buffer k % 2.  Returns with the copies queued on the stream.
    auto issue = [&](int k) -> int {
        const int b0 = k * slab, n = std::min(slab, nb - b0);
        (void)hipEventRecord(ev_c0, stream);
        if (czw_compress_field_range(field, nx, ny, nz, prec, eps, wtype, zerobits, shuffle, b0, n, d.records,
                                     d.offsets, d.workspace, stream) != CZW_OK)
            return CZH_EIO;
        (void)hipEventRecord(ev_c1, stream);
        uint64_t *oh = reinterpret_cast<uint64_t *>(off_h[k & 1].p);
        if (hipMemcpyAsync(oh, d.offsets, sizeof(uint64_t) * (n + 1), hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipEventRecord(ev_off, stream) != hipSuccess || hipEventSynchronize(ev_off) != hipSuccess ||
            czw_compress_status(d.workspace, stream) != CZW_OK)
            return CZH_EIO;
        float ms = 0;
        if (hipEventElapsedTime(&ms, ev_c0, ev_c1) == hipSuccess) t_gpu += ms * 1e-3;
        const auto tc = std::chrono::steady_clock::now();
        if (!rec_h[k & 1].reserve(oh[n] ? oh[n] : 1) ||
            hipMemcpyAsync(rec_h[k & 1].p, d.records, oh[n], hipMemcpyDeviceToHost, stream) != hipSuccess)
            return CZH_EIO;
        if (hipStreamSynchronize(stream) != hipSuccess) return CZH_EIO;
        t_copy += seconds_since(tc);
        return CZH_OK;
    };
    // Host side of slab k: the second stage of its chunks, which then go to
    // the file.
    auto encode = [&](int k) -> int {
        const auto th = std::chrono::steady_clock::now();
        const int b0 = k * slab, n = std::min(slab, nb - b0);
        std::vector<std::vector<uint8_t>> part;
        const int e = cz::encode_chunks(p, rec_h[k & 1].p, reinterpret_cast<const uint64_t *>(off_h[k & 1].p), n,
                                        nthreads, part);
        t_host += seconds_since(th);
        if (e) return CZH_EZLIB;
        const auto tw = std::chrono::steady_clock::now();
        for (const auto &c : part)
            if (writer.chunk(c.data(), c.size())) return CZH_EIO;
        t_write += seconds_since(tw);
        return CZH_OK;
    };

    rc = issue(0);
    for (int k = 0; k < nslabs && rc == CZH_OK; ++k) {
        if (overlap) {
            // the device works on slab k+1 while the host deflates slab k
            if (k + 1 < nslabs) {
                const int b0 = (k + 1) * slab, n = std::min(slab, nb - b0);
                (void)hipEventRecord(ev_c0, stream);
                if (czw_compress_field_range(field, nx, ny, nz, prec, eps, wtype, zerobits, shuffle, b0, n,
                                             d.records, d.offsets, d.workspace, stream) != CZW_OK) {
                    rc = CZH_EIO;
                    break;
                }
                (void)hipEventRecord(ev_c1, stream);
            }
            rc = encode(k);
            if (rc == CZH_OK && k + 1 < nslabs) {
                // the compress of slab k+1 is queued already; fetch its results
                const int b0 = (k + 1) * slab, n = std::min(slab, nb - b0);
                uint64_t *oh = reinterpret_cast<uint64_t *>(off_h[(k + 1) & 1].p);
                if (hipMemcpyAsync(oh, d.offsets, sizeof(uint64_t) * (n + 1), hipMemcpyDeviceToHost, stream) !=
                        hipSuccess ||
                    hipStreamSynchronize(stream) != hipSuccess || czw_compress_status(d.workspace, stream) != CZW_OK) {
                    rc = CZH_EIO;
                    break;
                }
                float ms = 0;
                if (hipEventElapsedTime(&ms, ev_c0, ev_c1) == hipSuccess) t_gpu += ms * 1e-3;
                const auto tc = std::chrono::steady_clock::now();
                if (!rec_h[(k + 1) & 1].reserve(oh[n] ? oh[n] : 1) ||
                    hipMemcpyAsync(rec_h[(k + 1) & 1].p, d.records, oh[n], hipMemcpyDeviceToHost, stream) !=
                        hipSuccess ||
                    hipStreamSynchronize(stream) != hipSuccess) {
                    rc = CZH_EIO;
                    break;
                }
                t_copy += seconds_since(tc);
            }
        } else {
            rc = encode(k);
            if (rc == CZH_OK && k + 1 < nslabs) rc = issue(k + 1);
        }
    }
    if (rc != CZH_OK) return rc;

    const auto tw = std::chrono::steady_clock::now();
    if (writer.finish()) return CZH_EIO;
    if (out.f) {
        const int c = std::fclose(out.f);
        out.f = nullptr;
        if (c != 0) return CZH_EIO;
    }
    out.keep = true;
    t_write += seconds_since(tw);
    if (file_bytes) *file_bytes = sink.bytes();
    if (times) {
        times[0] = seconds_since(t0);
        times[1] = t_gpu;
        times[2] = t_host;
        times[3] = t_copy;
        times[4] = t_write;
    }
    return CZH_OK;
}
