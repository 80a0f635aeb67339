// cz_pipeline.cpp -- a device field straight to a .cz file, with the host
// second stage (zlib, LZ4 or none) overlapped with the device compressor.
//
// The reference writer compresses a block, appends its record to the
// thread's chunk buffer and deflates the buffer every ENTRIES blocks
// (SerializerIO_WaveletCompression_MPI_Simple.h:187-345, 101-184).  Here the
// field is cut into slabs of whole chunks (a multiple of ENTRIES blocks).  The
// device compresses slab k+1 (czw_compress_field_range) while host threads
// encode the chunks of slab k.  The chunks, and so the file, are the same as
// those of one writer compressing the whole field in block order.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
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
    hipEvent_t ev_c0, ev_c1, ev_off;
    if (hipEventCreate(&ev_c0) != hipSuccess || hipEventCreate(&ev_c1) != hipSuccess ||
        hipEventCreate(&ev_off) != hipSuccess)
        return CZH_EIO;

    cz::Params p;
    p.prec = prec;
    for (int k = 0; k < 3; ++k) p.bpd[k] = p.sub_bpd[k] = bpd[k];
    p.wtype = wtype;
    p.threshold = eps;
    p.encoder = encoder;
    std::vector<std::vector<uint8_t>> chunks;
    chunks.reserve((nb + E - 1) / E);
    double t_gpu = 0, t_host = 0, t_copy = 0;
    int rc = CZH_OK;

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
            hipEventRecord(ev_off, stream) != hipSuccess || hipEventSynchronize(ev_off) != hipSuccess)
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
    // Host side of slab k: the second stage of its chunks.
    auto encode = [&](int k) -> int {
        const auto th = std::chrono::steady_clock::now();
        const int b0 = k * slab, n = std::min(slab, nb - b0);
        std::vector<std::vector<uint8_t>> part;
        const int e = cz::encode_chunks(p, rec_h[k & 1].p, reinterpret_cast<const uint64_t *>(off_h[k & 1].p), n,
                                        nthreads, part);
        for (auto &c : part) chunks.push_back(std::move(c));
        t_host += seconds_since(th);
        return e ? CZH_EZLIB : CZH_OK;
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
                    hipStreamSynchronize(stream) != hipSuccess) {
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
    (void)hipEventDestroy(ev_c0);
    (void)hipEventDestroy(ev_c1);
    (void)hipEventDestroy(ev_off);
    if (rc != CZH_OK) return rc;

    const auto tw = std::chrono::steady_clock::now();
    uint64_t fsize = 0;
    const int wrc = path ? cz::write_cz(path, p, chunks, nb, &fsize) : cz::size_cz(p, chunks, nb, &fsize);
    if (wrc) return wrc == -3 ? CZH_EIO : CZH_EINVAL;
    if (file_bytes) *file_bytes = fsize;
    if (times) {
        times[0] = seconds_since(t0);
        times[1] = t_gpu;
        times[2] = t_host;
        times[3] = t_copy;
        times[4] = seconds_since(tw);
    }
    return CZH_OK;
}
