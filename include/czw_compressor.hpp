// czw_compressor.hpp -- a drop-in for CubismZ's per-block compressor class
// WaveletCompressorGeneric<BS, DataType> (Compressor/source/WaveletCompressor.h:35-94)
// on top of the gfx950 C ABI (czw.h).
//
// Same member names and meanings:
//   uncompressed_data()    the 32^3 block, [z][y][x], written by the caller;
//   compress(eps, f16, wt) fwt + threshold + mask + survivors (+ zerobits,
//                          shuffle), returns the payload size in bytes, the
//                          payload is at compressed_data()
//                          (WaveletCompressor.cpp:153-182);
//   compressed_data()      the payload buffer (the caller memcpy's a payload
//                          in before decompressing, as Reader...h:700 does);
//   decompress(f16, n, wt[, out]) decode + iwt (WaveletCompressor.cpp:232-259),
//                          the result in uncompressed_data() or copied to out;
//   copy_to / copy_from    as in WaveletCompressor.h:73-93.
// The reference selects zerobits and the byte shuffle at compile time
// (_USE_ZEROBITS_/_ZEROBITS_, _USE_SHUFFLE3_); here they are constructor
// arguments.  float16 is ignored, as in the reference.
//
// One block per call keeps the GPU nearly idle; this class exists for code
// that is written around the per-block object.  Whole fields go through
// czw_compress_field / czw_decompress_field instead.
//
// The reference constructs one object per block on the stack
// (SerializerIO_WaveletCompression_MPI_Simple.h:214, Reader...h:683), so the
// object itself holds only host memory: its device buffers belong to the
// calling thread (one set per device and precision, allocated on the
// thread's first call and freed when the thread ends) and are shared by every
// object that thread creates.  Objects of one thread therefore must not be
// used concurrently on different streams.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "czw.h"

namespace czw_adapter {

// Device buffers of one block's compress or decompress.
struct Scratch {
    void *block = nullptr, *rec = nullptr, *ws = nullptr;
    uint64_t *off = nullptr;
    uint32_t *err = nullptr;
    bool ready = false;
    Scratch() = default;
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
    ~Scratch()
    {
        if (!ready) return;
        (void)hipFree(block);
        (void)hipFree(rec);
        (void)hipFree(ws);
        (void)hipFree(off);
        (void)hipFree(err);
    }
};

// Buffer sets allocated so far in this process (for tests and accounting).
inline std::atomic<unsigned long> &allocations()
{
    static std::atomic<unsigned long> n{0};
    return n;
}

// The calling thread's buffers on the current device for Real of PREC bytes.
template <int PREC>
Scratch &scratch()
{
    constexpr int MAXDEV = 16;
    thread_local Scratch sets[MAXDEV];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAXDEV) throw std::runtime_error("czw: no usable device");
    Scratch &x = sets[dev];
    if (!x.ready) {
        auto chk = [](hipError_t e) {
            if (e != hipSuccess) throw std::runtime_error(std::string("czw: hipMalloc: ") + hipGetErrorString(e));
        };
        chk(hipMalloc(&x.block, (size_t)PREC * 32768));
        chk(hipMalloc(&x.rec, czw_records_bound(1, PREC)));
        chk(hipMalloc(&x.ws, czw_compress_workspace_bytes(1, PREC)));
        chk(hipMalloc(reinterpret_cast<void **>(&x.off), 2 * sizeof(uint64_t)));
        chk(hipMalloc(reinterpret_cast<void **>(&x.err), sizeof(uint32_t)));
        x.ready = true;
        allocations() += 1;
    }
    return x;
}

}  // namespace czw_adapter

template <int BS, typename DataType>
class WaveletCompressorGPU
{
    static_assert(BS == 32, "the gfx950 kernels are built for 32^3 blocks");
    static_assert(sizeof(DataType) == 4 || sizeof(DataType) == 8, "float or double");

public:
    enum { BS3 = BS * BS * BS, BITSETSIZE = (BS3 + 7) / 8, BUFMAXSIZE = BITSETSIZE + sizeof(DataType) * BS3 };

    explicit WaveletCompressorGPU(int zerobits = 0, bool shuffle = false, hipStream_t stream = nullptr)
        : zerobits_(zerobits), shuffle_(shuffle), stream_(stream)
    {
    }
    WaveletCompressorGPU(const WaveletCompressorGPU &) = delete;
    WaveletCompressorGPU &operator=(const WaveletCompressorGPU &) = delete;

    DataType (&uncompressed_data())[BS][BS][BS] { return data_; }

    void *compressed_data() { return buf_ + 4; }

    size_t compress(const float threshold, const bool /*float16*/, int wtype)
    {
        const int p = (int)sizeof(DataType);
        czw_adapter::Scratch &d = czw_adapter::scratch<(int)sizeof(DataType)>();
        check(hipMemcpyAsync(d.block, &data_[0][0][0], sizeof(data_), hipMemcpyHostToDevice, stream_), "copy");
        status(czw_compress_blocks(d.block, 1, p, threshold, wtype, zerobits_, shuffle_ ? 1 : 0, d.rec, d.off, d.ws,
                                   stream_),
               "czw_compress_blocks");
        uint64_t off[2];
        check(hipMemcpyAsync(off, d.off, sizeof(off), hipMemcpyDeviceToHost, stream_), "copy");
        status(czw_compress_status(d.ws, stream_), "czw_compress_blocks (device)");
        if (off[1] > sizeof(buf_)) throw std::runtime_error("czw: record larger than a block's bound");
        check(hipMemcpyAsync(buf_, d.rec, off[1], hipMemcpyDeviceToHost, stream_), "copy");
        check(hipStreamSynchronize(stream_), "sync");
        int32_t n;
        std::memcpy(&n, buf_, 4);
        return (size_t)n;
    }

    void decompress(const bool /*float16*/, size_t bytes, int wtype)
    {
        if (bytes > (size_t)BUFMAXSIZE) throw std::runtime_error("czw: payload larger than a block's bound");
        const int p = (int)sizeof(DataType);
        const int32_t n = (int32_t)bytes;
        std::memcpy(buf_, &n, 4);
        const uint64_t off[2] = {0, 4 + (uint64_t)bytes};
        const uint32_t zero = 0;
        czw_adapter::Scratch &d = czw_adapter::scratch<(int)sizeof(DataType)>();
        check(hipMemcpyAsync(d.rec, buf_, 4 + bytes, hipMemcpyHostToDevice, stream_), "copy");
        check(hipMemcpyAsync(d.off, off, sizeof(off), hipMemcpyHostToDevice, stream_), "copy");
        check(hipMemcpyAsync(d.err, &zero, sizeof(zero), hipMemcpyHostToDevice, stream_), "copy");
        status(czw_decompress_blocks(d.rec, 4 + bytes, d.off, 1, p, wtype, shuffle_ ? 1 : 0, d.block, d.err, stream_),
               "czw_decompress_blocks");
        uint32_t err = 0;
        check(hipMemcpyAsync(&err, d.err, sizeof(err), hipMemcpyDeviceToHost, stream_), "copy");
        check(hipMemcpyAsync(&data_[0][0][0], d.block, sizeof(data_), hipMemcpyDeviceToHost, stream_), "copy");
        check(hipStreamSynchronize(stream_), "sync");
        if (err) throw std::runtime_error("czw: corrupt payload (error bits " + std::to_string(err) + ")");
    }

    void decompress(const bool float16, size_t ninputbytes, int wtype, DataType out[BS][BS][BS])
    {
        decompress(float16, ninputbytes, wtype);
        copy_to(out);
    }

    void copy_to(DataType out[BS][BS][BS]) { std::memcpy(&out[0][0][0], &data_[0][0][0], sizeof(data_)); }
    void copy_from(const DataType in[BS][BS][BS]) { std::memcpy(&data_[0][0][0], &in[0][0][0], sizeof(data_)); }

private:
    static void check(hipError_t e, const char *what)
    {
        if (e != hipSuccess) throw std::runtime_error(std::string("czw: ") + what + ": " + hipGetErrorString(e));
    }
    static void status(int rc, const char *what)
    {
        if (rc != CZW_OK) throw std::runtime_error(std::string("czw: ") + what + " failed (" + std::to_string(rc) + ")");
    }

    DataType data_[BS][BS][BS];
    unsigned char buf_[4 + BUFMAXSIZE];  // [int32 size][payload]
    int zerobits_;
    bool shuffle_;
    hipStream_t stream_;
};
