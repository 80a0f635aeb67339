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
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "czw.h"

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
        const int p = (int)sizeof(DataType);
        check(hipMalloc(&dblock_, sizeof(DataType) * BS3), "hipMalloc");
        check(hipMalloc(&drec_, czw_records_bound(1, p)), "hipMalloc");
        check(hipMalloc(&dws_, czw_compress_workspace_bytes(1, p)), "hipMalloc");
        check(hipMalloc(reinterpret_cast<void **>(&doff_), 2 * sizeof(uint64_t)), "hipMalloc");
        check(hipMalloc(reinterpret_cast<void **>(&derr_), sizeof(uint32_t)), "hipMalloc");
    }
    ~WaveletCompressorGPU()
    {
        (void)hipFree(dblock_);
        (void)hipFree(drec_);
        (void)hipFree(dws_);
        (void)hipFree(doff_);
        (void)hipFree(derr_);
    }
    WaveletCompressorGPU(const WaveletCompressorGPU &) = delete;
    WaveletCompressorGPU &operator=(const WaveletCompressorGPU &) = delete;

    DataType (&uncompressed_data())[BS][BS][BS] { return data_; }

    void *compressed_data() { return buf_ + 4; }

    size_t compress(const float threshold, const bool /*float16*/, int wtype)
    {
        const int p = (int)sizeof(DataType);
        check(hipMemcpyAsync(dblock_, &data_[0][0][0], sizeof(data_), hipMemcpyHostToDevice, stream_), "copy");
        status(czw_compress_blocks(dblock_, 1, p, threshold, wtype, zerobits_, shuffle_ ? 1 : 0, drec_, doff_, dws_,
                                   stream_),
               "czw_compress_blocks");
        uint64_t off[2];
        check(hipMemcpyAsync(off, doff_, sizeof(off), hipMemcpyDeviceToHost, stream_), "copy");
        check(hipStreamSynchronize(stream_), "sync");
        check(hipMemcpyAsync(buf_, drec_, off[1], hipMemcpyDeviceToHost, stream_), "copy");
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
        check(hipMemcpyAsync(drec_, buf_, 4 + bytes, hipMemcpyHostToDevice, stream_), "copy");
        check(hipMemcpyAsync(doff_, off, sizeof(off), hipMemcpyHostToDevice, stream_), "copy");
        check(hipMemsetAsync(derr_, 0, sizeof(uint32_t), stream_), "memset");
        status(czw_decompress_blocks(drec_, 4 + bytes, doff_, 1, p, wtype, shuffle_ ? 1 : 0, dblock_, derr_, stream_),
               "czw_decompress_blocks");
        uint32_t err = 0;
        check(hipMemcpyAsync(&err, derr_, sizeof(err), hipMemcpyDeviceToHost, stream_), "copy");
        check(hipMemcpyAsync(&data_[0][0][0], dblock_, sizeof(data_), hipMemcpyDeviceToHost, stream_), "copy");
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
    void *dblock_ = nullptr, *drec_ = nullptr, *dws_ = nullptr;
    uint64_t *doff_ = nullptr;
    uint32_t *derr_ = nullptr;
};
