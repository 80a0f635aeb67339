// adapter_check.cpp -- drives WaveletCompressorGPU (include/czw_compressor.hpp)
// the way CubismZ's writer and reader drive WaveletCompressorGeneric: copy a
// block in, compress, take the payload; memcpy a payload into
// compressed_data(), decompress into a block.
//
//   adapter_check <f32|f64> <block.raw> <eps> <wtype> <zerobits> <shuffle> <payload.out> <decoded.out> [repeat]
//
// With repeat, the block then goes through `repeat` more compress/decompress
// rounds, each with fresh objects (one per block, as the reference's writer
// constructs them); every round must give the same payload and block, and
// the line "allocations N" reports the device buffer sets allocated in all.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/czw_compressor.hpp"

template <typename T>
static int run(char **argv)
{
    std::FILE *f = std::fopen(argv[2], "rb");
    if (!f) return 2;
    std::vector<T> blk(32768);
    if (std::fread(blk.data(), sizeof(T), blk.size(), f) != blk.size()) return 2;
    std::fclose(f);
    const float eps = (float)std::atof(argv[3]);
    const int wtype = std::atoi(argv[4]), zb = std::atoi(argv[5]), sh = std::atoi(argv[6]);

    auto c = std::make_unique<WaveletCompressorGPU<32, T>>(zb, sh != 0);
    c->copy_from(reinterpret_cast<const T(*)[32][32]>(blk.data()));
    const size_t n = c->compress(eps, false, wtype);
    std::vector<unsigned char> payload(n);
    std::memcpy(payload.data(), c->compressed_data(), n);
    std::FILE *o = std::fopen(argv[7], "wb");
    if (!o || std::fwrite(payload.data(), 1, n, o) != n) return 3;
    std::fclose(o);

    auto d = std::make_unique<WaveletCompressorGPU<32, T>>(zb, sh != 0);
    std::memcpy(d->compressed_data(), payload.data(), n);
    std::vector<T> out(32768);
    d->decompress(false, n, wtype, reinterpret_cast<T(*)[32][32]>(out.data()));
    o = std::fopen(argv[8], "wb");
    if (!o || std::fwrite(out.data(), sizeof(T), out.size(), o) != out.size()) return 3;
    std::fclose(o);
    const int repeat = argv[9] ? std::atoi(argv[9]) : 0;
    for (int k = 0; k < repeat; ++k) {
        WaveletCompressorGPU<32, T> w(zb, sh != 0);
        w.copy_from(reinterpret_cast<const T(*)[32][32]>(blk.data()));
        const size_t m = w.compress(eps, false, wtype);
        if (m != n || std::memcmp(w.compressed_data(), payload.data(), n) != 0) return 5;
        WaveletCompressorGPU<32, T> r(zb, sh != 0);
        std::memcpy(r.compressed_data(), payload.data(), n);
        std::vector<T> again(32768);
        r.decompress(false, n, wtype, reinterpret_cast<T(*)[32][32]>(again.data()));
        if (std::memcmp(again.data(), out.data(), sizeof(T) * out.size()) != 0) return 6;
    }
    std::printf("allocations %lu\n", czw_adapter::allocations().load());
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 9 && argc != 10) {
        std::fprintf(stderr, "usage: %s f32|f64 block.raw eps wtype zerobits shuffle payload.out decoded.out [repeat]\n",
                     argv[0]);
        return 1;
    }
    try {
        return std::strcmp(argv[1], "f64") == 0 ? run<double>(argv) : run<float>(argv);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 4;
    }
}
