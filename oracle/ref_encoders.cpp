// ref_encoders.cpp -- C-ABI shim around the reference's lossless stage
// (Compressor/source/CompressionEncoders.h: deflate_inplace, zdecompress),
// compiled by oracle/Makefile with -D_USE_ZLIB_ into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: pins the host writer's chunk bytes.
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <zlib.h>
#include "CompressionEncoders.h"

extern "C" {

// Compress buf[0..len) in place exactly as the reference writer does
// (SerializerIO_WaveletCompression_MPI_Simple.h:121-134); returns the
// compressed size, or -1.
long rf_deflate_inplace(unsigned char *buf, unsigned len, unsigned max)
{
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    deflateInit(&zs, Z_DEFAULT_COMPRESSION);
    unsigned mah = max;
    const int err = deflate_inplace(&zs, buf, len, &mah);
    const long out = (err == Z_OK) ? (long)zs.total_out : -1;
    deflateEnd(&zs);
    return out;
}

} // extern "C"
