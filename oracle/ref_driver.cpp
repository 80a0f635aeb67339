// ref_driver.cpp -- a small C-ABI shim around the reference's own block
// compressor, compiled together with the reference sources where they lie
// (Compressor/source/WaveletCompressor.cpp) by oracle/Makefile into
// oracle/_ref/.  TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement
// and to write golden vectors; never part of the product path.
//
// The reference selects zerobits/shuffle at compile time (_USE_ZEROBITS_,
// _ZEROBITS_, _USE_SHUFFLE3_, Tools/Makefile:78-93), so one library is built
// per variant; every variant exports the same symbols.
#include <cstring>
#include <cstdlib>
#include "WaveletCompressor.h"

extern "C" {

int rf_sizeof_real(void) { return (int)sizeof(Real); }
long rf_sizeof_compressor(void) { return (long)sizeof(WaveletCompressor); }
int rf_zerobits(void)
{
#if defined(_USE_ZEROBITS_)
    return _ZEROBITS_;
#else
    return 0;
#endif
}
int rf_shuffle(void)
{
#if defined(_USE_SHUFFLE3_)
    return 1;
#else
    return 0;
#endif
}

// block: [z][y][x] Real; out: >= 4096 + 32768*sizeof(Real) bytes.
long rf_compress(const Real *block, float threshold, int wtype, unsigned char *out)
{
    WaveletCompressor *c = new WaveletCompressor;
    std::memcpy(&c->uncompressed_data()[0][0][0], block, sizeof(Real) * 32768);
    const size_t n = c->compress(threshold, false, wtype);
    std::memcpy(out, c->compressed_data(), n);
    delete c;
    return (long)n;
}

int rf_decompress(const unsigned char *payload, long nbytes, int wtype, Real *block)
{
    WaveletCompressor *c = new WaveletCompressor;
    std::memcpy(c->compressed_data(), payload, (size_t)nbytes);
    c->decompress(false, (size_t)nbytes, wtype, (Real(*)[32][32])block);
    delete c;
    return 0;
}

void rf_fwt(Real *block, int wtype)
{
    WaveletsOnInterval::FullTransform<32> *f = new WaveletsOnInterval::FullTransform<32>;
    std::memcpy(&f->data[0][0][0], block, sizeof(Real) * 32768);
    f->fwt(wtype);
    std::memcpy(block, &f->data[0][0][0], sizeof(Real) * 32768);
    delete f;
}

void rf_iwt(Real *block, int wtype)
{
    WaveletsOnInterval::FullTransform<32> *f = new WaveletsOnInterval::FullTransform<32>;
    std::memcpy(&f->data[0][0][0], block, sizeof(Real) * 32768);
    f->iwt(wtype);
    std::memcpy(block, &f->data[0][0][0], sizeof(Real) * 32768);
    delete f;
}

// The reference writer's block loop over a whole HDF5-ordered field
// [nx][ny][nz] (SerializerIO_WaveletCompression_MPI_Simple.h:192-240: an
// OpenMP loop over blocks, each gathered into uncompressed_data() and
// compressed), blocks x fastest.  One compressor object per thread; payloads
// are not kept, only their sizes.  Returns the payload bytes in all.
long rf_field_compress(const Real *field, int nx, int ny, int nz, float threshold, int wtype, int nthreads,
                       long *sizes)
{
    const int bx = nx / 32, by = ny / 32, bz = nz / 32, nb = bx * by * bz;
    long total = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : total)
    {
        WaveletCompressor *c = new WaveletCompressor;
#pragma omp for schedule(static)
        for (int b = 0; b < nb; ++b) {
            const int ix = b % bx, iy = (b / bx) % by, iz = b / (bx * by);
            Real *buf = &c->uncompressed_data()[0][0][0];
            for (int z = 0; z < 32; ++z)
                for (int y = 0; y < 32; ++y)
                    for (int x = 0; x < 32; ++x)
                        buf[x + 32 * (y + 32 * z)] =
                            field[((size_t)(32 * ix + x) * ny + (32 * iy + y)) * nz + (32 * iz + z)];
            const size_t n = c->compress(threshold, false, wtype);
            if (sizes) sizes[b] = (long)n;
            total += (long)n;
        }
        delete c;
    }
    return total;
}

// The reference reader's per-block decode (Reader_WaveletCompression.h:
// 700-754: decompress, then scatter into the field) for every record
// [int32 nbytes][payload] of a stream at offsets[b], OpenMP over blocks.
void rf_field_decompress(const unsigned char *records, const long long *offsets, int nx, int ny, int nz, int wtype,
                         int nthreads, Real *field)
{
    const int bx = nx / 32, by = ny / 32, bz = nz / 32, nb = bx * by * bz;
#pragma omp parallel num_threads(nthreads)
    {
        WaveletCompressor *c = new WaveletCompressor;
        Real(*out)[32][32] = new Real[32][32][32];
#pragma omp for schedule(static)
        for (int b = 0; b < nb; ++b) {
            int n;
            std::memcpy(&n, records + offsets[b], 4);
            std::memcpy(c->compressed_data(), records + offsets[b] + 4, (size_t)n);
            c->decompress(false, (size_t)n, wtype, out);
            const int ix = b % bx, iy = (b / bx) % by, iz = b / (bx * by);
            for (int z = 0; z < 32; ++z)
                for (int y = 0; y < 32; ++y)
                    for (int x = 0; x < 32; ++x)
                        field[((size_t)(32 * ix + x) * ny + (32 * iy + y)) * nz + (32 * iz + z)] = out[z][y][x];
        }
        delete[] out;
        delete c;
    }
}

void rf_row(Real *row, int n, int wtype, int forward)
{
    using WaveletsOnInterval::WI4;
    if (n == 8) { if (forward) WI4::transform<8, true>(row, wtype); else WI4::transform<8, false>(row, wtype); }
    if (n == 16) { if (forward) WI4::transform<16, true>(row, wtype); else WI4::transform<16, false>(row, wtype); }
    if (n == 32) { if (forward) WI4::transform<32, true>(row, wtype); else WI4::transform<32, false>(row, wtype); }
}

} // extern "C"
