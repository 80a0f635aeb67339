/*
 * CPU restatement of CubismZ's per-block wavelet compressor, instantiated
 * twice by czw_oracle.c: once with REAL = float (the reference "float"
 * precision build) and once with REAL = double (the "double" build).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle that the HIP kernels are
 * checked against; it is never part of the product path.
 *
 * Arithmetic contract (reference build: g++ -O3 -DNDEBUG, x86-64 SSE, no
 * -march, hence no FMA contraction):
 *   - stencil constants are binary64 literals, so every predictor is evaluated
 *     in binary64 and rounded to REAL when it is returned
 *     (Compressor/source/WaveletsOnInterval.h:33-83);
 *   - the differences/sums around the predictors are REAL operations;
 *   - this file must be compiled with -ffp-contract=off.
 *
 * Macros expected from the includer: REAL, SFX (function-name suffix).
 */

#define CZO_CAT2(a, b) a##b
#define CZO_CAT(a, b) CZO_CAT2(a, b)
#define CZO_FN(name) CZO_CAT(name, SFX)

/* ---- a1: 4-point interpolating predictors (WaveletsOnInterval.h:33-51) ---- */
static inline REAL CZO_FN(p4_first)(REAL f0, REAL f1, REAL f2, REAL f3)
{
    return (REAL)(((0.3125 * (double)f0 + 0.9375 * (double)f1) - 0.3125 * (double)f2) + 0.0625 * (double)f3);
}
static inline REAL CZO_FN(p4_mid)(REAL f0, REAL f1, REAL f2, REAL f3)
{
    return (REAL)(((-0.0625 * (double)f0 + 0.5625 * (double)f1) + 0.5625 * (double)f2) - 0.0625 * (double)f3);
}
static inline REAL CZO_FN(p4_onetolast)(REAL f0, REAL f1, REAL f2, REAL f3)
{
    return (REAL)(((0.0625 * (double)f0 - 0.3125 * (double)f1) + 0.9375 * (double)f2) + 0.3125 * (double)f3);
}
static inline REAL CZO_FN(p4_last)(REAL f0, REAL f1, REAL f2, REAL f3)
{
    return (REAL)(((-0.3125 * (double)f0 + 1.3125 * (double)f1) - 2.1875 * (double)f2) + 2.1875 * (double)f3);
}

/* ---- a2: 3-point average-interpolating predictors (WaveletsOnInterval.h:55-83).
 * q0* predicts the left half-cell, q1* the right half-cell. ---- */
static inline REAL CZO_FN(q0_first)(REAL a, REAL b, REAL c)
{ return (REAL)((1.375 * (double)a - 0.5 * (double)b) + 0.125 * (double)c); }
static inline REAL CZO_FN(q1_first)(REAL a, REAL b, REAL c)
{ return (REAL)((0.625 * (double)a + 0.5 * (double)b) - 0.125 * (double)c); }
static inline REAL CZO_FN(q0_mid)(REAL a, REAL b, REAL c)
{ return (REAL)((0.125 * (double)a + (double)b) - 0.125 * (double)c); }
static inline REAL CZO_FN(q1_mid)(REAL a, REAL b, REAL c)
{ return (REAL)((-0.125 * (double)a + (double)b) + 0.125 * (double)c); }
static inline REAL CZO_FN(q0_last)(REAL a, REAL b, REAL c)
{ return (REAL)((-0.125 * (double)a + 0.5 * (double)b) + 0.625 * (double)c); }
static inline REAL CZO_FN(q1_last)(REAL a, REAL b, REAL c)
{ return (REAL)((0.125 * (double)a - 0.5 * (double)b) + 1.375 * (double)c); }

/* ---- a3: one level of the 1D transform on a contiguous row of n values
 * (n in {8,16,32}); output layout [scalings(n/2) | details(n/2)].
 * Follows WaveletsOnInterval.h:85-201 (wtype 1/2: :97-153, wtype 0/3: :154-200). ---- */
static void CZO_FN(czo_row_fwd)(REAL *d, int n, int wtype)
{
    const int h = n / 2;
    REAL s[16], det[16];
    if (wtype == 1 || wtype == 2) {
        det[0] = d[1] - CZO_FN(p4_first)(d[0], d[2], d[4], d[6]);
        for (int i = 1; i < h - 2; ++i)
            det[i] = d[2 * i + 1] - CZO_FN(p4_mid)(d[2 * i - 2], d[2 * i], d[2 * i + 2], d[2 * i + 4]);
        det[h - 2] = d[n - 3] - CZO_FN(p4_onetolast)(d[n - 8], d[n - 6], d[n - 4], d[n - 2]);
        det[h - 1] = d[n - 1] - CZO_FN(p4_last)(d[n - 8], d[n - 6], d[n - 4], d[n - 2]);
        for (int i = 0; i < h; ++i) s[i] = d[2 * i];
        if (wtype == 2)
            for (int i = 0; i < h; ++i) s[i] = (REAL)((double)s[i] + 0.5 * (double)det[i]);
    } else {
        for (int i = 0; i < h; ++i) s[i] = (REAL)(0.5 * (double)(REAL)(d[2 * i] + d[2 * i + 1]));
        {
            const REAL pr = CZO_FN(q1_first)(s[0], s[1], s[2]) - CZO_FN(q0_first)(s[0], s[1], s[2]);
            det[0] = (REAL)(0.5 * (double)(REAL)((REAL)(d[1] - d[0]) - pr));
        }
        for (int i = 1; i < h - 1; ++i) {
            const REAL pr = CZO_FN(q1_mid)(s[i - 1], s[i], s[i + 1]) - CZO_FN(q0_mid)(s[i - 1], s[i], s[i + 1]);
            det[i] = (REAL)(0.5 * (double)(REAL)((REAL)(d[2 * i + 1] - d[2 * i]) - pr));
        }
        {
            const REAL pr = CZO_FN(q1_last)(s[h - 3], s[h - 2], s[h - 1]) - CZO_FN(q0_last)(s[h - 3], s[h - 2], s[h - 1]);
            det[h - 1] = (REAL)(0.5 * (double)(REAL)((REAL)(d[n - 1] - d[n - 2]) - pr));
        }
    }
    for (int i = 0; i < h; ++i) { d[i] = s[i]; d[h + i] = det[i]; }
}

static void CZO_FN(czo_row_inv)(REAL *d, int n, int wtype)
{
    const int h = n / 2;
    REAL s[16], det[16];
    for (int i = 0; i < h; ++i) { s[i] = d[i]; det[i] = d[h + i]; }
    if (wtype == 1 || wtype == 2) {
        if (wtype == 2)
            for (int i = 0; i < h; ++i) s[i] = (REAL)((double)s[i] - 0.5 * (double)det[i]);
        for (int i = 0; i < h; ++i) d[2 * i] = s[i];
        d[1] = CZO_FN(p4_first)(s[0], s[1], s[2], s[3]) + det[0];
        for (int i = 1; i < h - 2; ++i)
            d[2 * i + 1] = CZO_FN(p4_mid)(s[i - 1], s[i], s[i + 1], s[i + 2]) + det[i];
        d[n - 3] = CZO_FN(p4_onetolast)(s[h - 4], s[h - 3], s[h - 2], s[h - 1]) + det[h - 2];
        d[n - 1] = CZO_FN(p4_last)(s[h - 4], s[h - 3], s[h - 2], s[h - 1]) + det[h - 1];
    } else {
        d[0] = CZO_FN(q0_first)(s[0], s[1], s[2]) - det[0];
        d[1] = CZO_FN(q1_first)(s[0], s[1], s[2]) + det[0];
        for (int i = 1; i < h - 1; ++i) {
            d[2 * i] = CZO_FN(q0_mid)(s[i - 1], s[i], s[i + 1]) - det[i];
            d[2 * i + 1] = CZO_FN(q1_mid)(s[i - 1], s[i], s[i + 1]) + det[i];
        }
        d[n - 2] = CZO_FN(q0_last)(s[h - 3], s[h - 2], s[h - 1]) - det[h - 1];
        d[n - 1] = CZO_FN(q1_last)(s[h - 3], s[h - 2], s[h - 1]) + det[h - 1];
    }
}

void CZO_FN(czo_row)(REAL *d, int n, int wtype, int forward)
{
    if (forward) CZO_FN(czo_row_fwd)(d, n, wtype);
    else CZO_FN(czo_row_inv)(d, n, wtype);
}

/* ---- a4: one 3D level on the [0,bs)^3 corner of a 32^3 array a[a0][a1][a2],
 * with the reference's in-place swaps (WaveletsOnInterval.h:205-276):
 *   forward: rows; swap(a1,a2) per a0-slice; rows; swap(a0,a2); rows
 *   inverse: rows; swap(a0,a2); rows; swap(a1,a2) per a0-slice; rows          ---- */
#define CZO_IDX(a0, a1, a2) ((a2) + 32 * ((a1) + 32 * (a0)))

static void CZO_FN(rows)(REAL *a, int bs, int wtype, int fwd)
{
    for (int i0 = 0; i0 < bs; ++i0)
        for (int i1 = 0; i1 < bs; ++i1)
            CZO_FN(czo_row)(a + CZO_IDX(i0, i1, 0), bs, wtype, fwd);
}

static void CZO_FN(swap12)(REAL *a, int bs)
{
    for (int i0 = 0; i0 < bs; ++i0)
        for (int i = 0; i < bs; ++i)
            for (int j = i + 1; j < bs; ++j) {
                const REAL t = a[CZO_IDX(i0, i, j)];
                a[CZO_IDX(i0, i, j)] = a[CZO_IDX(i0, j, i)];
                a[CZO_IDX(i0, j, i)] = t;
            }
}

static void CZO_FN(swap02)(REAL *a, int bs)
{
    for (int i1 = 0; i1 < bs; ++i1)
        for (int i = 0; i < bs; ++i)
            for (int j = i + 1; j < bs; ++j) {
                const REAL t = a[CZO_IDX(i, i1, j)];
                a[CZO_IDX(i, i1, j)] = a[CZO_IDX(j, i1, i)];
                a[CZO_IDX(j, i1, i)] = t;
            }
}

static void CZO_FN(level)(REAL *a, int bs, int wtype, int fwd)
{
    if (fwd) {
        CZO_FN(rows)(a, bs, wtype, 1);
        CZO_FN(swap12)(a, bs);
        CZO_FN(rows)(a, bs, wtype, 1);
        CZO_FN(swap02)(a, bs);
        CZO_FN(rows)(a, bs, wtype, 1);
    } else {
        CZO_FN(rows)(a, bs, wtype, 0);
        CZO_FN(swap02)(a, bs);
        CZO_FN(rows)(a, bs, wtype, 0);
        CZO_FN(swap12)(a, bs);
        CZO_FN(rows)(a, bs, wtype, 0);
    }
}

/* ---- a5: multilevel engine (FullWaveletTransform.h:44-56, 158-160):
 * forward levels 32,16,8; inverse 8,16,32; the 4^3 corner is never touched. ---- */
void CZO_FN(czo_fwt)(REAL *a, int wtype)
{
    for (int bs = 32; bs >= 8; bs /= 2) CZO_FN(level)(a, bs, wtype, 1);
}

void CZO_FN(czo_iwt)(REAL *a, int wtype)
{
    for (int bs = 8; bs <= 32; bs *= 2) CZO_FN(level)(a, bs, wtype, 0);
}

/* ---- a6: eps test + stream-order survivor packing (FullWaveletTransform.h:59-96,
 * 163-176).  Stream order: the 4^3 corner in a0,a1,a2 order; then for bs = 8,16,32
 * and code = 1..7 the sub-box at (a2,a1,a0) origin h*(code&1, code>>1&1, code>>2&1),
 * h = bs/2, iterated a0,a1,a2 (a2 fastest).  Mask bit p = a2 + 32*(a1 + 32*a0).
 * Returns the survivor count; mask must hold 4096 zeroed bytes. ---- */
static int CZO_FN(threshold)(const REAL *a, REAL eps, unsigned char *mask, REAL *surv)
{
    int ns = 0;
    for (int i0 = 0; i0 < 4; ++i0)
        for (int i1 = 0; i1 < 4; ++i1)
            for (int i2 = 0; i2 < 4; ++i2) {
                const int p = CZO_IDX(i0, i1, i2);
                mask[p >> 3] |= (unsigned char)(1u << (p & 7));
                surv[ns++] = a[p];
            }
    for (int bs = 8; bs <= 32; bs *= 2) {
        const int h = bs / 2;
        for (int code = 1; code < 8; ++code) {
            const int o2 = h * (code & 1), o1 = h * ((code >> 1) & 1), o0 = h * ((code >> 2) & 1);
            for (int i0 = 0; i0 < h; ++i0)
                for (int i1 = 0; i1 < h; ++i1)
                    for (int i2 = 0; i2 < h; ++i2) {
                        const int p = CZO_IDX(o0 + i0, o1 + i1, o2 + i2);
                        const REAL v = a[p];
                        if (CZO_ABS(v) > eps) { /* NaN never survives */
                            mask[p >> 3] |= (unsigned char)(1u << (p & 7));
                            surv[ns++] = v;
                        }
                    }
        }
    }
    return ns;
}

/* ---- a10: inverse of a6 (FullWaveletTransform.h:100-150, 179-189).  The
 * reference pops survivors from the back in exact reverse stream order; walking
 * forward in stream order with a running index yields the same placement. ---- */
static int CZO_FN(load)(REAL *a, const unsigned char *mask, const REAL *surv, int ns)
{
    int k = 0;
    for (int i0 = 0; i0 < 4; ++i0)
        for (int i1 = 0; i1 < 4; ++i1)
            for (int i2 = 0; i2 < 4; ++i2) {
                if (k >= ns) return -1;
                a[CZO_IDX(i0, i1, i2)] = surv[k++];
            }
    for (int bs = 8; bs <= 32; bs *= 2) {
        const int h = bs / 2;
        for (int code = 1; code < 8; ++code) {
            const int o2 = h * (code & 1), o1 = h * ((code >> 1) & 1), o0 = h * ((code >> 2) & 1);
            for (int i0 = 0; i0 < h; ++i0)
                for (int i1 = 0; i1 < h; ++i1)
                    for (int i2 = 0; i2 < h; ++i2) {
                        const int p = CZO_IDX(o0 + i0, o1 + i1, o2 + i2);
                        if ((mask[p >> 3] >> (p & 7)) & 1) {
                            if (k >= ns) return -1;
                            a[p] = surv[k++];
                        } else {
                            a[p] = (REAL)0;
                        }
                    }
        }
    }
    return k == ns ? 0 : -1;
}

/* ---- a11 compress: fwt, threshold, mask (LSB-first bytes, WaveletCompressor.cpp:97-122),
 * zerobits on the first 32-bit word of each survivor (cpp:34-43, 166-173),
 * byte shuffle (cpp:56-73).  block is [z][y][x] (x fastest); out needs
 * 4096 + 32768*sizeof(REAL) bytes.  Returns the payload size. ---- */
size_t CZO_FN(czo_compress)(const REAL *block, float threshold, int wtype, int zerobits,
                            int shuffle, unsigned char *out)
{
    REAL *a = (REAL *)malloc(sizeof(REAL) * 32768);
    memcpy(a, block, sizeof(REAL) * 32768);
    CZO_FN(czo_fwt)(a, wtype);
    memset(out, 0, 4096);
    REAL *surv = (REAL *)malloc(sizeof(REAL) * 32768);
    const REAL eps = (REAL)threshold;
    const int ns = CZO_FN(threshold)(a, eps, out, surv);
    free(a);
    unsigned char *pay = out + 4096;
    memcpy(pay, surv, sizeof(REAL) * (size_t)ns);
    free(surv);
    if (zerobits > 0) {
        const uint32_t keep = ~((1u << zerobits) - 1u);
        for (int i = 0; i < ns; ++i) {
            uint32_t w;
            memcpy(&w, pay + (size_t)i * sizeof(REAL), 4);
            w &= keep;
            memcpy(pay + (size_t)i * sizeof(REAL), &w, 4);
        }
    }
    if (shuffle) {
        const size_t nb = (size_t)ns * sizeof(REAL);
        unsigned char *tmp = (unsigned char *)malloc(nb ? nb : 1);
        for (size_t b = 0, j = 0; b < sizeof(REAL); ++b)
            for (size_t i = 0; i < (size_t)ns; ++i) tmp[j++] = pay[i * sizeof(REAL) + b];
        memcpy(pay, tmp, nb);
        free(tmp);
    }
    return 4096 + sizeof(REAL) * (size_t)ns;
}

/* ---- a11 decompress (WaveletCompressor.cpp:232-259): mask -> popcount,
 * reshuffle (cpp:75-93), load, iwt.  Returns 0, or -1 if nbytes disagrees with
 * the mask popcount. ---- */
int CZO_FN(czo_decompress)(const unsigned char *payload, size_t nbytes, int wtype, int shuffle,
                           REAL *block)
{
    if (nbytes < 4096 || (nbytes - 4096) % sizeof(REAL) != 0) return -1;
    const int ns = (int)((nbytes - 4096) / sizeof(REAL));
    int pop = 0;
    for (int i = 0; i < 4096; ++i) pop += __builtin_popcount(payload[i]);
    if (pop != ns) return -1;
    REAL *surv = (REAL *)malloc(sizeof(REAL) * (ns ? ns : 1));
    const unsigned char *pay = payload + 4096;
    if (shuffle) {
        unsigned char *dst = (unsigned char *)surv;
        for (int i = 0; i < ns; ++i)
            for (size_t b = 0; b < sizeof(REAL); ++b)
                dst[(size_t)i * sizeof(REAL) + b] = pay[b * (size_t)ns + i];
    } else {
        memcpy(surv, pay, sizeof(REAL) * (size_t)ns);
    }
    const int rc = CZO_FN(load)(block, payload, surv, ns);
    free(surv);
    if (rc) return rc;
    CZO_FN(czo_iwt)(block, wtype);
    return 0;
}

/* ---- a13 per-block driver over a field: blocks are gathered from a field
 * stored h5-style f[X][Y][Z] (Z fastest) so that block(ix,iy,iz)[z][y][x] =
 * f[32ix+x][32iy+y][32iz+z] (Tools/Test_IO_Compressed.h:200-212,
 * Types_single.h:80); blocks numbered x fastest (Cubism Grid.h:68).
 * records: [int32 nbytes][payload] per block at rec_off[b] (cpp SerializerIO
 * :237-240).  Uses OpenMP over blocks when compiled with -fopenmp. ---- */
void CZO_FN(czo_gather_block)(const REAL *field, int NX, int NY, int NZ, int ix, int iy, int iz, REAL *blk)
{
    for (int z = 0; z < 32; ++z)
        for (int y = 0; y < 32; ++y)
            for (int x = 0; x < 32; ++x)
                blk[x + 32 * (y + 32 * z)] =
                    field[((size_t)(32 * ix + x) * NY + (size_t)(32 * iy + y)) * NZ + (size_t)(32 * iz + z)];
    (void)NX;
}

void CZO_FN(czo_scatter_block)(REAL *field, int NX, int NY, int NZ, int ix, int iy, int iz, const REAL *blk)
{
    for (int z = 0; z < 32; ++z)
        for (int y = 0; y < 32; ++y)
            for (int x = 0; x < 32; ++x)
                field[((size_t)(32 * ix + x) * NY + (size_t)(32 * iy + y)) * NZ + (size_t)(32 * iz + z)] =
                    blk[x + 32 * (y + 32 * z)];
    (void)NX;
}

/* Compress the blocks [b0, b1) of a field into per-block slots of
 * slot_bytes each; nbytes[b - b0] receives each payload size. */
void CZO_FN(czo_compress_field)(const REAL *field, int NX, int NY, int NZ, int b0, int b1,
                                float threshold, int wtype, int zerobits, int shuffle,
                                unsigned char *slots, size_t slot_bytes, uint32_t *nbytes)
{
    const int bx = NX / 32, by = NY / 32;
#pragma omp parallel for schedule(static)
    for (int b = b0; b < b1; ++b) {
        REAL *blk = (REAL *)malloc(sizeof(REAL) * 32768);
        const int ix = b % bx, iy = (b / bx) % by, iz = b / (bx * by);
        CZO_FN(czo_gather_block)(field, NX, NY, NZ, ix, iy, iz, blk);
        nbytes[b - b0] = (uint32_t)CZO_FN(czo_compress)(blk, threshold, wtype, zerobits, shuffle,
                                                        slots + (size_t)(b - b0) * slot_bytes);
        free(blk);
    }
}

int CZO_FN(czo_decompress_field)(REAL *field, int NX, int NY, int NZ, int b0, int b1, int wtype,
                                 int shuffle, const unsigned char *slots, size_t slot_bytes,
                                 const uint32_t *nbytes)
{
    const int bx = NX / 32, by = NY / 32;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int b = b0; b < b1; ++b) {
        REAL *blk = (REAL *)malloc(sizeof(REAL) * 32768);
        const int ix = b % bx, iy = (b / bx) % by, iz = b / (bx * by);
        if (CZO_FN(czo_decompress)(slots + (size_t)(b - b0) * slot_bytes, nbytes[b - b0], wtype,
                                   shuffle, blk))
            bad = 1;
        else
            CZO_FN(czo_scatter_block)(field, NX, NY, NZ, ix, iy, iz, blk);
        free(blk);
    }
    return bad ? -1 : 0;
}

#undef CZO_IDX
#undef CZO_FN
#undef CZO_CAT
#undef CZO_CAT2
