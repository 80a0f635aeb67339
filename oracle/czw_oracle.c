/*
 * czw_oracle.c -- CPU restatement of the CubismZ wavelet block compressor
 * (Compressor/source/WaveletsOnInterval.h, FullWaveletTransform.h,
 * WaveletCompressor.cpp), in both reference precisions.
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg, always as the checker or the CPU
 * baseline, never as the product path.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -fPIC -shared (see oracle/Makefile).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define REAL float
#define SFX _f32
#define CZO_ABS(v) fabsf(v)
#include "czw_oracle_impl.h"
#undef CZO_ABS
#undef SFX
#undef REAL

#define REAL double
#define SFX _f64
#define CZO_ABS(v) fabs(v)
#include "czw_oracle_impl.h"
#undef CZO_ABS
#undef SFX
#undef REAL

int czo_version(void) { return 1; }

#ifdef _OPENMP
#include <omp.h>
void czo_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int czo_get_threads(void) { return omp_get_max_threads(); }
#else
void czo_set_threads(int n) { (void)n; }
int czo_get_threads(void) { return 1; }
#endif

/* a17: cz2diff error metrics (Tools/cz2diff.cpp:103-261), accumulated in the
 * reference's order: blocks b = x + NBX*(y + NBY*z), then zb, yb, xb inside a
 * block; f1 = decoded, f2 = reference.  Fields are HDF5-ordered [X][Y][Z].
 * out = {e_inf, e_1, e_2, n_inf, n_1, n_2, min, max, n}. */
#define CZO_DIFF(SFXN, T)                                                                                 \
    void czo_diff_metrics##SFXN(const T *f1, const T *f2, int NX, int NY, int NZ, double *out)            \
    {                                                                                                     \
        const int NBX = NX / 32, NBY = NY / 32, NBZ = NZ / 32;                                            \
        double e_inf = 0, e_1 = 0, e_2 = 0, n_inf = 0, n_1 = 0, n_2 = 0;                                  \
        double maxdata = -1.7976931348623157e308, mindata = 1.7976931348623157e308;                       \
        long n = 0;                                                                                       \
        for (int b = 0; b < NBX * NBY * NBZ; ++b) {                                                       \
            const int z = b / (NBY * NBX), y = (b / NBX) % NBY, x = b % NBX;                              \
            for (int zb = 0; zb < 32; zb++)                                                               \
                for (int yb = 0; yb < 32; yb++)                                                           \
                    for (int xb = 0; xb < 32; xb++) {                                                     \
                        const size_t gi = ((size_t)(32 * x + xb) * NY + (size_t)(32 * y + yb)) * NZ +     \
                                          (size_t)(32 * z + zb);                                          \
                        const double a = (double)f1[gi], c = (double)f2[gi];                              \
                        if (c > maxdata) maxdata = c;                                                     \
                        if (c < mindata) mindata = c;                                                     \
                        const double v = fabs(c);                                                         \
                        if (v > n_inf) n_inf = v;                                                         \
                        n_1 += v;                                                                         \
                        n_2 += v * v;                                                                     \
                        const double err = fabs(a - c);                                                   \
                        if (err > e_inf) e_inf = err;                                                     \
                        e_1 += err;                                                                       \
                        e_2 += err * err;                                                                 \
                        n++;                                                                              \
                    }                                                                                     \
        }                                                                                                 \
        out[0] = e_inf; out[1] = e_1; out[2] = e_2; out[3] = n_inf; out[4] = n_1; out[5] = n_2;           \
        out[6] = mindata; out[7] = maxdata; out[8] = (double)n;                                           \
    }
CZO_DIFF(_f32, float)
CZO_DIFF(_f64, double)
