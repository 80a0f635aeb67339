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
