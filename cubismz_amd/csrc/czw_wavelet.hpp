// czw_wavelet.hpp -- register-resident one-level 1D wavelet transforms on rows
// of 32, 16 or 8 values, for the gfx950 kernels.
//
// Arithmetic follows the reference stencils exactly
// (Compressor/source/WaveletsOnInterval.h:33-201):
//  * every predictor is a sum of products of Real values by binary64
//    constants, evaluated left to right in binary64 and rounded to Real on
//    return;
//  * differences, sums and the 0.5 scalings around the predictors are Real
//    operations.
// For Real = float every product of a float by these constants (at most 6
// significant bits) is exact in binary64, so fma(k, x, t) rounds exactly the
// same value as the reference's t + k*x; the float path uses fma to save one
// instruction per term.  For Real = double the products are inexact and the
// double path keeps separate multiplies and adds (the library is compiled
// with -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>

namespace czw {

#define CZW_DI __device__ __forceinline__

template <typename R> struct Pred;

// Stencil constants of the binary32 path, read from constant memory so that
// they live in scalar registers: the fused multiply-adds can then use the
// three-address v_fma_f64 form instead of copying the shared addend.
// Values are exactly the reference's binary64 literals.
enum {
    K_0P0625 = 0, K_M0P0625, K_0P3125, K_M0P3125, K_0P5625, K_0P9375, K_1P3125, K_M2P1875, K_2P1875,
    K_0P125, K_M0P125, K_0P5, K_M0P5, K_0P625, K_1P375, K_COUNT
};
static __constant__ double czw_kc[K_COUNT] = {0.0625, -0.0625, 0.3125, -0.3125, 0.5625, 0.9375, 1.3125, -2.1875, 2.1875,
                                       0.125,  -0.125,  0.5,    -0.5,    0.625,  1.375};

template <> struct Pred<float> {
    typedef float R;
    static CZW_DI double K(int i) { return czw_kc[i]; }
    static CZW_DI float rnd(double t) { return (float)t; }
    static CZW_DI double F(double k, R x, double t) { return __builtin_fma(k, (double)x, t); }
    // wtype 1/2: WaveletsOnInterval.h:33-51
    static CZW_DI float p4_first(R a, R b, R c, R e)
    {
        double t = K(K_0P3125) * (double)a;
        t = F(K(K_0P9375), b, t);
        t = F(K(K_M0P3125), c, t);
        return rnd(F(K(K_0P0625), e, t));
    }
    static CZW_DI float p4_mid(R a, R b, R c, R e)
    {
        double t = K(K_M0P0625) * (double)a;
        t = F(K(K_0P5625), b, t);
        t = F(K(K_0P5625), c, t);
        return rnd(F(K(K_M0P0625), e, t));
    }
    static CZW_DI float p4_otl(R a, R b, R c, R e)
    {
        double t = K(K_0P0625) * (double)a;
        t = F(K(K_M0P3125), b, t);
        t = F(K(K_0P9375), c, t);
        return rnd(F(K(K_0P3125), e, t));
    }
    static CZW_DI float p4_last(R a, R b, R c, R e)
    {
        double t = K(K_M0P3125) * (double)a;
        t = F(K(K_1P3125), b, t);
        t = F(K(K_M2P1875), c, t);
        return rnd(F(K(K_2P1875), e, t));
    }
    // wtype 3: WaveletsOnInterval.h:55-83
    static CZW_DI float q0_first(R a, R b, R c)
    {
        const double t = F(K(K_M0P5), b, K(K_1P375) * (double)a);
        return rnd(F(K(K_0P125), c, t));
    }
    static CZW_DI float q1_first(R a, R b, R c)
    {
        const double t = F(K(K_0P5), b, K(K_0P625) * (double)a);
        return rnd(F(K(K_M0P125), c, t));
    }
    static CZW_DI float q0_mid(R a, R b, R c) { return rnd(F(K(K_M0P125), c, F(K(K_0P125), a, (double)b))); }
    static CZW_DI float q1_mid(R a, R b, R c) { return rnd(F(K(K_0P125), c, F(K(K_M0P125), a, (double)b))); }
    static CZW_DI float q0_last(R a, R b, R c)
    {
        const double t = F(K(K_0P5), b, K(K_M0P125) * (double)a);
        return rnd(F(K(K_0P625), c, t));
    }
    static CZW_DI float q1_last(R a, R b, R c)
    {
        const double t = F(K(K_M0P5), b, K(K_0P125) * (double)a);
        return rnd(F(K(K_1P375), c, t));
    }
    // wtype 2 lifting update s +- 0.5*d in binary64 (WaveletsOnInterval.h:124-126,137-139)
    static CZW_DI float lift_add(R s, R d) { return rnd(F(K(K_0P5), d, (double)s)); }
    static CZW_DI float lift_sub(R s, R d) { return rnd(F(K(K_M0P5), d, (double)s)); }
};

template <> struct Pred<double> {
    typedef double R;
    static CZW_DI double p4_first(R a, R b, R c, R e) { return ((0.3125 * a + 0.9375 * b) - 0.3125 * c) + 0.0625 * e; }
    static CZW_DI double p4_mid(R a, R b, R c, R e) { return ((-0.0625 * a + 0.5625 * b) + 0.5625 * c) - 0.0625 * e; }
    static CZW_DI double p4_otl(R a, R b, R c, R e) { return ((0.0625 * a - 0.3125 * b) + 0.9375 * c) + 0.3125 * e; }
    static CZW_DI double p4_last(R a, R b, R c, R e) { return ((-0.3125 * a + 1.3125 * b) - 2.1875 * c) + 2.1875 * e; }
    static CZW_DI double q0_first(R a, R b, R c) { return (1.375 * a - 0.5 * b) + 0.125 * c; }
    static CZW_DI double q1_first(R a, R b, R c) { return (0.625 * a + 0.5 * b) - 0.125 * c; }
    static CZW_DI double q0_mid(R a, R b, R c) { return (0.125 * a + b) - 0.125 * c; }
    static CZW_DI double q1_mid(R a, R b, R c) { return (-0.125 * a + b) + 0.125 * c; }
    static CZW_DI double q0_last(R a, R b, R c) { return (-0.125 * a + 0.5 * b) + 0.625 * c; }
    static CZW_DI double q1_last(R a, R b, R c) { return (0.125 * a - 0.5 * b) + 1.375 * c; }
    static CZW_DI double lift_add(R s, R d) { return s + 0.5 * d; }
    static CZW_DI double lift_sub(R s, R d) { return s - 0.5 * d; }
};

// One forward level on d[0..N): result [scalings | details]
// (WaveletsOnInterval.h:99-130 for wtype 1/2, :156-180 for wtype 0/3).
template <int N, int WT, typename R>
CZW_DI void fwd_row(R *d)
{
    typedef Pred<R> P;
    constexpr int H = N / 2;
    R s[H], det[H];
    if constexpr (WT == 1 || WT == 2) {
        det[0] = d[1] - P::p4_first(d[0], d[2], d[4], d[6]);
#pragma unroll
        for (int i = 1; i < H - 2; ++i)
            det[i] = d[2 * i + 1] - P::p4_mid(d[2 * i - 2], d[2 * i], d[2 * i + 2], d[2 * i + 4]);
        det[H - 2] = d[N - 3] - P::p4_otl(d[N - 8], d[N - 6], d[N - 4], d[N - 2]);
        det[H - 1] = d[N - 1] - P::p4_last(d[N - 8], d[N - 6], d[N - 4], d[N - 2]);
#pragma unroll
        for (int i = 0; i < H; ++i) s[i] = (WT == 2) ? P::lift_add(d[2 * i], det[i]) : d[2 * i];
    } else {
#pragma unroll
        for (int i = 0; i < H; ++i) s[i] = (R)0.5 * (d[2 * i] + d[2 * i + 1]);
        det[0] = (R)0.5 * ((d[1] - d[0]) - (P::q1_first(s[0], s[1], s[2]) - P::q0_first(s[0], s[1], s[2])));
#pragma unroll
        for (int i = 1; i < H - 1; ++i)
            det[i] = (R)0.5 * ((d[2 * i + 1] - d[2 * i]) -
                               (P::q1_mid(s[i - 1], s[i], s[i + 1]) - P::q0_mid(s[i - 1], s[i], s[i + 1])));
        det[H - 1] = (R)0.5 * ((d[N - 1] - d[N - 2]) -
                               (P::q1_last(s[H - 3], s[H - 2], s[H - 1]) - P::q0_last(s[H - 3], s[H - 2], s[H - 1])));
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
        d[i] = s[i];
        d[H + i] = det[i];
    }
}

// One inverse level (WaveletsOnInterval.h:131-152 for wtype 1/2, :181-198 for 0/3).
template <int N, int WT, typename R>
CZW_DI void inv_row(R *d)
{
    typedef Pred<R> P;
    constexpr int H = N / 2;
    R s[H], det[H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
        s[i] = d[i];
        det[i] = d[H + i];
    }
    if constexpr (WT == 1 || WT == 2) {
        if constexpr (WT == 2) {
#pragma unroll
            for (int i = 0; i < H; ++i) s[i] = P::lift_sub(s[i], det[i]);
        }
#pragma unroll
        for (int i = 0; i < H; ++i) d[2 * i] = s[i];
        d[1] = P::p4_first(s[0], s[1], s[2], s[3]) + det[0];
#pragma unroll
        for (int i = 1; i < H - 2; ++i) d[2 * i + 1] = P::p4_mid(s[i - 1], s[i], s[i + 1], s[i + 2]) + det[i];
        d[N - 3] = P::p4_otl(s[H - 4], s[H - 3], s[H - 2], s[H - 1]) + det[H - 2];
        d[N - 1] = P::p4_last(s[H - 4], s[H - 3], s[H - 2], s[H - 1]) + det[H - 1];
    } else {
        d[0] = P::q0_first(s[0], s[1], s[2]) - det[0];
        d[1] = P::q1_first(s[0], s[1], s[2]) + det[0];
#pragma unroll
        for (int i = 1; i < H - 1; ++i) {
            d[2 * i] = P::q0_mid(s[i - 1], s[i], s[i + 1]) - det[i];
            d[2 * i + 1] = P::q1_mid(s[i - 1], s[i], s[i + 1]) + det[i];
        }
        d[N - 2] = P::q0_last(s[H - 3], s[H - 2], s[H - 1]) - det[H - 1];
        d[N - 1] = P::q1_last(s[H - 3], s[H - 2], s[H - 1]) + det[H - 1];
    }
}

// Adjoint (transpose) of one inverse level: given g = d(loss)/d(out) on the
// N outputs of inv_row, returns d(loss)/d([s | det]) in the same [s | det]
// layout.  The inverse is linear in exact arithmetic with the predictor
// weights of WaveletsOnInterval.h:33-83; this applies the transposed weights
// (transposed lifting).  Evaluated in R.
template <int N, int WT, typename R>
CZW_DI void adj_inv_row(R *g)
{
    constexpr int H = N / 2;
    R gs[H], gd[H];
#pragma unroll
    for (int j = 0; j < H; ++j) gs[j] = 0;
    if constexpr (WT == 1 || WT == 2) {
        // out[2i] = s_i; out[2i+1] = P_i(s) + det_i
#pragma unroll
        for (int i = 0; i < H; ++i) {
            gs[i] += g[2 * i];
            gd[i] = g[2 * i + 1];
        }
        const R a = g[1];  // first: s0..s3
        gs[0] += (R)0.3125 * a; gs[1] += (R)0.9375 * a; gs[2] += (R)-0.3125 * a; gs[3] += (R)0.0625 * a;
#pragma unroll
        for (int i = 1; i < H - 2; ++i) {  // middle: s_{i-1}..s_{i+2}
            const R m = g[2 * i + 1];
            gs[i - 1] += (R)-0.0625 * m; gs[i] += (R)0.5625 * m; gs[i + 1] += (R)0.5625 * m; gs[i + 2] += (R)-0.0625 * m;
        }
        const R o = g[N - 3], l = g[N - 1];  // one-to-last and last: s_{H-4}..s_{H-1}
        gs[H - 4] += (R)0.0625 * o + (R)-0.3125 * l;
        gs[H - 3] += (R)-0.3125 * o + (R)1.3125 * l;
        gs[H - 2] += (R)0.9375 * o + (R)-2.1875 * l;
        gs[H - 1] += (R)0.3125 * o + (R)2.1875 * l;
        if constexpr (WT == 2) {  // the inverse first applies s -= det/2
#pragma unroll
            for (int i = 0; i < H; ++i) gd[i] -= (R)0.5 * gs[i];
        }
    } else {
        // out[2i] = Q0_i(s) - det_i; out[2i+1] = Q1_i(s) + det_i
#pragma unroll
        for (int i = 0; i < H; ++i) gd[i] = g[2 * i + 1] - g[2 * i];
        {
            const R e = g[0], o = g[1];
            gs[0] += (R)1.375 * e + (R)0.625 * o;
            gs[1] += (R)-0.5 * e + (R)0.5 * o;
            gs[2] += (R)0.125 * e + (R)-0.125 * o;
        }
#pragma unroll
        for (int i = 1; i < H - 1; ++i) {
            const R e = g[2 * i], o = g[2 * i + 1];
            gs[i - 1] += (R)0.125 * e + (R)-0.125 * o;
            gs[i] += e + o;
            gs[i + 1] += (R)-0.125 * e + (R)0.125 * o;
        }
        {
            const R e = g[N - 2], o = g[N - 1];
            gs[H - 3] += (R)-0.125 * e + (R)0.125 * o;
            gs[H - 2] += (R)0.5 * e + (R)-0.5 * o;
            gs[H - 1] += (R)0.625 * e + (R)1.375 * o;
        }
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
        g[i] = gs[i];
        g[H + i] = gd[i];
    }
}

template <int N, bool FWD, int WT, typename R>
CZW_DI void row(R *d)
{
    if constexpr (FWD) fwd_row<N, WT, R>(d);
    else inv_row<N, WT, R>(d);
}

} // namespace czw
