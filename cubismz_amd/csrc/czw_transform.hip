// czw_transform.hip -- whole-block multilevel transforms (forward, inverse and
// the adjoint of the inverse) on blocks in the kernels' orientation
// O[z][y][x]: binary32 blocks in one launch, each block held in LDS for all
// nine passes; binary64 (a 256 KB block does not fit in LDS) one launch per
// 1D pass.
//
// Used by the differentiable inverse (cubismz_amd.autograd): the inverse
// transform with a frozen survivor mask is linear in the coefficients, and
// its gradient is the adjoint of that linear map.  The pass sequence follows
// the reference's schedule (WaveletsOnInterval.h:240-276 with the axis
// permutation it implies):
//   forward  level 32: x, y, z;  16: z, x, y;   8: y, z, x
//   inverse  level  8: x, z, y;  16: y, x, z;  32: z, y, x
//   adjoint of the inverse: the inverse passes in reverse order, each
//   replaced by its transpose.
// Coefficients stay in O orientation: the reference's stored ("physical")
// coefficient array is the piecewise axis permutation of it documented in
// czw_kernels.hip.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/czw.h"
#include "czw_wavelet.hpp"

namespace czw {

enum { OP_FWD = 0, OP_INV = 1, OP_INV_ADJ = 2 };

// Rows of length N along axis `ax` (0: z, 1: y, 2: x) inside the [0,N)^3
// corner of every block; one thread per row.
template <typename R, int WT, int OP, int N>
__global__ __launch_bounds__(256) void k_rowpass(R *__restrict__ blocks, int nblocks, int ax)
{
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (long long)nblocks * N * N) return;
    const int b = (int)(r / (N * N)), i = (int)((r / N) % N), j = (int)(r % N);
    const int stride = ax == 0 ? 1024 : (ax == 1 ? 32 : 1);
    // the two other coordinates, slower first
    const int si = ax == 0 ? 32 : 1024, sj = ax == 2 ? 32 : 1;
    R *p = blocks + (size_t)b * 32768 + (size_t)i * si + (size_t)j * sj;
    R v[N];
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = p[k * stride];
    if constexpr (OP == OP_FWD) fwd_row<N, WT, R>(v);
    else if constexpr (OP == OP_INV) inv_row<N, WT, R>(v);
    else adj_inv_row<N, WT, R>(v);
#pragma unroll
    for (int k = 0; k < N; ++k) p[k * stride] = v[k];
}

// Binary32: a persistent grid, one block per workgroup at a time.  The block
// is read once into LDS (O[z][y][x] at z*TSP + y*TRP + x, the padding that
// makes rows along x, y and z conflict-free), all nine passes of the
// schedule run there with the same row functions and order as k_rowpass,
// and the block is written back once: 256 KB of HBM traffic per block
// instead of nine times that.  The workgroup's next block is loaded into
// registers while the current one is transformed.
constexpr int TRP = 33, TSP = 32 * 33 + 1, TT = 512;

template <int WT, int OP, int N>
CZW_DI void lds_pass(float *A, int t, int ax)
{
    const int stride = ax == 0 ? TSP : (ax == 1 ? TRP : 1);
    const int si = ax == 0 ? TRP : TSP, sj = ax == 2 ? TRP : 1;  // the other two coordinates, slower first
    constexpr int RPT = (N * N + TT - 1) / TT;                   // rows per thread
    float v[RPT][N];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {  // every row loaded before the first is written back
        const int r = t + TT * q;
        if (r < N * N) {
            const float *p = A + (r / N) * si + (r % N) * sj;
#pragma unroll
            for (int k = 0; k < N; ++k) v[q][k] = p[k * stride];
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = t + TT * q;
        if (r < N * N) {
            if constexpr (OP == OP_FWD) fwd_row<N, WT, float>(v[q]);
            else if constexpr (OP == OP_INV) inv_row<N, WT, float>(v[q]);
            else adj_inv_row<N, WT, float>(v[q]);
            float *p = A + (r / N) * si + (r % N) * sj;
#pragma unroll
            for (int k = 0; k < N; ++k) p[k * stride] = v[q][k];
        }
    }
    __syncthreads();
}

// The nine passes of one block in A (ends with a barrier).
template <int WT, int OP>
CZW_DI void transform_block(float *A, int t)
{
    enum { Z = 0, Y = 1, X = 2 };
    if constexpr (OP == OP_FWD) {  // level 32: x, y, z; 16: z, x, y; 8: y, z, x
        lds_pass<WT, OP, 32>(A, t, X); lds_pass<WT, OP, 32>(A, t, Y); lds_pass<WT, OP, 32>(A, t, Z);
        lds_pass<WT, OP, 16>(A, t, Z); lds_pass<WT, OP, 16>(A, t, X); lds_pass<WT, OP, 16>(A, t, Y);
        lds_pass<WT, OP, 8>(A, t, Y); lds_pass<WT, OP, 8>(A, t, Z); lds_pass<WT, OP, 8>(A, t, X);
    } else if constexpr (OP == OP_INV) {  // level 8: x, z, y; 16: y, x, z; 32: z, y, x
        lds_pass<WT, OP, 8>(A, t, X); lds_pass<WT, OP, 8>(A, t, Z); lds_pass<WT, OP, 8>(A, t, Y);
        lds_pass<WT, OP, 16>(A, t, Y); lds_pass<WT, OP, 16>(A, t, X); lds_pass<WT, OP, 16>(A, t, Z);
        lds_pass<WT, OP, 32>(A, t, Z); lds_pass<WT, OP, 32>(A, t, Y); lds_pass<WT, OP, 32>(A, t, X);
    } else {  // the inverse's passes in reverse order, each transposed
        lds_pass<WT, OP, 32>(A, t, X); lds_pass<WT, OP, 32>(A, t, Y); lds_pass<WT, OP, 32>(A, t, Z);
        lds_pass<WT, OP, 16>(A, t, Z); lds_pass<WT, OP, 16>(A, t, X); lds_pass<WT, OP, 16>(A, t, Y);
        lds_pass<WT, OP, 8>(A, t, Y); lds_pass<WT, OP, 8>(A, t, Z); lds_pass<WT, OP, 8>(A, t, X);
    }
}

template <int WT, int OP>
__global__ __launch_bounds__(TT) void k_transform_lds_f32(float *__restrict__ blocks, int nblocks)
{
    __shared__ float A[32 * TSP];
    const int t = threadIdx.x;
    constexpr int PER = 32768 / TT;
    float nxt[PER];
    auto load = [&](int bb) {
        const float *g = blocks + (size_t)bb * 32768;
#pragma unroll
        for (int k = 0; k < PER; ++k) nxt[k] = g[t + TT * k];
    };
    if ((int)blockIdx.x < nblocks) load(blockIdx.x);
    for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = t + TT * k;
            A[(e >> 10) * TSP + ((e >> 5) & 31) * TRP + (e & 31)] = nxt[k];
        }
        __syncthreads();
        if (b + (int)gridDim.x < nblocks) load(b + gridDim.x);
        transform_block<WT, OP>(A, t);
        float *g = blocks + (size_t)b * 32768;
#pragma unroll 8
        for (int k = 0; k < PER; ++k) {
            const int e = t + TT * k;
            g[e] = A[(e >> 10) * TSP + ((e >> 5) & 31) * TRP + (e & 31)];
        }
        __syncthreads();  // A is read before the next block is written
    }
}

// Workgroups of k_transform_lds_f32 resident on the current device (its LDS
// block allows one per CU).
static int transform_groups(const void *fn)
{
    static int res[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (res[dev] == 0) {
        int cus = 0, per = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, TT, 0) != hipSuccess || per <= 0) per = 1;
        res[dev] = cus * per;
    }
    return res[dev];
}

template <int WT>
static void transform_lds_f32(float *blocks, int nblocks, int op, hipStream_t s)
{
    const void *fn = (const void *)k_transform_lds_f32<WT, OP_FWD>;
    const dim3 grid(std::min(nblocks, transform_groups(fn))), blk(TT);
    if (op == OP_FWD) hipLaunchKernelGGL((k_transform_lds_f32<WT, OP_FWD>), grid, blk, 0, s, blocks, nblocks);
    if (op == OP_INV) hipLaunchKernelGGL((k_transform_lds_f32<WT, OP_INV>), grid, blk, 0, s, blocks, nblocks);
    if (op == OP_INV_ADJ) hipLaunchKernelGGL((k_transform_lds_f32<WT, OP_INV_ADJ>), grid, blk, 0, s, blocks, nblocks);
}

template <typename R, int WT, int OP>
static void pass(R *blocks, int nblocks, int n, int ax, hipStream_t s)
{
    const long long rows = (long long)nblocks * n * n;
    const dim3 grid((unsigned)((rows + 255) / 256)), blk(256);
    if (n == 32) hipLaunchKernelGGL((k_rowpass<R, WT, OP, 32>), grid, blk, 0, s, blocks, nblocks, ax);
    if (n == 16) hipLaunchKernelGGL((k_rowpass<R, WT, OP, 16>), grid, blk, 0, s, blocks, nblocks, ax);
    if (n == 8) hipLaunchKernelGGL((k_rowpass<R, WT, OP, 8>), grid, blk, 0, s, blocks, nblocks, ax);
}

enum { AX_Z = 0, AX_Y = 1, AX_X = 2 };

template <typename R, int WT>
static void transform(R *blocks, int nblocks, int op, hipStream_t s)
{
    if constexpr (sizeof(R) == 4) {
        transform_lds_f32<WT>(blocks, nblocks, op, s);
        return;
    }
    if (op == OP_FWD) {
        const int sched[3][3] = {{AX_X, AX_Y, AX_Z}, {AX_Z, AX_X, AX_Y}, {AX_Y, AX_Z, AX_X}};
        for (int l = 0; l < 3; ++l)
            for (int k = 0; k < 3; ++k) pass<R, WT, OP_FWD>(blocks, nblocks, 32 >> l, sched[l][k], s);
    } else {
        const int sched[3][3] = {{AX_X, AX_Z, AX_Y}, {AX_Y, AX_X, AX_Z}, {AX_Z, AX_Y, AX_X}};  // levels 8, 16, 32
        if (op == OP_INV) {
            for (int l = 0; l < 3; ++l)
                for (int k = 0; k < 3; ++k) pass<R, WT, OP_INV>(blocks, nblocks, 8 << l, sched[l][k], s);
        } else {
            for (int l = 2; l >= 0; --l)
                for (int k = 2; k >= 0; --k) pass<R, WT, OP_INV_ADJ>(blocks, nblocks, 8 << l, sched[l][k], s);
        }
    }
}

template <typename R>
static int transform_t(R *blocks, int nblocks, int wtype, int op, hipStream_t s)
{
    switch (wtype) {
    case 1: transform<R, 1>(blocks, nblocks, op, s); break;
    case 2: transform<R, 2>(blocks, nblocks, op, s); break;
    case 0:
    case 3: transform<R, 3>(blocks, nblocks, op, s); break;
    default: return CZW_EINVAL;
    }
    return hipGetLastError() == hipSuccess ? CZW_OK : CZW_EHIP;
}

}  // namespace czw

extern "C" int czw_transform_blocks(void *blocks, int nblocks, int prec, int wtype, int op, hipStream_t stream)
{
    if (!blocks || nblocks <= 0 || op < 0 || op > 2) return CZW_EINVAL;
    if (prec == 4) return czw::transform_t(static_cast<float *>(blocks), nblocks, wtype, op, stream);
    if (prec == 8) return czw::transform_t(static_cast<double *>(blocks), nblocks, wtype, op, stream);
    return CZW_EINVAL;
}
