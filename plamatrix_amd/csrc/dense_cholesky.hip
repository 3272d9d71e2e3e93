// Dense FP64 Cholesky solve on the device: the reference's DenseCpu Schur
// backend (block_schur_dense_solver.h:126-214 factorization, 199-227
// triangular solves, 300-398 residual check) re-designed for gfx950.
//
// Storage: the lower triangle of A in a dense n x n column-major array in the
// workspace, plus the inverses of the 128 x 128 diagonal blocks.
//
// Factorization, right-looking over 256-column panels, each panel as two
// 128-column blocks:
//   k_dc_potrf   one workgroup factors a 128 x 128 diagonal block in LDS
//                (packed lower triangle) and inverts its factor;
//   k_dc_gemm    every other step is a rank-K product on FP64 MFMA
//                (v_mfma_f64_16x16x4f64, 128 x 128 tiles, K staged through
//                LDS 16 columns at a time):
//                  - the panel solve  L21 = A21 L11^-T   (C = P W^T, in place),
//                  - the in-panel update of the second block (C -= P Q^T),
//                  - the trailing update of the lower triangle with K = 256.
// The trailing update carries n^3/3 of the n^3/3 + O(n^2) flops; with K = 256
// each 128 x 128 tile does 32 flops per byte of its read-modify-write, above
// MI355X's FP64 balance point (78.6 TFLOP/s against 8 TB/s).
//
// Solves: forward L y = b and backward L^T x = y by 128-row blocks, each block
// one launch that applies the block's inverse and updates the remaining rows
// (columns-contiguous reads of L).  The residual ||b - A x|| uses A's CSR.
// Every reduction has a fixed order, so repeated solves are bitwise equal.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "dense_cholesky.hpp"
#include "pm_common.hpp"
#include "tuning.hpp"
#include "workspace.hpp"

namespace pm
{

namespace
{

typedef double dc_d4 __attribute__((ext_vector_type(4)));

constexpr int kDcTile = 128;                          // diagonal block and C tile edge
constexpr int kDcPacked = kDcTile * (kDcTile + 1) / 2;  // packed lower triangle
constexpr int kDcKc = 16;                             // K columns staged per step
constexpr int kDcLd = kDcTile + 16;                   // LDS row pitch (bank offset per k row)
constexpr int kDcPanel = 2 * kDcTile;

__device__ __forceinline__ int pk(int i, int j) { return i * (i + 1) / 2 + j; }

// Broadcast lane `src` (a constant) of a double through two v_readlane_b32.
__device__ __forceinline__ double lane_value(double v, int src)
{
    const unsigned long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits & 0xffffffffull), src);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), src);
    return __longlong_as_double((static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                static_cast<unsigned>(lo));
}

__global__ void k_dc_flag_reset(int* fail, unsigned* counter)
{
    *fail = 0;
    *counter = 0u;
}

// Zero the lower triangle (column-major, ld = n) of the dense array.
__global__ __launch_bounds__(kBlock) void k_dc_zero_lower(int64_t n, double* a)
{
    const int64_t total = n * n;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += stride)
    {
        const int64_t r = idx % n, c = idx / n;
        if (r >= c) a[idx] = 0.0;
    }
}

// Scatter the lower-triangle entries of a CSR row into the dense array; one
// thread per row walks its entries in order, so duplicates sum in CSR order.
// Offsets are clamped to [0, nnz] and columns outside [0, r] are skipped, so a
// malformed CSR cannot address memory outside the arrays.
template <typename I>
__global__ __launch_bounds__(kBlock) void k_dc_scatter(int64_t n, int64_t nnz, const I* __restrict__ ro,
                                                      const I* __restrict__ ci, const double* __restrict__ v,
                                                      double* a)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += stride)
    {
        const int64_t begin = max(static_cast<int64_t>(ro[r]), int64_t(0));
        const int64_t end = min(static_cast<int64_t>(ro[r + 1]), nnz);
        for (int64_t p = begin; p < end; ++p)
        {
            const int64_t c = ci[p];
            if (c >= 0 && c <= r) a[r + c * n] += v[p];
        }
    }
}

// Factor the w x w (w <= 128) diagonal block at (c0, c0) in place and write
// the inverse of its factor, zero-padded to 128 x 128 column-major, to inv.
// A pivot that is not positive and finite raises *fail and is replaced by 1
// so that later arithmetic stays finite.
//
// Both steps are blocked by 16 inside LDS (packed lower triangles of L and
// X = L^-1), so the block costs about 40 workgroup barriers:
//   factor, per 16-column step: one wavefront factors the 16 x 16 diagonal
//   block in registers (a row per lane, readlane broadcasts), then one thread per row
//   solves the rows below against it, then the trailing lower triangle takes
//   the rank-16 update;
//   inverse: each wavefront inverts one 16 x 16 diagonal block (a column per
//   lane), then for block distance d = 1..7 every X[j+d][j] is
//   -X[j+d][j+d] sum_{j<=k<j+d} L[j+d][k] X[k][j], through an LDS scratch.
constexpr int kDcSub = 16;
constexpr int kDcSubBlocks = kDcTile / kDcSub;

__global__ __launch_bounds__(1024) void k_dc_potrf(int64_t n, int64_t c0, int w, double* a, double* inv, int* fail)
{
    __shared__ double L[kDcPacked];
    __shared__ double X[kDcPacked];
    __shared__ double U[(kDcSubBlocks - 1) * kDcSub * kDcSub];
    __shared__ int bad;
    if (*fail) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // column j, rows i >= j: thread (i, j) = (tid % 128, tid / 128 + 8 m)
    const int ti = tid & (kDcTile - 1), tj = tid >> 7;
    for (int j = tj; j < w; j += 1024 / kDcTile)
        if (ti >= j && ti < w) L[pk(ti, j)] = a[(c0 + ti) + (c0 + j) * n];
    if (tid == 0) bad = 0;
    __syncthreads();

    for (int j0 = 0; j0 < w; j0 += kDcSub)
    {
        const int jb = min(kDcSub, w - j0);
        if (wave == 0)
        {
            double row[kDcSub];
#pragma unroll
            for (int c = 0; c < kDcSub; ++c) row[c] = (lane < jb && c <= lane) ? L[pk(j0 + lane, j0 + c)] : 0.0;
#pragma unroll
            for (int c = 0; c < kDcSub; ++c)
            {
                if (c < jb)
                {
                    double d = lane_value(row[c], c);
                    if (!(d > 0.0) || !isfinite(d))
                    {
                        if (lane == 0) bad = 1;
                        d = 1.0;
                    }
                    d = sqrt(d);
                    if (lane == c) row[c] = d;
                    if (lane > c && lane < jb) row[c] /= d;
#pragma unroll
                    for (int m = c + 1; m < kDcSub; ++m)
                    {
                        const double dm = lane_value(row[c], m);
                        if (lane >= m && lane < jb) row[m] -= row[c] * dm;
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < kDcSub; ++c)
                if (lane < jb && c <= lane) L[pk(j0 + lane, j0 + c)] = row[c];
        }
        __syncthreads();
        const int rest = j0 + jb;
        if (rest + tid < w)
        {
            const int i = rest + tid;
            double x[kDcSub];
#pragma unroll
            for (int c = 0; c < kDcSub; ++c) x[c] = c < jb ? L[pk(i, j0 + c)] : 0.0;
#pragma unroll
            for (int c = 0; c < kDcSub; ++c)
            {
                if (c < jb)
                {
                    double v = x[c];
#pragma unroll
                    for (int t = 0; t < c; ++t) v -= x[t] * L[pk(j0 + c, j0 + t)];
                    x[c] = v / L[pk(j0 + c, j0 + c)];
                }
            }
#pragma unroll
            for (int c = 0; c < kDcSub; ++c)
                if (c < jb) L[pk(i, j0 + c)] = x[c];
        }
        __syncthreads();
        const int i = rest + (tid >> 3);
        if (i < w)
        {
            double li[kDcSub];
#pragma unroll
            for (int c = 0; c < kDcSub; ++c) li[c] = c < jb ? L[pk(i, j0 + c)] : 0.0;
            for (int t = rest + (tid & 7); t <= i; t += 8)
            {
                double s = 0.0;
#pragma unroll
                for (int c = 0; c < kDcSub; ++c)
                    if (c < jb) s += li[c] * L[pk(t, j0 + c)];
                L[pk(i, t)] -= s;
            }
        }
        __syncthreads();
    }

    // inverse: diagonal 16 x 16 blocks, one wavefront each, a column per lane
    const int nblk = (w + kDcSub - 1) / kDcSub;
    if (wave < nblk)
    {
        const int b0 = wave * kDcSub;
        const int bs = min(kDcSub, w - b0);
        const int c = lane;
        double x[kDcSub];
#pragma unroll
        for (int r = 0; r < kDcSub; ++r)
        {
            double v = 0.0;
            if (c < bs && r < bs && r >= c)
            {
                v = r == c ? 1.0 : 0.0;
#pragma unroll
                for (int t = 0; t < r; ++t)
                    if (t >= c) v -= L[pk(b0 + r, b0 + t)] * x[t];
                v /= L[pk(b0 + r, b0 + r)];
            }
            x[r] = v;
        }
#pragma unroll
        for (int r = 0; r < kDcSub; ++r)
            if (c < bs && r < bs && r >= c) X[pk(b0 + r, b0 + c)] = x[r];
    }
    __syncthreads();
    for (int d = 1; d < nblk; ++d)
    {
        const int count = (nblk - d) * kDcSub * kDcSub;
        for (int e = tid; e < count; e += 1024)
        {
            const int j = e / (kDcSub * kDcSub), r = (e / kDcSub) % kDcSub, c = e % kDcSub;
            const int gi = (j + d) * kDcSub + r, gc = j * kDcSub + c;
            double s = 0.0;
            if (gi < w)
                for (int k = j; k < j + d; ++k)
                    for (int t = 0; t < kDcSub; ++t)
                    {
                        const int gt = k * kDcSub + t;
                        if (gt >= gc) s += L[pk(gi, gt)] * X[pk(gt, gc)];
                    }
            U[e] = s;
        }
        __syncthreads();
        for (int e = tid; e < count; e += 1024)
        {
            const int j = e / (kDcSub * kDcSub), r = (e / kDcSub) % kDcSub, c = e % kDcSub;
            const int b0 = (j + d) * kDcSub, gi = b0 + r, gc = j * kDcSub + c;
            if (gi < w)
            {
                double s = 0.0;
                for (int t = 0; t <= r; ++t)
                    if (b0 + t < w) s += X[pk(gi, b0 + t)] * U[j * kDcSub * kDcSub + t * kDcSub + c];
                X[pk(gi, gc)] = -s;
            }
        }
        __syncthreads();
    }

    for (int j = tj; j < kDcTile; j += 1024 / kDcTile)
    {
        if (j < w && ti >= j && ti < w) a[(c0 + ti) + (c0 + j) * n] = L[pk(ti, j)];
        inv[ti + j * kDcTile] = (ti < w && j <= ti) ? X[pk(ti, j)] : 0.0;
    }
    if (tid == 0 && bad) *fail = 1;
}

// C[r][c] (-)= sum_k P[r][k] Q[c][k] over rows x cols, element (r, k) of P at
// P[r + k * ldp] and so on (column-major views).  One 128 x 128 tile per
// workgroup of four wavefronts, each wavefront a 64 x 64 quadrant as 4 x 4
// MFMA tiles.  v_mfma_f64_16x16x4f64 takes A[i][k] and B[k][j] from lane
// i + 16 k (resp. j + 16 k) and returns D[i][j] in lane j + 16 (i % 4),
// register i / 4.  Its A operand walks C's columns and its B operand C's
// rows, so that a lane's four accumulators are four columns of one row and
// 16 consecutive lanes hold 16 consecutive rows: the epilogue's reads and
// writes of column-major C are 128-byte runs.
//   tri:        tiles enumerated over the lower triangle (rows == cols);
//   mask_upper: skip elements with r < c (rows and cols share an origin).
// P may alias C (the in-place panel solve): each workgroup reads all K columns
// of its own rows before it writes them, and workgroups own disjoint rows.
template <bool SUB>
__global__ __launch_bounds__(256, 2) void k_dc_gemm(int64_t rows, int64_t cols, int K, const double* P, int64_t ldp,
                                                    const double* __restrict__ Q, int64_t ldq, double* C, int64_t ldc,
                                                    int tri, int mask_upper, const int* fail)
{
    if (*fail) return;
    int64_t ti, tj;
    const int64_t tr = (rows + kDcTile - 1) / kDcTile;
    // each XCD takes a contiguous run of tiles (mostly whole tile rows), so a
    // row's P slice stays in that XCD's L2
    const int64_t b = xcd_logical_block(blockIdx.x, gridDim.x);
    if (tri)
    {
        int64_t t = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(b) + 1.0) - 1.0) * 0.5);
        while (t * (t + 1) / 2 > b) --t;
        while ((t + 1) * (t + 2) / 2 <= b) ++t;
        ti = t;
        tj = b - t * (t + 1) / 2;
    }
    else
    {
        ti = b % tr;
        tj = b / tr;
    }
    const int64_t r0 = ti * kDcTile, c0 = tj * kDcTile;

    // two K chunks in flight: the MFMAs read one LDS buffer while the next
    // chunk is stored into the other, one barrier per chunk
    __shared__ double Ps[2][kDcKc * kDcLd];
    __shared__ double Qs[2][kDcKc * kDcLd];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
    const int li = lane & 15, lk = lane >> 4;

    dc_d4 acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = dc_d4{0.0, 0.0, 0.0, 0.0};

    constexpr int kPer = kDcKc * kDcTile / 256;
    double pr[kPer], qr[kPer];
    auto fetch = [&](int kc) {
#pragma unroll
        for (int e = 0; e < kPer; ++e)
        {
            const int idx = tid + e * 256;
            const int i = idx & (kDcTile - 1), k = idx / kDcTile;
            const int kg = kc + k;
            pr[e] = (r0 + i < rows && kg < K) ? P[(r0 + i) + static_cast<int64_t>(kg) * ldp] : 0.0;
            qr[e] = (c0 + i < cols && kg < K) ? Q[(c0 + i) + static_cast<int64_t>(kg) * ldq] : 0.0;
        }
    };
    auto stage = [&](int buf) {
#pragma unroll
        for (int e = 0; e < kPer; ++e)
        {
            const int idx = tid + e * 256;
            const int i = idx & (kDcTile - 1), k = idx / kDcTile;
            Ps[buf][k * kDcLd + i] = pr[e];
            Qs[buf][k * kDcLd + i] = qr[e];
        }
    };
    fetch(0);
    stage(0);
    if (kDcKc < K) fetch(kDcKc);
    __syncthreads();
    int buf = 0;
    for (int kc = 0; kc < K; kc += kDcKc)
    {
        const double* ps = Ps[buf];
        const double* qs = Qs[buf];
#pragma unroll
        for (int k4 = 0; k4 < kDcKc; k4 += 4)
        {
            const int k = k4 + lk;
            double av[4], bv[4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                av[t] = qs[k * kDcLd + wc + 16 * t + li];
                bv[t] = ps[k * kDcLd + wr + 16 * t + li];
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc[x][y], 0, 0, 0);
        }
        if (kc + kDcKc < K)
        {
            stage(buf ^ 1);
            if (kc + 2 * kDcKc < K) fetch(kc + 2 * kDcKc);
        }
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int64_t c = c0 + wc + 16 * x + lk + 4 * v;
                const int64_t r = r0 + wr + 16 * y + li;
                if (r < rows && c < cols && (!mask_upper || r >= c))
                {
                    double* dst = C + r + c * ldc;
                    *dst = SUB ? *dst - acc[x][y][v] : acc[x][y][v];
                }
            }
}

// Forward block step: y_k = inv_k z_k (z_k final), y_k -> yout, and
// z[r] -= L[r][k-block] y_k for the rows below the block.  inv_k is a full
// 128 x 128 column-major array with zeros above the diagonal and outside the
// block, so every dot runs over all 128 slots with independent loads.
//   y_k: row i by two threads (64 columns each), halves added in fixed order;
//   update: 64 rows per workgroup, lane = row, wave v = columns 32v..32v+31,
//   the four partial sums added in fixed order.
constexpr int kDcSolveRows = 64;
static_assert(kBlock == 4 * kWave && kDcSolveRows == kWave && kDcTile == 2 * kWave,
              "the block solves assume 4 wavefronts of 64 lanes and 128-wide blocks");

__global__ __launch_bounds__(kBlock) void k_dc_forward(int64_t n, int64_t c0, int w, const double* __restrict__ a,
                                                      const double* __restrict__ inv, double* z, double* yout,
                                                      const int* fail)
{
    __shared__ double zb[kDcTile];
    __shared__ double yk[kDcTile];
    __shared__ double halves[2][kDcTile];
    __shared__ double part[kBlock / kWave][kDcSolveRows];
    if (*fail) return;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    if (tid < kDcTile) zb[tid] = tid < w ? z[c0 + tid] : 0.0;
    __syncthreads();
    {
        const int i = tid & (kDcTile - 1), h = tid / kDcTile;
        double s = 0.0;
#pragma unroll 16
        for (int jj = 0; jj < kDcTile / 2; ++jj)
        {
            const int j = h * (kDcTile / 2) + jj;
            s += inv[i + j * kDcTile] * zb[j];
        }
        halves[h][i] = s;
    }
    __syncthreads();
    if (tid < kDcTile)
    {
        const double y = halves[0][tid] + halves[1][tid];
        yk[tid] = tid < w ? y : 0.0;
        if (blockIdx.x == 0 && tid < w) yout[c0 + tid] = y;
    }
    __syncthreads();
    const int64_t r = c0 + w + static_cast<int64_t>(blockIdx.x) * kDcSolveRows + lane;
    double s = 0.0;
    if (r < n)
    {
        const double* col = a + r + (c0 + wave * (kDcTile / 4)) * n;
#pragma unroll
        for (int jj = 0; jj < kDcTile / 4; ++jj)
            if (wave * (kDcTile / 4) + jj < w) s += col[jj * n] * yk[wave * (kDcTile / 4) + jj];
    }
    part[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && r < n) z[r] -= ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
}

// Backward block step: x_k = inv_k^T y_k (y_k final), x_k -> x, and
// y[c] -= sum_i L[c0+i][c] x_k[i] for the columns before the block.
//   x_k: output c by one wavefront (lanes over the 128 slots of inv_k's
//   column c, coalesced), butterfly sum; 32 outputs per wavefront;
//   update: 16 lanes per column over the block's rows, butterfly sum.
__global__ __launch_bounds__(kBlock) void k_dc_backward(int64_t n, int64_t c0, int w, const double* __restrict__ a,
                                                       const double* __restrict__ inv, double* y, double* x,
                                                       const int* fail)
{
    __shared__ double yb[kDcTile];
    __shared__ double xk[kDcTile];
    if (*fail) return;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    if (tid < kDcTile) yb[tid] = tid < w ? y[c0 + tid] : 0.0;
    __syncthreads();
    const double y0 = yb[lane], y1 = yb[lane + kWave];
#pragma unroll 8
    for (int cc = 0; cc < kDcTile / 4; ++cc)
    {
        const int c = wave * (kDcTile / 4) + cc;
        double s = inv[lane + c * kDcTile] * y0 + inv[lane + kWave + c * kDcTile] * y1;
#pragma unroll
        for (int m = 1; m < kWave; m <<= 1) s += __shfl_xor(s, m);
        if (lane == 0)
        {
            xk[c] = c < w ? s : 0.0;
            if (blockIdx.x == 0 && c < w) x[c0 + c] = s;
        }
    }
    __syncthreads();
    const int64_t col = static_cast<int64_t>(blockIdx.x) * (kBlock / 16) + (tid >> 4);
    const int q = tid & 15;
    double s = 0.0;
    if (col < c0)
        for (int i = q; i < w; i += 16) s += a[(c0 + i) + col * n] * xk[i];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    s += __shfl_xor(s, 4);
    s += __shfl_xor(s, 8);
    if (col < c0 && q == 0) y[col] -= s;
}

struct DenseLayout
{
    double* a;
    double* inv;
    double* y;
    int* fail;
    double* norms;  // [||b - A x||^2, ||b||^2]
    unsigned* counter;
};

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

template <bool SUB>
int launch_gemm(int64_t rows, int64_t cols, int K, const double* p, int64_t ldp, const double* q, int64_t ldq,
                double* c, int64_t ldc, bool tri, bool mask_upper, const int* fail, hipStream_t stream)
{
    if (rows <= 0 || cols <= 0) return PM_OK;
    const int64_t tr = ceil_div(rows, kDcTile), tc = ceil_div(cols, kDcTile);
    const int64_t blocks = tri ? tr * (tr + 1) / 2 : tr * tc;
    hipLaunchKernelGGL(k_dc_gemm<SUB>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, rows, cols, K, p,
                       ldp, q, ldq, c, ldc, tri ? 1 : 0, mask_upper ? 1 : 0, fail);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

// The factorization of a 128-column block at c0 of width w: factor and invert
// the diagonal block, then solve the rows below.
int factor_block(int64_t n, int64_t c0, int w, const DenseLayout& d, hipStream_t stream)
{
    double* inv = d.inv + (c0 / kDcTile) * kDcTile * kDcTile;
    hipLaunchKernelGGL(k_dc_potrf, dim3(1), dim3(1024), 0, stream, n, c0, w, d.a, inv, d.fail);
    PM_HIP(hipGetLastError());
    const int64_t below = n - c0 - w;
    double* panel = d.a + (c0 + w) + c0 * n;
    return launch_gemm<false>(below, w, w, panel, n, inv, kDcTile, panel, n, false, false, d.fail, stream);
}

} // namespace

int dense_cholesky_solve(const pm_csr* a, const double* b, double* x, pm_workspace* ws, DenseCholeskyOutcome* out,
                         hipStream_t stream)
{
    const int64_t n = a->rows;
    *out = DenseCholeskyOutcome{};
    if (n == 0) return PM_OK;
    if (n > kDenseCholeskyMaxRows)
        return fail(PM_INVALID_ARGUMENT, "dense Cholesky: %lld rows exceed the dense limit of %lld",
                    static_cast<long long>(n), static_cast<long long>(kDenseCholeskyMaxRows));
    const int64_t blocks = ceil_div(n, kDcTile);
    const size_t bytes_a = align_up(sizeof(double) * static_cast<size_t>(n) * static_cast<size_t>(n));
    const size_t bytes_inv = align_up(sizeof(double) * static_cast<size_t>(blocks) * kDcTile * kDcTile);
    const size_t bytes_y = align_up(sizeof(double) * static_cast<size_t>(n));
    PM_TRY(ws->dense.reserve(bytes_a + bytes_inv + bytes_y + 1024));
    PM_TRY(ws->partials.reserve(sizeof(double) * kMaxPartials * 8));
    char* base = ws->dense.as<char>();
    DenseLayout d;
    d.a = reinterpret_cast<double*>(base);
    d.inv = reinterpret_cast<double*>(base + bytes_a);
    d.y = reinterpret_cast<double*>(base + bytes_a + bytes_inv);
    char* tail = base + bytes_a + bytes_inv + bytes_y;
    d.fail = reinterpret_cast<int*>(tail);
    d.counter = reinterpret_cast<unsigned*>(tail + 256);
    d.norms = reinterpret_cast<double*>(tail + 512);

    hipLaunchKernelGGL(k_dc_flag_reset, dim3(1), dim3(1), 0, stream, d.fail, d.counter);
    const int zgrid = static_cast<int>(std::min<int64_t>(ceil_div(n * n, kBlock), 16384));
    hipLaunchKernelGGL(k_dc_zero_lower, dim3(zgrid), dim3(kBlock), 0, stream, n, d.a);
    const int sgrid = static_cast<int>(std::min<int64_t>(ceil_div(n, kBlock), 4096));
    if (a->row_offsets32 && a->col_indices32)
        hipLaunchKernelGGL(k_dc_scatter<int32_t>, dim3(sgrid), dim3(kBlock), 0, stream, n, a->nnz, a->row_offsets32,
                           a->col_indices32, static_cast<const double*>(a->values), d.a);
    else
        hipLaunchKernelGGL(k_dc_scatter<int64_t>, dim3(sgrid), dim3(kBlock), 0, stream, n, a->nnz, a->row_offsets,
                           a->col_indices, static_cast<const double*>(a->values), d.a);
    PM_HIP(hipGetLastError());

    // Lookahead: the trailing update of panel p is split into the next
    // panel's columns (on `stream`, right away) and the rest (on the
    // workspace's second stream), so that factoring panel p+1 overlaps the
    // bulk update.  Panel q is written by the bulk update of panel q-2 at the
    // latest, which `stream` waits for before the update that precedes it.
    PM_TRY(ws->ensure_aux());
    hipStream_t aux = ws->aux_stream;
    hipEvent_t panel_done = ws->aux_events[0], bulk_done = ws->aux_events[1];
    bool bulk_pending = false;
    for (int64_t p0 = 0; p0 < n; p0 += kDcPanel)
    {
        const int w0 = static_cast<int>(std::min<int64_t>(kDcTile, n - p0));
        PM_TRY(factor_block(n, p0, w0, d, stream));
        const int64_t c1 = p0 + kDcTile;
        if (w0 == kDcTile && c1 < n)
        {
            const int w1 = static_cast<int>(std::min<int64_t>(kDcTile, n - c1));
            // second block of the panel: A[c1:, c1:c1+w1] -= L[c1:, p0:c1] L[c1:c1+w1, p0:c1]^T
            const double* left = d.a + c1 + p0 * n;
            PM_TRY(launch_gemm<true>(n - c1, w1, kDcTile, left, n, left, n, d.a + c1 + c1 * n, n, false, true, d.fail,
                                     stream));
            PM_TRY(factor_block(n, c1, w1, d, stream));
        }
        const int K = static_cast<int>(std::min<int64_t>(kDcPanel, n - p0));
        const int64_t t0 = p0 + K;
        if (t0 >= n) break;
        const double* panel = d.a + t0 + p0 * n;
        if (bulk_pending)
        {
            PM_HIP(hipStreamWaitEvent(stream, bulk_done, 0));
            bulk_pending = false;
        }
        // the next panel's columns
        const int kn = static_cast<int>(std::min<int64_t>(kDcPanel, n - t0));
        PM_TRY(launch_gemm<true>(n - t0, kn, K, panel, n, panel, n, d.a + t0 + t0 * n, n, false, true, d.fail,
                                 stream));
        // everything right of them
        const int64_t r0 = t0 + kn;
        if (r0 < n)
        {
            const double* rest = d.a + r0 + p0 * n;
            PM_HIP(hipEventRecord(panel_done, stream));
            PM_HIP(hipStreamWaitEvent(aux, panel_done, 0));
            PM_TRY(launch_gemm<true>(n - r0, n - r0, K, rest, n, rest, n, d.a + r0 + r0 * n, n, true, true, d.fail,
                                     aux));
            PM_HIP(hipEventRecord(bulk_done, aux));
            bulk_pending = true;
        }
    }
    if (bulk_pending) PM_HIP(hipStreamWaitEvent(stream, bulk_done, 0));

    // x <- b, forward into y, backward into x
    PM_HIP(hipMemcpyAsync(x, b, sizeof(double) * static_cast<size_t>(n), hipMemcpyDeviceToDevice, stream));
    for (int64_t k = 0; k < blocks; ++k)
    {
        const int64_t c0 = k * kDcTile;
        const int w = static_cast<int>(std::min<int64_t>(kDcTile, n - c0));
        const int64_t grid = std::max<int64_t>(1, ceil_div(n - c0 - w, kDcSolveRows));
        hipLaunchKernelGGL(k_dc_forward, dim3(static_cast<unsigned>(grid)), dim3(kBlock), 0, stream, n, c0, w, d.a,
                           d.inv + k * kDcTile * kDcTile, x, d.y, d.fail);
    }
    for (int64_t k = blocks - 1; k >= 0; --k)
    {
        const int64_t c0 = k * kDcTile;
        const int w = static_cast<int>(std::min<int64_t>(kDcTile, n - c0));
        const int64_t grid = std::max<int64_t>(1, ceil_div(c0, kBlock / 16));
        hipLaunchKernelGGL(k_dc_backward, dim3(static_cast<unsigned>(grid)), dim3(kBlock), 0, stream, n, c0, w, d.a,
                           d.inv + k * kDcTile * kDcTile, d.y, x, d.fail);
    }
    PM_HIP(hipGetLastError());

    PM_TRY(relres_launch(a, b, x, PM_F64, ws->partials.as<double>(), d.counter, d.norms, stream));
    struct
    {
        double norms[2];
        int fail;
    } host{};
    PM_HIP(hipMemcpyAsync(host.norms, d.norms, sizeof(host.norms), hipMemcpyDeviceToHost, stream));
    PM_HIP(hipMemcpyAsync(&host.fail, d.fail, sizeof(int), hipMemcpyDeviceToHost, stream));
    PM_HIP(hipStreamSynchronize(stream));
    out->not_positive_definite = host.fail;
    out->residual_norm = std::sqrt(host.norms[0]);
    out->rhs_norm = std::sqrt(host.norms[1]);
    return PM_OK;
}

} // namespace pm

using namespace pm;

extern "C" {

int pm_dense_cholesky_solve(const pm_csr* a, const double* b, double* x, pm_dense_report* report, pm_workspace* ws,
                            hipStream_t stream)
{
    if (!a || !report || !ws) return fail(PM_INVALID_ARGUMENT, "dense Cholesky: null argument");
    std::memset(report, 0, sizeof(*report));
    PM_TRY(check_csr(a, "dense Cholesky"));
    if (a->dtype != PM_F64) return fail(PM_INVALID_ARGUMENT, "dense Cholesky: the matrix must be FP64");
    if (a->rows != a->cols) return fail(PM_INVALID_ARGUMENT, "dense Cholesky: the matrix must be square");
    if (a->rows > 0 && (!b || !x)) return fail(PM_INVALID_ARGUMENT, "dense Cholesky: null vector");
    if (b && b == x) return fail(PM_INVALID_ARGUMENT, "dense Cholesky: rhs and solution must not alias");
    if (a->nnz > 0 && !a->row_offsets32 && !a->row_offsets)
        return fail(PM_INVALID_ARGUMENT, "dense Cholesky: missing row offsets");
    PM_TRY(ws->bind(stream));
    DenseCholeskyOutcome out;
    PM_TRY(dense_cholesky_solve(a, b, x, ws, &out, stream));
    report->not_positive_definite = out.not_positive_definite;
    report->rhs_norm = out.rhs_norm;
    report->residual_norm = out.residual_norm;
    return PM_OK;
}

} // extern "C"
