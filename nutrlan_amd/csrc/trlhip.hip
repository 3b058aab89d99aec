// trlhip.hip -- CDNA4 (gfx950) device layer of the thick-restart Lanczos
// solver: device CSR operator, Lanczos basis in HBM, and the hot-path
// kernels (SpMV + alpha, recurrence, classical Gram-Schmidt projection and
// update, FP64-MFMA restart GEMM).  C ABI in include/trlhip.h.
//
// Storage: the basis is FP64 (real), FP64 complex (interleaved re/im) or
// FP32 (mixed precision); every reduction and every Gram-Schmidt
// coefficient is accumulated in FP64.
//
// Reductions (alpha, ||r||^2, V^H r) never leave the device until the final
// scalar: each block reduces its rows (wave shuffles, then LDS) into a
// partial slot, and a finalize kernel (or the consuming kernel) sums the
// partials in block order, so results are deterministic run to run.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>
#include <thread>
#include <type_traits>
#include <unordered_map>

#include "trlan_internal.h"
#include "trlhip.h"

namespace {

constexpr int kBlock = 256;            // 4 waves of 64
constexpr int kMaxGrid = 1024;         // partial slots per streaming reduction
constexpr int kSpmvGrid = 2048;        // SpMV blocks (and alpha partial slots)
constexpr int kProjGrid = 2048;        // row blocks (partial rows) of the projection
// projection blocks wanted over all column chunks, by the bytes of a basis
// element read (FP32 shadow, FP64, c128); measured on the 128^3 Laplacian
// (n = 2^21, m up to 64): FP32 reads 8192 blocks 25.30, 4096 25.15, 2048
// 24.99 eigenpairs/s; FP64 reads 2048 17.56, 4096 16.96, 8192 17.16
constexpr int kProjBlocks32 = 8192, kProjBlocks64 = 2048, kProjBlocksC128 = 4096;
constexpr int kNnzBlk = 2048;          // nonzeros staged per SpMV row block
constexpr int kRowsBlk = kBlock;       // rows per SpMV row block (at most)
// Smallest share of work per block: rows per update or projection block
// (4 rows per thread), row blocks per CSR-stream SpMV block.  Measured on
// the batched 2D Anderson workload: 2x, 4x and 8x these (fewer, longer
// blocks) run 3 %, 4 % and 10 % slower.
constexpr int kMinRowsUpd = 4 * kBlock;
constexpr int kMinSpmvRowBlocks = 1;
constexpr int kChainMax = 32;  // Lanczos steps one batched context may record per round
constexpr uint64_t kCsrMagic = 0x7472686970637372ull;  // "trhipcsr"


#define TRY(expr)                                  \
    do {                                           \
        hipError_t e_ = (expr);                    \
        if (e_ != hipSuccess) return TRLHIP_EHIP;  \
    } while (0)

// ------------------------------------------------------------ reductions

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Transpose-reduce N per-lane values across the 64 lanes of a wave:
// afterwards lane l holds the wave total of value index l / (64 / N).
// Costs N-1+log2(64/N) shuffles instead of N*6.
template <int N>
__device__ __forceinline__ double wave_transpose_sum(double (&v)[N], int lane)
{
    static_assert(N >= 1 && N <= 64 && (N & (N - 1)) == 0, "N power of 2");
    int o = 32;
#pragma unroll
    for (int cnt = N; cnt > 1; cnt >>= 1, o >>= 1) {
        const bool hi = (lane & o) != 0;
#pragma unroll
        for (int i = 0; i < cnt / 2; ++i) {
            // select values, not array elements: a conditional lvalue
            // would become a dynamically indexed (scratch) array access
            const double lo_v = v[i], hi_v = v[i + cnt / 2];
            const double send = hi ? lo_v : hi_v;
            const double keep = hi ? hi_v : lo_v;
            v[i] = keep + __shfl_xor(send, o);
        }
    }
#pragma unroll
    for (; o > 0; o >>= 1) v[0] += __shfl_xor(v[0], o);
    return v[0];
}

// Sum of one double over the block; valid in thread 0.
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double sm[kBlock / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sm[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < kBlock / 64; ++i) t += sm[i];
    }
    __syncthreads();
    return t;
}

// Cross-block reductions never synchronize blocks inside a kernel (no
// device-scope fences or tickets, which are costly on this memory model):
// every block writes its partial, and a later kernel -- the finalize
// kernels below, or the consumer itself -- sums the partials in block
// order after the kernel boundary has made them visible.

// Sum of part[b * stride + p] over b < nb, four independent accumulators
// so the loads overlap; the order depends only on nb.
__device__ __forceinline__ double strided_sum(const double *part, int nb, size_t stride, size_t p,
                                              int b0, int step)
{
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int b = b0;
    for (; b + 3 * step < nb; b += 4 * step) {
        s0 += part[(size_t)b * stride + p];
        s1 += part[(size_t)(b + step) * stride + p];
        s2 += part[(size_t)(b + 2 * step) * stride + p];
        s3 += part[(size_t)(b + 3 * step) * stride + p];
    }
    for (; b < nb; b += step) s0 += part[(size_t)b * stride + p];
    return (s0 + s1) + (s2 + s3);
}

// Ordered sum of nb scalar partials by one block; result in thread 0.
__device__ __forceinline__ double ordered_sum(const double *part, int nb)
{
    return block_sum(strided_sum(part, nb, 1, 0, threadIdx.x, kBlock));
}

// Finalize a scalar reduction: out[0] = sum of nb partials (one block).
__global__ __launch_bounds__(kBlock) void k_sum(const double *__restrict__ part, int nb,
                                                double *__restrict__ out)
{
    const double t = ordered_sum(part, nb);
    if (threadIdx.x == 0) out[0] = t;
}

// Two scalar reductions at once: block 0 sums p0 into o0, block 1 p1 into
// o1.  With a host mirror (pinned, host-coherent memory), the results and
// the scalars `dscal` already holds are also written there, so the host
// reads them after a stream synchronize without a copy.
__device__ __forceinline__ void sum2_body(const double *__restrict__ p0, int n0,
                                                 double *__restrict__ o0,
                                                 const double *__restrict__ p1, int n1,
                                                 double *__restrict__ o1,
                                                 const double *__restrict__ dscal,
                                                 double *__restrict__ hmirror, int bid, int nblk)
{
    const bool second = bid == 1;
    const double t = ordered_sum(second ? p1 : p0, second ? n1 : n0);
    if (threadIdx.x == 0) {
        *(second ? o1 : o0) = t;
        if (hmirror) {
            if (second) {
                hmirror[o1 - dscal] = t;
            } else {
                hmirror[o0 - dscal] = t;
                for (int i = 0; i < 4; ++i)
                    if (dscal + i != o0 && dscal + i != o1) hmirror[i] = dscal[i];
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_sum2(const double *__restrict__ p0, int n0,
                                                 double *__restrict__ o0,
                                                 const double *__restrict__ p1, int n1,
                                                 double *__restrict__ o1,
                                                 const double *__restrict__ dscal,
                                                 double *__restrict__ hmirror)
{
    sum2_body(p0, n0, o0, p1, n1, o1, dscal, hmirror, blockIdx.x, gridDim.x);
}

// Finalize a vector reduction: hout[p] = sum_b part[p * nb + b], p < np
// (each coefficient's nb partials are contiguous).  One block per p.
__device__ __forceinline__ void colsum_body(const double *__restrict__ part, int nb,
                                                   double *__restrict__ hout, int bid, int nblk)
{
    const double t = ordered_sum(part + (size_t)bid * nb, nb);
    if (threadIdx.x == 0) hout[bid] = t;
}

__global__ __launch_bounds__(kBlock) void k_colsum(const double *__restrict__ part, int nb,
                                                   double *__restrict__ hout)
{
    colsum_body(part, nb, hout, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------- SpMV

// a + b * c rounded twice (no FMA contraction), so that kernels that form
// the same sums in different code agree bit for bit
__device__ __forceinline__ double add_prod(double a, double b, double c)
{
#pragma clang fp contract(off)
    const double p = b * c;
    return a + p;
}

// y = A x, CSR-stream: a block walks row blocks [rb[b], rb[b+1]) (grid
// stride); it stages each row block's nonzeros (at most kNnzBlk) as
// products val*x[col] in LDS with fully coalesced loads, then each thread
// sums one row from LDS.  A row block holding one very long row is reduced
// by the whole block.  Epilogue: the block's partial of Re(x^H y) in
// part[blockIdx.x] (summed by the consumer).
// S = vector storage, SV = matrix value storage, E = doubles per element.
// FUSE: x is fl(x_src * xscale) formed on the fly from the unscaled
// source (see spmv_vec_body); the row owner also stores it to xq (and its
// FP32 rounding to xq32).
// PK (real only): the nonzeros come packed instead of ci/val, one 32-bit
// word each: bits 31..24 a value index (255: the diagonal value dg[col],
// else tab[index]), bits 23..0 the column minus the row block's first row
// (signed).  The decoded values are the CSR's values bit for bit, so the
// products and sums are the unpacked kernel's.
template <typename S, typename SV, int E, bool FUSE, bool PK = false>
__device__ __forceinline__ void spmv_stream_body(
    int nb, const int *__restrict__ rb, const int *__restrict__ rp, const int *__restrict__ ci,
    const SV *__restrict__ val, const S *__restrict__ x, S *__restrict__ y,
    double *__restrict__ part, double xscale, S *__restrict__ xq, float *__restrict__ xq32, int bid,
    int nblk, const uint32_t *__restrict__ pk = nullptr, const double *__restrict__ dg = nullptr,
    const double *__restrict__ tab = nullptr)
{
    static_assert(!PK || E == 1, "packed nonzeros are real");
    __shared__ double prod[kNnzBlk * E];
    auto xv = [&](int c, int e) -> double {
        if constexpr (FUSE)
            return (double)(S)((double)x[(size_t)E * c + e] * xscale);
        else
            return (double)x[(size_t)E * c + e];
    };
    auto own = [&](int row) {  // the row's x value, stored when fused
        if constexpr (FUSE) {
            for (int e = 0; e < E; ++e) {
                const double v = xv(row, e);
                if (xq) xq[(size_t)E * row + e] = (S)v;
                if (E == 1 && xq32) xq32[row] = (float)v;
            }
        }
    };
    double acc = 0.0;
    for (int blk = bid; blk < nb; blk += nblk) {
        const int r0 = rb[blk], r1 = rb[blk + 1];
        const int k0 = rp[r0], nz = rp[r1] - k0;
        // nonzero k of the block: its column and (real) value
        auto nzr = [&](int k, int &c) -> double {
            if constexpr (PK) {
                const uint32_t w = pk[k0 + k];
                c = r0 + ((int)(w << 8) >> 8);
                const uint32_t vi = w >> 24;
                return vi == 255u ? dg[c] : tab[vi];
            } else {
                c = ci[k0 + k];
                return (double)val[k0 + k];
            }
        };
        if (r1 - r0 > 1 || nz <= kNnzBlk) {
#pragma unroll 4
            for (int k = threadIdx.x; k < nz; k += kBlock) {
                if constexpr (E == 1) {
                    int c;
                    const double v = nzr(k, c);
                    prod[k] = v * xv(c, 0);
                } else {
                    const int c = ci[k0 + k];
                    const double ar = val[2 * (k0 + k)], ai = val[2 * (k0 + k) + 1];
                    const double xr = xv(c, 0), xi = xv(c, 1);
                    prod[2 * k] = ar * xr - ai * xi;
                    prod[2 * k + 1] = ar * xi + ai * xr;
                }
            }
            __syncthreads();
            const int t = threadIdx.x;
            if (t < r1 - r0) {
                const int row = r0 + t;
                const int a = rp[row] - k0, b = rp[row + 1] - k0;
                own(row);
                if constexpr (E == 1) {
                    double s = 0.0;
                    for (int k = a; k < b; ++k) s += prod[k];
                    y[row] = (S)s;
                    acc = add_prod(acc, xv(row, 0), (double)(S)s);
                } else {
                    double sr = 0.0, si = 0.0;
                    for (int k = a; k < b; ++k) {
                        sr += prod[2 * k];
                        si += prod[2 * k + 1];
                    }
                    y[2 * row] = sr;
                    y[2 * row + 1] = si;
                    acc += xv(row, 0) * sr + xv(row, 1) * si;
                }
            }
            __syncthreads();
        } else {
            double sr = 0.0, si = 0.0;
            for (int k = threadIdx.x; k < nz; k += kBlock) {
                if constexpr (E == 1) {
                    int c;
                    const double v = nzr(k, c);
                    sr += v * xv(c, 0);
                } else {
                    const int c = ci[k0 + k];
                    const double ar = val[2 * (k0 + k)], ai = val[2 * (k0 + k) + 1];
                    sr += ar * xv(c, 0) - ai * xv(c, 1);
                    si += ar * xv(c, 1) + ai * xv(c, 0);
                }
            }
            sr = block_sum(sr);
            if (E == 2) si = block_sum(si);
            if (threadIdx.x == 0) {
                own(r0);
                if constexpr (E == 1) {
                    y[r0] = (S)sr;
                    acc = add_prod(acc, xv(r0, 0), (double)(S)sr);
                } else {
                    y[2 * r0] = sr;
                    y[2 * r0 + 1] = si;
                    acc += xv(r0, 0) * sr + xv(r0, 1) * si;
                }
            }
        }
    }
    const double bs = block_sum(acc);
    if (threadIdx.x == 0) part[bid] = bs;
}

template <typename S, typename SV, int E, bool FUSE>
__global__ __launch_bounds__(kBlock) void k_spmv(
    int nb, const int *__restrict__ rb, const int *__restrict__ rp, const int *__restrict__ ci,
    const SV *__restrict__ val, const S *__restrict__ x, S *__restrict__ y,
    double *__restrict__ part, double xscale, S *__restrict__ xq, float *__restrict__ xq32)
{
    spmv_stream_body<S, SV, E, FUSE>(nb, rb, rp, ci, val, x, y, part, xscale, xq, xq32, blockIdx.x,
                                     gridDim.x);
}

// CSR-stream SpMV over packed nonzeros (real).
template <bool FUSE>
__global__ __launch_bounds__(kBlock) void k_spmv_pk(
    int nb, const int *__restrict__ rb, const int *__restrict__ rp, const uint32_t *__restrict__ pk,
    const double *__restrict__ dg, const double *__restrict__ tab, const double *__restrict__ x,
    double *__restrict__ y, double *__restrict__ part, double xscale, double *__restrict__ xq,
    float *__restrict__ xq32)
{
    spmv_stream_body<double, double, 1, FUSE, true>(nb, rb, rp, nullptr, nullptr, x, y, part, xscale,
                                                    xq, xq32, blockIdx.x, gridDim.x, pk, dg, tab);
}

// y = A x for a periodic nearest-neighbour stencil on an L^D lattice (D = 2
// or 3; site x + L y (+ L^2 z)) whose off-diagonal entries all hold `off`
// and whose diagonal is dg: the 2D+1 columns of a row are computed from
// its lattice coordinates instead of read, sorted ascending as the CSR
// stores them, and the products (each rounded before it is added, as the
// CSR-stream kernel's LDS-staged products are) summed in that order.  Row
// blocks, the row-to-thread map and the alpha partials follow
// spmv_stream_body, so y and the partials equal the CSR-stream kernel's on
// the same matrix bit for bit, while the only matrix bytes read are the
// diagonal's.
// One row of the periodic-stencil product: its K = 2D+1 columns formed
// from the lattice position and sorted ascending (the CSR column order),
// the products rounded and summed left to right as add_prod does in the
// CSR-stream kernel, so y[row] has the CSR kernel's bits.  Returns y[row]
// and sets xr = the (scaled) x[row] for the alpha product; FUSE stores the
// scaled x to xq / xq32.
template <int D>
__device__ __forceinline__ void stencil_cols(int row, int L, int *c)
{
    constexpr int K = 2 * D + 1;
    {
        const int xs = row % L, rest = row / L;
        const int ys = D == 2 ? rest : rest % L;
        const int base = row - xs;  // the site with x = 0 on this row's line
        c[0] = row;
        c[1] = base + (xs == 0 ? L - 1 : xs - 1);
        c[2] = base + (xs == L - 1 ? 0 : xs + 1);
        const int ybase = row - ys * L;
        c[3] = ybase + (ys == 0 ? L - 1 : ys - 1) * L;
        c[4] = ybase + (ys == L - 1 ? 0 : ys + 1) * L;
        if constexpr (D == 3) {
            const int zs = rest / L;
            const int L2 = L * L;
            const int zbase = row - zs * L2;
            c[5] = zbase + (zs == 0 ? L - 1 : zs - 1) * L2;
            c[6] = zbase + (zs == L - 1 ? 0 : zs + 1) * L2;
        }
    }
    // odd-even transposition sort: fixed indices, so c stays in registers
#pragma unroll
    for (int r = 0; r < K; ++r)
#pragma unroll
        for (int i = r & 1; i + 1 < K; i += 2) {
            const int lo = min(c[i], c[i + 1]), hi = max(c[i], c[i + 1]);
            c[i] = lo;
            c[i + 1] = hi;
        }
}

template <int D, bool FUSE>
__device__ __forceinline__ double stencil_row(int row, int L, double off,
                                              const double *__restrict__ dg,
                                              const double *__restrict__ x, double xscale,
                                              double *__restrict__ xq, float *__restrict__ xq32,
                                              double &xr)
{
    constexpr int K = 2 * D + 1;
    auto xv = [&](int c) -> double {
        if constexpr (FUSE)
            return (double)(x[c] * xscale);
        else
            return x[c];
    };
    int c[K];
    stencil_cols<D>(row, L, c);
    xr = xv(row);
    if constexpr (FUSE) {
        if (xq) xq[row] = xr;
        if (xq32) xq32[row] = (float)xr;
    }
    const double d = dg[row];
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) s = add_prod(s, c[k] == row ? d : off, xv(c[k]));
    return s;
}

// y = A x for a periodic nearest-neighbour stencil (no matrix arrays: the
// columns come from the lattice, the off-diagonal is one value, the
// diagonal dg).  Same row blocks and block -> row-block map as the
// CSR-stream kernel (block bid takes row blocks bid, bid + nblk, ...), one
// row per thread, alpha accumulated over a thread's rows in that order, one
// partial per block: results, alpha partials included, have the CSR-stream
// kernel's bits.  A thread's next kStU row blocks are formed together so
// their loads are in flight at once; the accumulation order is unchanged.
constexpr int kStU = 4;

template <int D, bool FUSE>
__device__ __forceinline__ void spmv_stencil_body(
    int n, int nb, int L, double off, const double *__restrict__ dg,
    const double *__restrict__ x, double *__restrict__ y, double *__restrict__ part, double xscale,
    double *__restrict__ xq, float *__restrict__ xq32, int bid, int nblk)
{
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int blk0 = bid; blk0 < nb; blk0 += kStU * nblk) {
        double sv[kStU], xr[kStU];
        bool on[kStU];
#pragma unroll
        for (int u = 0; u < kStU; ++u) {
            const int blk = blk0 + u * nblk;
            on[u] = false;
            if (blk < nb) {
                const int r0 = blk * kRowsBlk, r1 = min(n, r0 + kRowsBlk);
                if (t < r1 - r0) {
                    on[u] = true;
                    sv[u] = stencil_row<D, FUSE>(r0 + t, L, off, dg, x, xscale, xq, xq32, xr[u]);
                    y[r0 + t] = sv[u];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kStU; ++u)
            if (on[u]) acc = add_prod(acc, xr[u], sv[u]);
    }
    const double bs = block_sum(acc);
    if (threadIdx.x == 0) part[bid] = bs;
}

constexpr int kStMaxL = 512;  // widest 2D lattice whose x window is staged in LDS

// The batched form: workgroup w of a problem takes row blocks w * RB ..
// w * RB + RB - 1 and writes each one's alpha partial to part[blk], as RB
// separate one-row-block blocks would (same per-row arithmetic, same
// wave_sum / block_sum order), so with nblk = nb partials the results equal
// spmv_stencil_body's bit for bit while a workgroup keeps RB rows' loads in
// flight per thread.
template <int D, int RB>
__device__ __forceinline__ void spmv_stencil_rb(int n, int nb, int L, double off,
                                                const double *__restrict__ dg,
                                                const double *__restrict__ x, double *__restrict__ y,
                                                double *__restrict__ part, double xscale,
                                                double *__restrict__ xq, float *__restrict__ xq32,
                                                int w)
{
    __shared__ double sm[RB][kBlock / 64];
    const int t = threadIdx.x;
    double acc[RB];
    bool staged = false;
    if constexpr (D == 2) {
        // 2D lattices of side L <= kStMaxL: the workgroup's rows
        // [R0, R0 + RB * kRowsBlk) read x only in the window
        // [R0 - L, R0 + RB * kRowsBlk + L) but for the periodic wrap in y,
        // so the window is staged (scaled) in LDS with coalesced loads and
        // the 2D+1 neighbours are read from there: one global x load per
        // row and a halo instead of five.  Same products in the same order.
        staged = L <= kStMaxL;
        if (staged) {
            __shared__ double xw[RB * kRowsBlk + 2 * kStMaxL];
            constexpr int kW = (RB * kRowsBlk + 2 * kStMaxL + kBlock - 1) / kBlock;
            const int R0 = w * RB * kRowsBlk;
            const int lo = max(0, R0 - L), hi = min(n, R0 + RB * kRowsBlk + L);
            double xt[kW], dv[RB];
#pragma unroll
            for (int k = 0; k < kW; ++k) {
                const int i = lo + k * kBlock + t;
                if (i < hi) xt[k] = (double)(x[i] * xscale);
            }
#pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int row = R0 + u * kRowsBlk + t;
                if (row < n) dv[u] = dg[row];
            }
#pragma unroll
            for (int k = 0; k < kW; ++k) {
                const int i = lo + k * kBlock + t;
                if (i < hi) xw[i - lo] = xt[k];
            }
            __syncthreads();
            auto xv = [&](int c) -> double {
                return (c >= lo && c < hi) ? xw[c - lo] : (double)(x[c] * xscale);
            };
#pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int row = R0 + u * kRowsBlk + t;
                acc[u] = 0.0;
                if (row < n) {
                    int c[2 * D + 1];
                    stencil_cols<D>(row, L, c);
                    const double xr = xw[row - lo];
                    if (xq) xq[row] = xr;
                    if (xq32) xq32[row] = (float)xr;
                    double s = 0.0;
#pragma unroll
                    for (int k = 0; k < 2 * D + 1; ++k)
                        s = add_prod(s, c[k] == row ? dv[u] : off, xv(c[k]));
                    y[row] = s;
                    acc[u] = add_prod(0.0, xr, s);
                }
            }
        }
    }
    if (!staged) {
#pragma unroll
        for (int u = 0; u < RB; ++u) {
            const int blk = w * RB + u;
            acc[u] = 0.0;
            if (blk < nb) {
                const int r0 = blk * kRowsBlk, r1 = min(n, r0 + kRowsBlk);
                if (t < r1 - r0) {
                    double xr;
                    const double s =
                        stencil_row<D, true>(r0 + t, L, off, dg, x, xscale, xq, xq32, xr);
                    y[r0 + t] = s;
                    acc[u] = add_prod(0.0, xr, s);
                }
            }
        }
    }
    const int lane = t & 63, wv = t >> 6;
#pragma unroll
    for (int u = 0; u < RB; ++u) {
        const double v = wave_sum(acc[u]);
        if (lane == 0) sm[u][wv] = v;
    }
    __syncthreads();
    if (t < RB && w * RB + t < nb) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < kBlock / 64; ++i) s += sm[t][i];
        part[w * RB + t] = s;
    }
}

template <int D, bool FUSE>
__global__ __launch_bounds__(kBlock) void k_spmv_st(
    int n, int nb, int L, double off, const double *__restrict__ dg,
    const double *__restrict__ x, double *__restrict__ y, double *__restrict__ part, double xscale,
    double *__restrict__ xq, float *__restrict__ xq32)
{
    spmv_stencil_body<D, FUSE>(n, nb, L, off, dg, x, y, part, xscale, xq, xq32, blockIdx.x,
                               gridDim.x);
}

// Operators of at most kSpmvGrid row blocks: workgroup w takes RB row
// blocks (alpha partials one per row block, as above with one row block per
// block).  xscale = 1 with no xq / xq32 is the plain product: x * 1 is x.
template <int D, int RB>
__global__ __launch_bounds__(kBlock) void k_spmv_st_rb(
    int n, int nb, int L, double off, const double *__restrict__ dg,
    const double *__restrict__ x, double *__restrict__ y, double *__restrict__ part, double xscale,
    double *__restrict__ xq, float *__restrict__ xq32)
{
    spmv_stencil_rb<D, RB>(n, nb, L, off, dg, x, y, part, xscale, xq, xq32, blockIdx.x);
}

// Column-panel SpMV, for large operators whose rows gather x from all
// over (random sparse patterns): CSR-stream gathers then move a whole
// cache line from the fabric per 8-byte x element.  Here the nonzeros are
// kept once more, split by column into panels of kPanelCols columns (x's
// share of a panel, 2 MiB, fits an XCD's 4 MiB L2), each panel's in row
// order and cut into chunks of kPanelChunk.  Stage 1 (k_spmv_panel): block
// w takes chunk perm[w]; every chunk of a panel goes to blocks of one XCD
// (workgroups go to the 8 XCDs round-robin by index), back to back, so the
// panel's x is read from HBM about once and then served by that L2.  It
// forms the products (rounded, as the CSR-stream kernel's are), and one
// partial per run of equal rows, summed in nonzero order by the thread
// owning the run's first nonzero, into sp (runs numbered in chunk order,
// from seg_base[chunk]).  Stage 2 (k_spmv_panel_rows): each row sums its
// partials in panel order (rs_ptr / rs_idx), stores y and the fused
// normalised x, and the alpha partials as the other kernels do.  The sums
// are grouped by panel, so results differ from the CSR-stream kernel's in
// the last bits; they are deterministic run to run.
constexpr int kPanelCols = 1 << 18;
constexpr int kPanelChunk = 2048;

template <bool FUSE>
__global__ __launch_bounds__(kBlock) void k_spmv_panel(
    const int *__restrict__ perm, const long *__restrict__ cstart, const int *__restrict__ seg_base,
    const int *__restrict__ erow, const int *__restrict__ ecol, const double *__restrict__ eval,
    const double *__restrict__ x, double xscale, double *__restrict__ sp)
{
    constexpr int PER = kPanelChunk / kBlock;
    __shared__ double prod[kPanelChunk];
    __shared__ int rows[kPanelChunk + 1];
    __shared__ int wsum[kBlock / 64];
    const int chunk = perm[blockIdx.x];
    if (chunk < 0) return;  // an idle slot of the XCD map: the whole block leaves
    const long k0 = cstart[chunk];
    const int nz = (int)(cstart[chunk + 1] - k0);
    for (int k = threadIdx.x; k < nz; k += kBlock) {
        const int c = ecol[k0 + k];
        const double xv = FUSE ? (double)(x[c] * xscale) : x[c];
        prod[k] = eval[k0 + k] * xv;
        rows[k] = erow[k0 + k];
    }
    if (threadIdx.x == 0) rows[nz] = -1;
    __syncthreads();
    const int a = threadIdx.x * PER;
    int starts = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int k = a + i;
        if (k < nz && (k == 0 || rows[k] != rows[k - 1])) ++starts;
    }
    // exclusive scan of the run starts over the block: wave scan, then LDS
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = starts;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int id = seg_base[chunk] + incl - starts;
    for (int i = 0; i < w; ++i) id += wsum[i];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int k = a + i;
        if (k < nz && (k == 0 || rows[k] != rows[k - 1])) {
            double sum = 0.0;
            int e = k;
            do {
                sum += prod[e];
                ++e;
            } while (rows[e] == rows[k]);
            sp[id++] = sum;
        }
    }
}

template <bool FUSE>
__global__ __launch_bounds__(kBlock) void k_spmv_panel_rows(
    int n, const int *__restrict__ rs_ptr, const int *__restrict__ rs_idx,
    const double *__restrict__ sp, const double *__restrict__ x, double *__restrict__ y,
    double *__restrict__ part, double xscale, double *__restrict__ xq, float *__restrict__ xq32)
{
    double acc = 0.0;
    for (int row = blockIdx.x * kBlock + threadIdx.x; row < n; row += gridDim.x * kBlock) {
        double sum = 0.0;
        for (int k = rs_ptr[row]; k < rs_ptr[row + 1]; ++k) sum += sp[rs_idx[k]];
        const double xv = FUSE ? (double)(x[row] * xscale) : x[row];
        if constexpr (FUSE) {
            if (xq) xq[row] = xv;
            if (xq32) xq32[row] = (float)xv;
        }
        y[row] = sum;
        acc = add_prod(acc, xv, sum);
    }
    const double bs = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// y = A x, vector CSR: LPR lanes per row (a power of two), 64/LPR rows per
// wave and U rows per lane group in flight (their column-index, value and
// x loads are issued before any product is used).  Lane `sub` of a row
// sums nonzeros sub, sub+LPR, ... and the LPR partial sums are combined
// with a butterfly over the lane group; no LDS and no barriers.  FUSE: x
// is fl(xs * xscale) formed on the fly from the unscaled source xs, and the
// kernel also stores that scaled vector to xq (the basis column it
// normalises), so the separate scale-copy pass disappears; xq32 (real
// only, may be null) receives its FP32 shadow.  Epilogue: the block's
// partial of Re(x^H y) in part[blockIdx.x].
template <typename S, typename SV, int E, int LPR, int U, bool FUSE>
__device__ __forceinline__ void spmv_vec_body(
    int n, const int *__restrict__ rp, const int *__restrict__ ci, const SV *__restrict__ val,
    const S *__restrict__ x, S *__restrict__ y, double *__restrict__ part, double xscale,
    S *__restrict__ xq, float *__restrict__ xq32, int bid, int nblk)
{
    static_assert(LPR >= 1 && LPR <= 64 && (LPR & (LPR - 1)) == 0, "LPR power of 2");
    constexpr int RPW = 64 / LPR;
    const int lane = threadIdx.x & 63;
    const int sub = lane % LPR, grp = lane / LPR;
    const long wave = (long)bid * (kBlock / 64) + (threadIdx.x >> 6);
    const long nw = (long)nblk * (kBlock / 64);
    auto xval = [&](int c, int e) -> double {
        if constexpr (FUSE)
            return (double)(S)((double)x[(size_t)E * c + e] * xscale);
        else
            return (double)x[(size_t)E * c + e];
    };
    double acc = 0.0;
    for (long base = wave * RPW * U; base < n; base += nw * RPW * U) {
        int kb[U], ke[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long row = base + (long)u * RPW + grp;
            kb[u] = row < n ? rp[row] + sub : 0;
            ke[u] = row < n ? rp[row + 1] : 0;
        }
        double sr[U], si[U];
        int cc[U];
        SV vr[U], vi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            sr[u] = si[u] = 0.0;
            if (kb[u] < ke[u]) {
                cc[u] = ci[kb[u]];
                vr[u] = val[(size_t)E * kb[u]];
                if constexpr (E == 2) vi[u] = val[(size_t)E * kb[u] + 1];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (kb[u] < ke[u]) {
                if constexpr (E == 1) {
                    sr[u] = (double)vr[u] * xval(cc[u], 0);
                } else {
                    const double xr = xval(cc[u], 0), xi = xval(cc[u], 1);
                    sr[u] = (double)vr[u] * xr - (double)vi[u] * xi;
                    si[u] = (double)vr[u] * xi + (double)vi[u] * xr;
                }
            }
            // rows longer than LPR
            for (int k = kb[u] + LPR; k < ke[u]; k += LPR) {
                const int c = ci[k];
                if constexpr (E == 1) {
                    sr[u] += (double)val[k] * xval(c, 0);
                } else {
                    const double ar = val[2 * (size_t)k], ai = val[2 * (size_t)k + 1];
                    const double xr = xval(c, 0), xi = xval(c, 1);
                    sr[u] += ar * xr - ai * xi;
                    si[u] += ar * xi + ai * xr;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1) {
                sr[u] += __shfl_xor(sr[u], o);
                if constexpr (E == 2) si[u] += __shfl_xor(si[u], o);
            }
            const long row = base + (long)u * RPW + grp;
            if (sub == 0 && row < n) {
                if constexpr (E == 1) {
                    const double xr = xval((int)row, 0);
                    y[row] = (S)sr[u];
                    if constexpr (FUSE) {
                        if (xq) xq[row] = (S)xr;
                        if (xq32) xq32[row] = (float)xr;
                    }
                    acc += xr * (double)(S)sr[u];
                } else {
                    const double xr = xval((int)row, 0), xi = xval((int)row, 1);
                    y[2 * row] = sr[u];
                    y[2 * row + 1] = si[u];
                    if constexpr (FUSE) {
                        xq[2 * row] = xr;
                        xq[2 * row + 1] = xi;
                    }
                    acc += xr * sr[u] + xi * si[u];
                }
            }
        }
    }
    const double bs = block_sum(acc);
    if (threadIdx.x == 0) part[bid] = bs;
}

template <typename S, typename SV, int E, int LPR, int U, bool FUSE>
__global__ __launch_bounds__(kBlock) void k_spmv_vec(
    int n, const int *__restrict__ rp, const int *__restrict__ ci, const SV *__restrict__ val,
    const S *__restrict__ x, S *__restrict__ y, double *__restrict__ part, double xscale,
    S *__restrict__ xq, float *__restrict__ xq32)
{
    spmv_vec_body<S, SV, E, LPR, U, FUSE>(n, rp, ci, val, x, y, part, xscale, xq, xq32, blockIdx.x, gridDim.x);
}

// Re(x^H y) over ne doubles: per-block partials in part[blockIdx.x].
template <typename S>
__global__ __launch_bounds__(kBlock) void k_dot(long ne, const S *__restrict__ x,
                                                const S *__restrict__ y, double *__restrict__ part)
{
    double acc = 0.0;
    for (long i = blockIdx.x * kBlock + threadIdx.x; i < ne; i += (long)gridDim.x * kBlock)
        acc += (double)x[i] * (double)y[i];
    const double bs = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// --------------------------------------------------------- Gram-Schmidt

// h = V(:, c0:c0+m)^H r.  gx*gy blocks: each takes one chunk of CB columns
// (of gy) and one row block rb (of gx), whose grid-stride rows it walks RU
// rows of the stride per iteration, so RU*CB loads are in flight per
// thread.  Column indices past the chunk's end are clamped to its last
// column (the extra loads hit cache and are discarded).  Each thread keeps
// CB*E running sums, which are transposed-and-reduced over the wave and
// combined over waves in LDS; the block writes its partial of coefficient p
// to part[p*gx + rb] (k_colsum sums each coefficient's partials in row
// block order).  VS = basis storage (float for the FP32 shadow of a
// mixed-precision basis), RS = storage of r.
// Four consecutive elements p[0..3] (p 16-B aligned for float, 32-B for
// double) with 16-byte loads/stores, as doubles.
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void load4(const float *p, double v[4])
{
    // the FP32 shadow is streamed once per kernel: a non-temporal load
    const f32x4 t = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(p));
    v[0] = t.x;
    v[1] = t.y;
    v[2] = t.z;
    v[3] = t.w;
}
__device__ __forceinline__ void load4(const double *p, double v[4])
{
    const double2 a = reinterpret_cast<const double2 *>(p)[0];
    const double2 b = reinterpret_cast<const double2 *>(p)[1];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
}
// Four consecutive basis entries.  The FP32 shadow is always streamed
// (load4 above); FP64 columns are streamed when NT: the recurrence's few
// columns (measured: the 128^3 Laplacian's mixed path +6.7 %), but not
// the plain path's Gram-Schmidt reads of the FP64 basis (its batch 14 %
// slower streamed).
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <bool NT>
__device__ __forceinline__ void loadv4(const float *p, double v[4])
{
    load4(p, v);
}
template <bool NT>
__device__ __forceinline__ void loadv4(const double *p, double v[4])
{
    if constexpr (NT) {
        const f64x2 a = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(p));
        const f64x2 b = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(p) + 1);
        v[0] = a.x;
        v[1] = a.y;
        v[2] = b.x;
        v[3] = b.y;
    } else {
        load4(p, v);
    }
}
__device__ __forceinline__ void store4(double *p, const double v[4])
{
    reinterpret_cast<double2 *>(p)[0] = double2{v[0], v[1]};
    reinterpret_cast<double2 *>(p)[1] = double2{v[2], v[3]};
}

template <typename VS, typename RS, int E, int CB, int RU>
__device__ __forceinline__ void proj_body(
    int n, const VS *__restrict__ V, long ld, int c0, int m, const RS *__restrict__ r,
    double *__restrict__ part, int gx, int gy, int bid, int nblk)
{
    constexpr int N = CB * E;
    static_assert(N <= 64, "chunk too wide");
    __shared__ double sm[kBlock / 64][N];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // Workgroups go to the 8 XCDs round-robin by linear index.  With gx a
    // multiple of 8, XCD x runs row blocks x, x+8, ... and takes all gy
    // chunks of one row block back to back, so r's stretch is read from
    // HBM once and then served by that XCD's L2.
    int rb, chunk;
    if ((gx & 7) == 0) {
        const int local = bid >> 3;
        rb = (local / gy) * 8 + (bid & 7);
        chunk = local % gy;
    } else {
        rb = bid / gy;
        chunk = bid % gy;
    }
    const int cb = chunk * CB;
    const int last = min(CB, m - cb) - 1;
    // wave-uniform column bases; rows index them
    const VS *col[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) col[c] = V + ((size_t)(c0 + cb + min(c, last)) * ld) * E;
    double acc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = 0.0;
    // FP32 basis: four consecutive rows per thread, one 16-byte load per
    // column.  (For an FP64 basis the strided rows below measured faster:
    // 4.6 vs 4.0 TB/s on the 128^3 Laplacian.)
    constexpr bool kContig = E == 1 && RU == 4 && std::is_same<VS, float>::value;
    if constexpr (kContig) {
        const int stride = gx * kBlock * 4;
        int row = (rb * kBlock + threadIdx.x) * 4;
        for (; row + 3 < n; row += stride) {
            double rv[4], vv[CB][4];
            load4(r + row, rv);
#pragma unroll
            for (int c = 0; c < CB; ++c) loadv4<false>(col[c] + row, vv[c]);
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int c = 0; c < CB; ++c) acc[c] += vv[c][j] * rv[j];
        }
        for (; row < n; ++row) {
            const double rv = (double)r[row];
#pragma unroll
            for (int c = 0; c < CB; ++c) acc[c] += (double)col[c][row] * rv;
        }
    }
    const int stride = gx * kBlock * RU;
    int row = rb * kBlock * RU + threadIdx.x;
    if constexpr (kContig) {
    } else if constexpr (E == 1) {
        for (; row + (RU - 1) * kBlock < n; row += stride) {
            double rv[RU], vv[RU][CB];
#pragma unroll
            for (int j = 0; j < RU; ++j) {
                rv[j] = (double)r[row + j * kBlock];
#pragma unroll
                for (int c = 0; c < CB; ++c) vv[j][c] = (double)col[c][row + j * kBlock];
            }
#pragma unroll
            for (int j = 0; j < RU; ++j)
#pragma unroll
                for (int c = 0; c < CB; ++c) acc[c] += vv[j][c] * rv[j];
        }
#pragma unroll
        for (int j = 0; j < RU; ++j) {
            const int rr = row + j * kBlock;
            if (rr < n) {
                const double rv = (double)r[rr];
#pragma unroll
                for (int c = 0; c < CB; ++c) acc[c] += (double)col[c][rr] * rv;
            }
        }
    } else {
        const double2 *r2 = reinterpret_cast<const double2 *>(r);
        for (; row + (RU - 1) * kBlock < n; row += stride) {
            double2 rv[RU], vv[RU][CB];
#pragma unroll
            for (int j = 0; j < RU; ++j) {
                rv[j] = r2[row + j * kBlock];
#pragma unroll
                for (int c = 0; c < CB; ++c)
                    vv[j][c] = reinterpret_cast<const double2 *>(col[c])[row + j * kBlock];
            }
#pragma unroll
            for (int j = 0; j < RU; ++j)
#pragma unroll
                for (int c = 0; c < CB; ++c) {
                    acc[2 * c] += vv[j][c].x * rv[j].x + vv[j][c].y * rv[j].y;  // conj(v) * r
                    acc[2 * c + 1] += vv[j][c].x * rv[j].y - vv[j][c].y * rv[j].x;
                }
        }
#pragma unroll
        for (int j = 0; j < RU; ++j) {
            const int rr = row + j * kBlock;
            if (rr < n) {
                const double2 rv = r2[rr];
#pragma unroll
                for (int c = 0; c < CB; ++c) {
                    const double2 v = reinterpret_cast<const double2 *>(col[c])[rr];
                    acc[2 * c] += v.x * rv.x + v.y * rv.y;
                    acc[2 * c + 1] += v.x * rv.y - v.y * rv.x;
                }
            }
        }
    }
    const double ws = wave_transpose_sum<N>(acc, lane);
    constexpr int grp = 64 / N;
    if (lane % grp == 0) sm[w][lane / grp] = ws;
    __syncthreads();
    if (threadIdx.x < (last + 1) * E) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < kBlock / 64; ++i) t += sm[i][threadIdx.x];
        part[(size_t)(cb * E + threadIdx.x) * gx + rb] = t;
    }
}

template <typename VS, typename RS, int E, int CB, int RU>
__global__ __launch_bounds__(kBlock) void k_proj(
    int n, const VS *__restrict__ V, long ld, int c0, int m, const RS *__restrict__ r,
    double *__restrict__ part, int gx, int gy)
{
    proj_body<VS, RS, E, CB, RU>(n, V, ld, c0, m, r, part, gx, gy, blockIdx.x, gridDim.x);
}

// Rows row + j*kBlock (j < R) of rout = rin - Vc hs; returns the sum of
// their |rout|^2.  Per row the products are summed into four interleaved
// partial sums (column c into sum c mod 4 within each group of 8 columns,
// E == 1; pairs of columns, E == 2), the same order for any R; R rows at a
// time put R*8 (R*4) column loads in flight per thread.
// r -= V h on the four consecutive rows [row, row+4) of a real basis (row
// a multiple of 4): one 16-byte load per column (two for FP64).  Each row
// sums its columns in the order of update_rows.
template <typename VS, typename RS, bool NTV = false>
__device__ __forceinline__ double update_rows4(int row, const VS *__restrict__ Vc, long ld, int m,
                                               const double *hs, const RS *rin, RS *rout,
                                               double *ro_out = nullptr)
{
    double s[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j][0] = s[j][1] = s[j][2] = s[j][3] = 0.0;
    int c = 0;
    for (; c + 8 <= m; c += 8) {
        double v[8][4];
#pragma unroll
        for (int q = 0; q < 8; ++q) loadv4<NTV>(Vc + (size_t)(c + q) * ld + row, v[q]);
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) s[j][q & 3] += v[q][j] * hs[c + q];
    }
    for (; c < m; ++c) {
        double v[4];
        loadv4<NTV>(Vc + (size_t)c * ld + row, v);
#pragma unroll
        for (int j = 0; j < 4; ++j) s[j][0] += v[j] * hs[c];
    }
    double ri[4], ro[4];
    load4(rin + row, ri);
    double nrm = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        ro[j] = (double)(RS)(ri[j] - ((s[j][0] + s[j][1]) + (s[j][2] + s[j][3])));
        nrm += ro[j] * ro[j];
    }
    store4(rout + row, ro);
    if (ro_out)
#pragma unroll
        for (int j = 0; j < 4; ++j) ro_out[j] = ro[j];
    return nrm;
}

// One complex basis entry; NT streams it (the recurrence's columns).
template <bool NT>
__device__ __forceinline__ double2 ldc(const double2 *p)
{
    if constexpr (NT) {
        const f64x2 t = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(p));
        return make_double2(t.x, t.y);
    } else {
        return *p;
    }
}

template <typename VS, typename RS, int E, int R, bool NTV = false>
__device__ __forceinline__ double update_rows(int row, const VS *__restrict__ Vc, long ld, int m,
                                              const double *hs, const RS *rin, RS *rout)
{
    double nrm = 0.0;
    if constexpr (E == 1) {
        double s[R][4];
#pragma unroll
        for (int j = 0; j < R; ++j) s[j][0] = s[j][1] = s[j][2] = s[j][3] = 0.0;
        int c = 0;
        for (; c + 8 <= m; c += 8) {
            double v[R][8];
#pragma unroll
            for (int j = 0; j < R; ++j)
#pragma unroll
                for (int q = 0; q < 8; ++q) v[j][q] = (double)Vc[(size_t)(c + q) * ld + row + j * kBlock];
#pragma unroll
            for (int j = 0; j < R; ++j)
#pragma unroll
                for (int q = 0; q < 8; ++q) s[j][q & 3] += v[j][q] * hs[c + q];
        }
        for (; c < m; ++c)
#pragma unroll
            for (int j = 0; j < R; ++j) s[j][0] += (double)Vc[(size_t)c * ld + row + j * kBlock] * hs[c];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int rr = row + j * kBlock;
            const RS rn = (RS)((double)rin[rr] - ((s[j][0] + s[j][1]) + (s[j][2] + s[j][3])));
            rout[rr] = rn;
            nrm += (double)rn * (double)rn;
        }
    } else {
        const double2 *V2 = reinterpret_cast<const double2 *>(Vc);
        double ar[R], ai[R], br[R], bi[R];
#pragma unroll
        for (int j = 0; j < R; ++j) ar[j] = ai[j] = br[j] = bi[j] = 0.0;
        int c = 0;
        for (; c + 4 <= m; c += 4) {
            double2 v[R][4];
#pragma unroll
            for (int j = 0; j < R; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) v[j][q] = ldc<NTV>(V2 + (size_t)(c + q) * ld + row + j * kBlock);
#pragma unroll
            for (int j = 0; j < R; ++j)
#pragma unroll
                for (int q = 0; q < 4; q += 2) {
                    const double h0r = hs[2 * (c + q)], h0i = hs[2 * (c + q) + 1];
                    const double h1r = hs[2 * (c + q + 1)], h1i = hs[2 * (c + q + 1) + 1];
                    ar[j] += v[j][q].x * h0r - v[j][q].y * h0i;
                    ai[j] += v[j][q].x * h0i + v[j][q].y * h0r;
                    br[j] += v[j][q + 1].x * h1r - v[j][q + 1].y * h1i;
                    bi[j] += v[j][q + 1].x * h1i + v[j][q + 1].y * h1r;
                }
        }
        for (; c < m; ++c)
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const double2 v = ldc<NTV>(V2 + (size_t)c * ld + row + j * kBlock);
                ar[j] += v.x * hs[2 * c] - v.y * hs[2 * c + 1];
                ai[j] += v.x * hs[2 * c + 1] + v.y * hs[2 * c];
            }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int rr = row + j * kBlock;
            const double2 ri = reinterpret_cast<const double2 *>(rin)[rr];
            double2 ro;
            ro.x = ri.x - (ar[j] + br[j]);
            ro.y = ri.y - (ai[j] + bi[j]);
            reinterpret_cast<double2 *>(rout)[rr] = ro;
            nrm += ro.x * ro.x + ro.y * ro.y;
        }
    }
    return nrm;
}

// rout = rin - V(:, c0:c0+m) h ; per-block partials of ||rout||^2 in
// part[blockIdx.x] ; ||h||^2 -> out_hh.  h (E*m doubles, FP64) is staged in
// LDS.  When cpart is given, coefficient `cslot` is not read from h but is
// the sum of the cnp partials in cpart (the alpha that the preceding
// SpMV/dot left as per-block partials); every block forms that sum in the
// same order, and block 0 also stores it to out_coef.  Without h the
// coefficients are (h0, 0, ..., 0) apart from cslot.  A block takes RU
// rows of its grid stride per iteration.  rin may equal rout.  VS = basis
// storage, RS = storage of r.
template <typename VS, typename RS, int E, int RU, bool NTV = false>
__device__ __forceinline__ void update_body(
    int n, const VS *__restrict__ V, long ld, int c0, int m, const double *__restrict__ h, double h0,
    const double *__restrict__ cpart, int cnp, int cslot, const RS *rin, RS *rout,
    double *__restrict__ part, double *__restrict__ out_hh, double *__restrict__ out_coef, int bid, int nblk)
{
    extern __shared__ double hs[];
    for (int i = threadIdx.x; i < E * m; i += kBlock) hs[i] = h ? h[i] : (i == 0 ? h0 : 0.0);
    if (cpart) {
        const double a = ordered_sum(cpart, cnp);  // contains __syncthreads
        if (threadIdx.x == 0) {
            hs[E * cslot] = a;
            if (E == 2) hs[E * cslot + 1] = 0.0;
            if (bid == 0 && out_coef) out_coef[0] = a;
        }
    }
    __syncthreads();
    const VS *Vc = V + (size_t)c0 * ld * E;
    double acc = 0.0;
    if constexpr (E == 1 && RU == 4) {
        // four consecutive rows per thread (16-byte loads)
        const int stride = nblk * kBlock * 4;
        int row = (bid * kBlock + threadIdx.x) * 4;
        for (; row + 3 < n; row += stride) acc += update_rows4<VS, RS, NTV>(row, Vc, ld, m, hs, rin, rout);
        for (; row < n; ++row) acc += update_rows<VS, RS, E, 1>(row, Vc, ld, m, hs, rin, rout);
    } else {
        const int stride = nblk * kBlock * RU;
        int row = bid * kBlock * RU + threadIdx.x;
        for (; row + (RU - 1) * kBlock < n; row += stride)
            acc += update_rows<VS, RS, E, RU, NTV>(row, Vc, ld, m, hs, rin, rout);
#pragma unroll
        for (int j = 0; j < RU - 1; ++j)
            if (row + j * kBlock < n)
                acc += update_rows<VS, RS, E, 1, NTV>(row + j * kBlock, Vc, ld, m, hs, rin, rout);
    }
    const double bs = block_sum(acc);
    if (threadIdx.x == 0) part[bid] = bs;
    if (out_hh && bid == 0 && threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < E * m; ++i) t += hs[i] * hs[i];
        out_hh[0] = t;
    }
}

template <typename VS, typename RS, int E, int RU, bool NTV = false>
__global__ __launch_bounds__(kBlock) void k_update(
    int n, const VS *__restrict__ V, long ld, int c0, int m, const double *__restrict__ h, double h0,
    const double *__restrict__ cpart, int cnp, int cslot, const RS *rin, RS *rout,
    double *__restrict__ part, double *__restrict__ out_hh, double *__restrict__ out_coef)
{
    update_body<VS, RS, E, RU, NTV>(n, V, ld, c0, m, h, h0, cpart, cnp, cslot, rin, rout, part, out_hh, out_coef, blockIdx.x, gridDim.x);
}

// The recurrence and the first Gram-Schmidt projection of a real step in
// one pass, over the FP32 shadow (VS = float) or the FP64 basis (VS =
// double) (n <= kMaxGrid * kMinRowsUpd rows):
// block rb owns rows [1024 rb, 1024 rb + 1024), four consecutive rows per
// thread.  It forms alpha from the SpMV partials and r = y - V(:, c0:c0+rm)
// h in place exactly as update_body does for the recurrence (the same
// rows per block and thread, so its ||r||^2 partial is the recurrence's,
// at part_recur[rb]), keeps its rows of r in registers, and for each chunk
// of CB columns of the basis forms the projection partial exactly
// as proj_body does for one row block (part_proj[p * gx + rb], gx the row
// blocks): r is neither written nor read twice, and one kernel replaces
// two.
template <int CB, typename VS>
__device__ __forceinline__ void recproj_body(int n, const double *__restrict__ V, long ld, int c0,
                                             int rm, const double *__restrict__ h, double h0,
                                             const double *__restrict__ cpart, int cnp,
                                             double *__restrict__ r, double *__restrict__ part_recur,
                                             double *__restrict__ out_coef,
                                             const VS *__restrict__ VG, int m,
                                             double *__restrict__ part_proj, int gx, int rb)
{
    extern __shared__ double hs[];
    __shared__ double sm[kBlock / 64][CB];
    for (int i = threadIdx.x; i < rm; i += kBlock) hs[i] = h ? h[i] : (i == 0 ? h0 : 0.0);
    {
        const double a = ordered_sum(cpart, cnp);  // contains __syncthreads
        if (threadIdx.x == 0) {
            hs[rm - 1] = a;
            if (rb == 0 && out_coef) out_coef[0] = a;
        }
    }
    __syncthreads();
    const double *Vc = V + (size_t)c0 * ld;
    const int row = (rb * kBlock + threadIdx.x) * 4;
    double rv[4] = {0.0, 0.0, 0.0, 0.0};
    double acc = 0.0;
    const bool four = row + 3 < n;
    if (four) {
        acc += update_rows4<double, double, true>(row, Vc, ld, rm, hs, r, r, rv);
    } else {
        for (int j = 0; row + j < n; ++j) {
            acc += update_rows<double, double, 1, 1>(row + j, Vc, ld, rm, hs, r, r);
            rv[j] = r[row + j];
        }
    }
    const double bs = block_sum(acc);
    if (threadIdx.x == 0) part_recur[rb] = bs;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int cb = 0; cb < m; cb += CB) {
        const int last = min(CB, m - cb) - 1;
        const VS *col[CB];
#pragma unroll
        for (int c = 0; c < CB; ++c) col[c] = VG + (size_t)(cb + min(c, last)) * ld;
        double pacc[CB];
#pragma unroll
        for (int c = 0; c < CB; ++c) pacc[c] = 0.0;
        if (four) {
            double vv[CB][4];
#pragma unroll
            for (int c = 0; c < CB; ++c) loadv4<false>(col[c] + row, vv[c]);
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int c = 0; c < CB; ++c) pacc[c] += vv[c][j] * rv[j];
        } else {
            for (int j = 0; row + j < n; ++j)
#pragma unroll
                for (int c = 0; c < CB; ++c) pacc[c] += (double)col[c][row + j] * rv[j];
        }
        const double ws = wave_transpose_sum<CB>(pacc, lane);
        constexpr int grp = 64 / CB;
        if (lane % grp == 0) sm[w][lane / grp] = ws;
        __syncthreads();
        if (threadIdx.x <= last) {
            double t = 0.0;
#pragma unroll
            for (int i = 0; i < kBlock / 64; ++i) t += sm[i][threadIdx.x];
            part_proj[(size_t)(cb + threadIdx.x) * gx + rb] = t;
        }
        __syncthreads();  // sm is reused by the next chunk
    }
}

template <int CB, typename VS>
__global__ __launch_bounds__(kBlock) void k_recproj(int n, const double *__restrict__ V, long ld,
                                                    int c0, int rm, const double *__restrict__ h,
                                                    double h0, const double *__restrict__ cpart,
                                                    int cnp, double *__restrict__ r,
                                                    double *__restrict__ part_recur,
                                                    double *__restrict__ out_coef,
                                                    const VS *__restrict__ VG, int m,
                                                    double *__restrict__ part_proj)
{
    recproj_body<CB, VS>(n, V, ld, c0, rm, h, h0, cpart, cnp, r, part_recur, out_coef, VG, m,
                         part_proj, gridDim.x, blockIdx.x);
}

// dst = s * src over ne storage elements (dst may equal src); with dst32
// also its FP32 rounding (the shadow column of a mixed-precision basis).
template <typename S>
__global__ __launch_bounds__(kBlock) void k_scale_copy(long ne, const S *src, S *dst, double s,
                                                       float *__restrict__ dst32)
{
    for (long i = blockIdx.x * kBlock + threadIdx.x; i < ne; i += (long)gridDim.x * kBlock) {
        const S v = (S)((double)src[i] * s);
        dst[i] = v;
        if (dst32) dst32[i] = (float)v;
    }
}

// dst = FP64 of an FP32 device vector (and its FP32 shadow copy).
__global__ __launch_bounds__(kBlock) void k_copy_f32(long n, const float *__restrict__ src,
                                                     double *__restrict__ dst, float *__restrict__ dst32)
{
    for (long i = blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        const float v = src[i];
        dst[i] = (double)v;
        if (dst32) dst32[i] = v;
    }
}

// ------------------------------------------------------------- restart

// In-place restart GEMM V(:, lck:lck+k) <- V(:, lck:lck+jl) * Y on FP64
// MFMA (v_mfma_f64_16x16x4_f64).  Each wave owns 16-row tiles (rows
// disjoint across waves), reads its rows over all jl input columns into
// the accumulators before writing any output column, so the update is safe
// in place.  Y (jl x k) is staged in LDS as Ys[kk * kp + col], padded with
// zeros to jl4 = ceil4(jl) rows and kp = 16*KT columns.  With V32 (the
// FP32 shadow slab of a mixed-precision basis, same ld) each output is
// also stored rounded to FP32.
// MFMA f64 16x16x4 operand layout (lane l): A[l%16][l/16], B[l/16][l%16],
// D[l/16 + 4*r][l%16] for accumulator element r (measured on gfx950).
// The wave computes the transposed tile, out^T = Y^T V^T (A = a 16-column
// slice of Y^T, B = V^T), so that accumulator element r of lane l is
// output row l%16 of column l/16 + 4r: each store covers four columns of
// 16 consecutive rows (128-byte segments) instead of sixteen 32-byte ones.
// The products and their order over kk are those of V * Y.
typedef double d4 __attribute__((ext_vector_type(4)));
// 16-row tiles per wave of the restart GEMM.  Swept on the batched 2D
// Anderson workload: 1 tile 2.65, 2 tiles 2.54, 4 tiles 2.33, 8 tiles 1.93
// TB/s (more blocks beat amortising the LDS staging of Y).
constexpr int kRitzTiles = 1;
// Input columns whose loads a wave issues before their first MFMA.
constexpr int kRitzLoads = 8;

// One block's rows of the restart GEMM (block bid; tiles_per_wave 16-row
// tiles per wave); Ys is the block's LDS (jl4 * 16 * KT doubles).
template <int KT>
__device__ __forceinline__ void ritz_body(long nrows, double *__restrict__ V, long ld, int lck, int jl,
                                          int k, const double *__restrict__ Y, int ldy,
                                          int tiles_per_wave, float *__restrict__ V32,
                                          double *__restrict__ Ys, int bid)
{
    constexpr int kp = 16 * KT;
    const int jl4 = (jl + 3) & ~3;
    for (int i = threadIdx.x; i < jl4 * kp; i += kBlock) {
        const int kk = i / kp, col = i % kp;
        Ys[i] = (kk < jl && col < k) ? Y[(size_t)col * ldy + kk] : 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lq = lane >> 4;
    double *Vin = V + (size_t)lck * ld;
    for (int t = 0; t < tiles_per_wave; ++t) {
        const long row0 = ((long)(bid * (kBlock / 64) + w) * tiles_per_wave + t) * 16;
        if (row0 >= nrows) break;
        d4 acc[KT];
#pragma unroll
        for (int q = 0; q < KT; ++q) acc[q] = d4{0.0, 0.0, 0.0, 0.0};
        const long arow = row0 + li;
        for (int k0 = 0; k0 < jl4; k0 += 4 * kRitzLoads) {
            double a[kRitzLoads];
#pragma unroll
            for (int u = 0; u < kRitzLoads; ++u) {
                const int col = k0 + 4 * u + lq;
                a[u] = (arow < nrows && col < jl) ? Vin[(size_t)col * ld + arow] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kRitzLoads; ++u) {
                const int kk = k0 + 4 * u;
                if (kk >= jl4) break;
#pragma unroll
                for (int q = 0; q < KT; ++q) {
                    const double b = Ys[(kk + lq) * kp + q * 16 + li];
                    acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a[u], acc[q], 0, 0, 0);
                }
            }
        }
        if (arow < nrows) {
#pragma unroll
            for (int q = 0; q < KT; ++q) {
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int col = q * 16 + lq + 4 * rr;
                    if (col < k) {
                        Vin[(size_t)col * ld + arow] = acc[q][rr];
                        if (V32) V32[(size_t)(lck + col) * ld + arow] = (float)acc[q][rr];
                    }
                }
            }
        }
    }
}

template <int KT>
__global__ __launch_bounds__(kBlock) void k_ritz_mfma(
    long nrows, double *__restrict__ V, long ld, int lck, int jl, int k, const double *__restrict__ Y,
    int ldy, int tiles_per_wave, float *__restrict__ V32)
{
    extern __shared__ double Ys[];
    ritz_body<KT>(nrows, V, ld, lck, jl, k, Y, ldy, tiles_per_wave, V32, Ys, blockIdx.x);
}

// ---------------------------------------------------------------- stats

struct Stats {
    int on = 0;      // 0 off, else time every on-th launch of each kind
    long seen[TRLHIP_K_COUNT] = {};
    double launches[TRLHIP_K_COUNT] = {};
    double ms[TRLHIP_K_COUNT] = {};
    double bytes[TRLHIP_K_COUNT] = {};
    double flops[TRLHIP_K_COUNT] = {};
    struct Pending {
        hipEvent_t a, b;
        int k;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
};
Stats g_stats;
std::mutex g_stats_mu;  // batched solves launch from several host threads

hipEvent_t ev_get()
{
    std::lock_guard<std::mutex> lk(g_stats_mu);
    if (!g_stats.pool.empty()) {
        hipEvent_t e = g_stats.pool.back();
        g_stats.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Resolve timings of completed launches (called after a stream sync);
// launches still running on another stream stay pending.
void stats_resolve()
{
    std::lock_guard<std::mutex> lk(g_stats_mu);
    size_t keep = 0;
    for (auto &p : g_stats.pending) {
        if (hipEventQuery(p.b) == hipErrorNotReady) {
            g_stats.pending[keep++] = p;
            continue;
        }
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) g_stats.ms[p.k] += ms;
        g_stats.pool.push_back(p.a);
        g_stats.pool.push_back(p.b);
    }
    g_stats.pending.resize(keep);
}

// Start/stop events armed for the next kernel launch.  KLAUNCH passes them
// to hipExtLaunchKernelGGL, which stamps them from the kernel's own
// dispatch (start and end of the kernel, no marker packets in the stream).
struct LaunchEvents {
    hipEvent_t a = nullptr, b = nullptr;
};
thread_local LaunchEvents g_launch_ev;

#define KLAUNCH(kernel, grid, block, shm, stream, ...)                                      \
    do {                                                                                    \
        const LaunchEvents ev_ = g_launch_ev;                                               \
        g_launch_ev = LaunchEvents{};                                                       \
        hipExtLaunchKernelGGL(kernel, grid, block, shm, stream, ev_.a, ev_.b, 0, __VA_ARGS__); \
    } while (0)

// The first of several kernels that one timed scope launches back to back:
// it takes the start event and leaves the stop event armed for the scope's
// last launch (a plain KLAUNCH), so the timing spans all of them.
#define KLAUNCH_FIRST(kernel, grid, block, shm, stream, ...)                                    \
    do {                                                                                        \
        const LaunchEvents ev_ = g_launch_ev;                                                   \
        g_launch_ev = LaunchEvents{nullptr, ev_.b};                                             \
        hipExtLaunchKernelGGL(kernel, grid, block, shm, stream, ev_.a, nullptr, 0, __VA_ARGS__); \
    } while (0)

// Times the first kernel launched in its scope when enabled: one scope in
// g_stats.on of each kind, picked by a multiplicative hash of its ordinal
// (a sample that does not alias with the periodic growth of the basis
// during a restart cycle).  Launches and bytes count the timed launches
// only, so bytes/ms stays the timed launches' rate.
bool timed_pick(int k)
{
    if (!g_stats.on) return false;
    std::lock_guard<std::mutex> lk(g_stats_mu);
    const uint64_t i = (uint64_t)g_stats.seen[k]++;
    return g_stats.on == 1 || ((i * 0x9E3779B97F4A7C15ull) >> 40) % (uint64_t)g_stats.on == 0;
}

struct Timed {
    hipEvent_t a = nullptr, b = nullptr;
    int k;
    double bytes, flops;
    Timed(int kind, hipStream_t, double nbytes, double nflops = 0.0)
        : k(kind), bytes(nbytes), flops(nflops)
    {
        if (!timed_pick(k)) return;
        a = ev_get();
        b = ev_get();
        if (a && b) g_launch_ev = LaunchEvents{a, b};
    }
    ~Timed()
    {
        if (!a && !b) return;
        std::lock_guard<std::mutex> lk(g_stats_mu);
        if (!a || !b) {
            if (a) g_stats.pool.push_back(a);
            if (b) g_stats.pool.push_back(b);
            return;
        }
        if (g_launch_ev.a == a) {  // nothing was launched in the scope
            g_launch_ev = LaunchEvents{};
            g_stats.pool.push_back(a);
            g_stats.pool.push_back(b);
            return;
        }
        g_stats.launches[k] += 1.0;
        g_stats.bytes[k] += bytes;
        g_stats.flops[k] += flops;
        g_stats.pending.push_back({a, b, k});
    }
};

}  // namespace

// ------------------------------------------------------------ structures

struct trlhip_csr {
    uint64_t magic;
    int n;
    long nnz;
    int vtype;
    int E;          // doubles per value (complex 2)
    int nb;         // SpMV row blocks (CSR-stream kernel)
    int lpr;        // lanes per row of the vector-CSR kernel in use; 0: CSR-stream
    int lpr_vec;    // the vector kernel's lanes per row for this matrix (0: unsuitable)
    int vgrid;      // blocks of the vector-CSR kernel
    int *rb;        // nb + 1 row-block starts
    int *rp;
    int *ci;
    double *val;    // FP64 values (E per nonzero)
    // packed nonzeros for the CSR-stream kernel (real matrices with at most
    // 255 distinct off-diagonal values; null otherwise): per nonzero a value
    // index and the column relative to its row block's first row
    uint32_t *pk;
    double *dg;     // diagonal values (n), value index 255
    double *tab;    // the off-diagonal values, indices 0..254
    int ntab;
    // periodic nearest-neighbour stencil on an st_L^st_dim lattice with one
    // off-diagonal value st_off (and the packed diagonal dg); st_dim = 0:
    // not a stencil (see spmv_stencil_body)
    int st_dim, st_L;
    double st_off;
    // column panels (see k_spmv_panel); pn_on: the SpMV in use
    int pn_on, pn_nchunk, pn_nslot, pn_grid;
    long pn_nseg;
    int *pn_perm, *pn_segbase, *pn_erow, *pn_ecol, *pn_rsptr, *pn_rsidx;
    long *pn_cstart;
    double *pn_eval, *pn_sp;
    // staging for host-called matvec
    double *dx, *dy, *part, *dscal;
    hipStream_t s;
};

struct trlhip_ctx {
    int n;
    int ld;         // elements per column
    int ncols;      // user columns; column ncols is an internal scratch
    int vtype;
    int E;          // doubles per element (complex 2)
    size_t esz;     // bytes per element
    char *V;
    float *V32;     // FP32 shadow of the basis (mixed precision), same ld
    double *part;
    size_t partcap; // doubles
    int np_alpha;   // partial count of the last SpMV / dot
    int np_recur;   // partial count of a deferred recurrence norm
    double *dscal;  // 16 device scalars
    double *dcoef;  // 2*(ncols+2) recurrence / restart-step coefficients
    double *dh;     // 2*(ncols+2) CGS projection
    double *dY;     // restart matrix
    size_t dYcap;
    double *hbuf;   // pinned host mirror: 16 scalars + coefficients
    double *hres;   // pinned host-coherent results: a fused step's 4 scalars, a
                    // batched column norm at kHresNorm, then 3 per refined pair
                    // from index 8 (hres_count doubles)
    double *dres;   // its device address
    int pend_dst;   // column of a deferred scale-copy (-1: none)
    double pend_s;  // its scale
    int pend_src;   // its source: the scratch column, or (a deferred column copy) a basis column
    size_t hcap;
    double *hY;     // pinned staging for Y
    hipStream_t s;
    trlhip_batch *batch;  // batch the context steps in (its stream is the batch's), or null
    int arena;            // memory comes from the batch's arenas (freed with the batch)
    int nrec;             // recorded restart GEMMs / scale-copies not yet flushed (batch)
    char *hops, *dops;    // staging of a refinement outside a batch (pinned / device)
    size_t ocap;
    char *hstage;         // batch: pinned staging of one column for col_set / col_get
    // chained steps (batch): the solver's hint for its next step, and the
    // followers recorded with the last step
    int ch_budget, ch_kept, ch_ntot;  // hint: steps the solver allows (<= 1: none)
    int ch_n, ch_next;                // steps in the chain; next follower to hand out (0: none)
    int ch_copy_ok;                   // that follower's fused normalisation was claimed
    int ch_qa[kChainMax], ch_qb[kChainMax], ch_rcol[kChainMax], ch_jnd[kChainMax];
};

// One recorded Lanczos step of a batched context (see the batch section).
struct BatchStep {
    trlhip_ctx *c;
    const trlhip_csr *A;
    int first, qa, qb, rcol, jnd;
    double beta_prev;
    int fuse;         // q_a is a deferred scale-copy, of column fuse_src
    double fuse_s;
    long coff;        // first step after a restart: its jnd-1 coefficients at cbuf[coff]
    int fuse_src;
    int ci;           // position in its context's chain (0: the step the solver asked for)
    int decide;       // a follower comes next: decide on the device whether it runs
    int kept, ntot;   // the solver's kept and info->ntot (for that decision)
};

// A recorded restart GEMM, scale-copy, refinement or column norm of a
// batched context; `phase` is its ordinal among the context's recorded
// operations since the last flush (a flush runs phase 0 of every context,
// then phase 1, ...).
enum { OP_RITZ = 0, OP_COPY = 1, OP_REFINE = 2, OP_NORM = 3 };
struct BatchOp {
    trlhip_ctx *c;
    int type, phase;
    int lck, jl, k, ldy;  // restart GEMM; Y at ybuf[yoff], ldy x k
    size_t yoff;
    int dst, src;         // scale-copy dst = s * src; norm: ||dst||^2
    double s;
    const float *src32;   // copy from an FP32 device vector instead of column src
    const trlhip_csr *A;  // refinement of columns [0, k) (k above), `avoid` kept
    int avoid;
};

namespace {
struct BDesc;  // defined with the batched kernels
}

// Lock for the batch's record lists: held for one push_back by the
// coroutines' threads, so a waiter spins (pause) instead of sleeping in the
// kernel; a std::mutex under that contention costs futex round trips.
// After a few thousand spins (the holder was descheduled) a waiter yields
// its CPU between checks.
struct SpinLock {
    std::atomic<int> f{0};
    void lock()
    {
        for (int spins = 0;;) {
            if (!f.exchange(1, std::memory_order_acquire)) return;
            while (f.load(std::memory_order_relaxed)) {
                if (++spins < 4096)
                    __builtin_ia32_pause();
                else
                    std::this_thread::yield();
            }
        }
    }
    void unlock() { f.store(0, std::memory_order_release); }
};

struct trlhip_batch {
    hipStream_t s = nullptr;      // the batched launches
    bool own_stream = true;       // false: s belongs to another batch
    hipEvent_t done = nullptr;    // recorded after each launch (trlhip_batch_wait)
    // the contexts' own (immediate) work: copies, extra Gram-Schmidt passes.
    // Kept off s, so a blocking copy from pageable memory waits only for
    // this batch's finished work, never for another lane's queued launch;
    // each launch waits for it (side_done) to keep program order.
    hipStream_t side = nullptr;
    hipEvent_t side_done = nullptr;
    void (*yield)(void *) = nullptr;
    void *user = nullptr;
    int err = 0;                  // sticky device error of a flush
    SpinLock mu;                  // guards the records (coroutines run on several host threads)
    std::mutex arena_mu;          // guards the arenas (contexts are created on several threads)
    std::vector<BatchStep> steps;
    std::vector<double> cbuf;     // restart-step coefficients of the recorded steps
    std::vector<BatchOp> ops;     // recorded restart GEMMs and scale-copies
    std::vector<double> ybuf;     // their restart matrices
    // staging of the recorded operations and steps, copied with one
    // transfer per launch
    char *hstg = nullptr, *dstg = nullptr;      // pinned / device
    size_t scap = 0;              // bytes each holds
    // arenas the batch's contexts allocate from (freed with the batch):
    // device memory zeroed when a chunk is taken, and pinned host-coherent
    // memory with its device address
    struct Chunk {
        char *p, *dp;
        size_t size, used;
    };
    std::vector<Chunk> dev, host;
};

namespace {

enum { S_ALPHA = 0, S_RR2 = 1, S_HH = 2, S_RR2B = 3 };
// Chained batched steps (see trlhip_step_cgs): the step before a follower
// leaves the follower's go flag, its fused normalisation 1/beta and beta
// (the follower's beta_prev) in these device scalars.
enum { S_GO = 4, S_XS = 5, S_BETA = 6 };

int grid_for(long work, int per_block_min, int cap)
{
    long g = (work + per_block_min - 1) / per_block_min;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}


template <typename S>
S *colp(trlhip_ctx *c, int col)
{
    return reinterpret_cast<S *>(c->V + (size_t)col * c->ld * c->esz);
}

// Bump allocation from a batch arena (256-byte aligned).  Device chunks of
// at least 1 GiB are zeroed on the batch stream when taken; host chunks of
// at least 64 MiB are pinned, host-coherent and mapped (dptr: the device
// address of the returned block).
// Arena chunks outlive their batch: a destroyed batch hands them here and
// the next batch reuses them (a pinned or device allocation of this size
// costs milliseconds).  Bytes past a chunk's `used` mark are zero; a reused
// chunk has its used part zeroed again.  Held until trlhip_batch_trim() or
// the end of the process.
std::mutex g_chunk_mu;
std::vector<trlhip_batch::Chunk> g_free_chunks[2];  // device, host

// Pinned staging buffers for batched multi-column transfers to the host
// (trlhip_cols_get): taken for one transfer and given back, sizes powers of
// two from 1 MiB, kept until trlhip_batch_trim() or the end of the process
// (no zeroing: a transfer overwrites what it reads back).
struct StageBuf {
    char *p;
    size_t size;
};
std::mutex g_stage_mu;
std::vector<StageBuf> g_stage_free;

// A batch's record staging (pinned + device, `cap` bytes each) outlives it
// too: the next batch takes it instead of allocating (a pinned and a
// device allocation cost milliseconds, and freeing device memory
// synchronizes the device).  Guarded by g_chunk_mu.
struct StagePair {
    char *h, *d;
    size_t cap;
};
std::vector<StagePair> g_stage_pairs;

StageBuf stage_take(size_t bytes)
{
    size_t sz = (size_t)1 << 20;
    while (sz < bytes) sz <<= 1;
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        for (size_t i = 0; i < g_stage_free.size(); ++i)
            if (g_stage_free[i].size == sz) {
                const StageBuf b = g_stage_free[i];
                g_stage_free.erase(g_stage_free.begin() + (long)i);
                return b;
            }
    }
    StageBuf b{nullptr, sz};
    if (hipHostMalloc((void **)&b.p, sz, 0) != hipSuccess) b.p = nullptr;
    return b;
}

void stage_give(const StageBuf &b)
{
    std::lock_guard<std::mutex> lk(g_stage_mu);
    g_stage_free.push_back(b);
}

void *arena_get(trlhip_batch *b, bool host, size_t bytes, void **dptr = nullptr)
{
    bytes = (bytes + 255) & ~(size_t)255;
    auto &v = host ? b->host : b->dev;
    if (v.empty() || v.back().size - v.back().used < bytes) {
        const size_t sz = std::max(bytes, host ? (size_t)64 << 20 : (size_t)1 << 30);
        trlhip_batch::Chunk ch{nullptr, nullptr, sz, 0};
        bool reused = false;
        {
            std::lock_guard<std::mutex> lk(g_chunk_mu);
            auto &fl = g_free_chunks[host ? 1 : 0];
            for (size_t i = 0; i < fl.size(); ++i)
                if (fl[i].size >= bytes) {
                    ch = fl[i];
                    fl.erase(fl.begin() + (long)i);
                    reused = true;
                    break;
                }
        }
        if (reused) {
            if (host)
                memset(ch.p, 0, ch.used);
            else if (ch.used && hipMemsetAsync(ch.p, 0, ch.used, b->side) != hipSuccess)
                return nullptr;
            ch.used = 0;
        } else if (host) {
            if (hipHostMalloc((void **)&ch.p, sz, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess)
                return nullptr;
            if (hipHostGetDevicePointer((void **)&ch.dp, ch.p, 0) != hipSuccess) {
                (void)hipHostFree(ch.p);
                return nullptr;
            }
            memset(ch.p, 0, sz);
        } else {
            if (hipMalloc((void **)&ch.p, sz) != hipSuccess) return nullptr;
            if (hipMemsetAsync(ch.p, 0, sz, b->side) != hipSuccess) {
                (void)hipFree(ch.p);
                return nullptr;
            }
            ch.dp = ch.p;
        }
        v.push_back(ch);
    }
    auto &ch = v.back();
    void *r = ch.p + ch.used;
    if (dptr) *dptr = ch.dp + ch.used;
    ch.used += bytes;
    return r;
}

int ensure_part(trlhip_ctx *c, size_t need)
{
    if (c->partcap >= need) return 0;
    if (c->part) {
        (void)hipStreamSynchronize(c->s);
        (void)hipFree(c->part);
    }
    c->part = nullptr;
    c->partcap = 0;
    if (hipMalloc((void **)&c->part, sizeof(double) * need) != hipSuccess) return TRLHIP_ENOMEM;
    c->partcap = need;
    return 0;
}

// Partial-sum regions of c->part: alpha partials (SpMV / dot), update
// partials (||r||^2), projection partial rows.
double *part_alpha(trlhip_ctx *c) { return c->part; }
double *part_update(trlhip_ctx *c) { return c->part + kSpmvGrid; }
double *part_recur(trlhip_ctx *c) { return c->part + kSpmvGrid + kMaxGrid; }
double *part_proj(trlhip_ctx *c) { return c->part + kSpmvGrid + 2 * kMaxGrid; }
size_t hres_chain(int ncols) { return 8 + 3 * (size_t)(ncols + 1); }
constexpr int kHresNorm = 4;  // a batched column norm (OP_NORM)
// host-coherent results: a step's 4 scalars, 3 per refined pair from index
// 8, then 7 per chain follower from hres_chain (alpha, ||r||^2, ||h||^2,
// ||r - V h||^2, go, 1/beta, beta)
size_t hres_count(int ncols) { return hres_chain(ncols) + 7 * (size_t)kChainMax; }
size_t part_need(int E, int ncols)
{
    return (size_t)kSpmvGrid + 2 * (size_t)kMaxGrid + (size_t)kProjGrid * E * (ncols + 1) + 64;
}

double spmv_bytes(const trlhip_csr *A, int E, size_t vsz, size_t xsz)
{
    // values + col_idx + row_ptr + x read once + y written once; packed:
    // one word per nonzero + row_ptr + the diagonal + x + y.  Column panels
    // count as the CSR they were built from: their (row, col, value)
    // triples and partial sums are the method's overhead, not the product's
    // bytes.
    if (A->st_dim && !A->lpr)  // stencil: row blocks + the diagonal + x + y
        return 4.0 * (A->nb + 1) + 8.0 * A->n + 2.0 * (double)xsz * A->n;
    if (A->pk && !A->lpr)
        return 4.0 * (double)A->nnz + 4.0 * (A->n + 1) + 8.0 * A->n + 8.0 * A->ntab +
               2.0 * (double)xsz * A->n;
    return (double)A->nnz * ((double)vsz * E + 4.0) + 4.0 * (A->n + 1) + 2.0 * (double)xsz * E * A->n;
}

// The SpMV kernel a real matrix runs: lanes per row of the vector kernel,
// 0 for CSR-stream, -1 for CSR-stream over packed nonzeros, -2 / -3 for
// the 2D / 3D stencil kernel.
int spmv_kind(const trlhip_csr *A)
{
    return A->lpr ? A->lpr : (A->pn_on ? -4 : (A->st_dim ? -A->st_dim : (A->pk ? -1 : 0)));
}

// The kernel a batched step runs: the column-panel kernel has no batched
// form (its operators are far larger than a batch's), they run CSR-stream.
int spmv_kind_batch(const trlhip_csr *A) { return A->pn_on && !A->lpr ? 0 : spmv_kind(A); }

int spmv_grid_stream(const trlhip_csr *A)
{
    const int g = (A->nb + kMinSpmvRowBlocks - 1) / kMinSpmvRowBlocks;
    return g < kSpmvGrid ? g : kSpmvGrid;
}

int spmv_grid(const trlhip_csr *A)
{
    if (A->lpr) return A->vgrid;
    if (A->pn_on) return A->pn_grid;
    const int g = (A->nb + kMinSpmvRowBlocks - 1) / kMinSpmvRowBlocks;
    return g < kSpmvGrid ? g : kSpmvGrid;
}

// The fused recurrence + first projection (recproj_body, over the FP32
// shadow or the FP64 basis) applies to real contexts whose rows fit
// kMaxGrid row blocks of kMinRowsUpd (one recurrence block per row block).
bool recproj_ok(const trlhip_ctx *c)
{
    return c->E == 1 && (long)c->n <= (long)kMaxGrid * kMinRowsUpd;
}

// Workgroups of the stencil SpMV: kStRB row blocks each (one alpha partial
// per row block) when A has at most kSpmvGrid row blocks, else the
// CSR-stream grid.
constexpr int kStRB = 4;
int stencil_wgs(const trlhip_csr *A)
{
    return A->nb <= kSpmvGrid ? (A->nb + kStRB - 1) / kStRB : spmv_grid(A);
}

template <int E, bool FUSE>
void launch_vec(const trlhip_csr *A, const void *x, void *y, double *part, hipStream_t s,
                double xscale, void *xq, float *xq32)
{
    constexpr int U = E == 1 ? 4 : 2;
    const dim3 g(A->vgrid), b(kBlock);
    const double *xs = (const double *)x;
    double *ys = (double *)y, *q = (double *)xq;
#define VEC_CASE(L)                                                                            \
    case L:                                                                                    \
        KLAUNCH((k_spmv_vec<double, double, E, L, U, FUSE>), g, b, 0, s, A->n, A->rp, A->ci,   \
                A->val, xs, ys, part, xscale, q, xq32);                                        \
        break;
    switch (A->lpr) {
        VEC_CASE(2) VEC_CASE(4) VEC_CASE(8) VEC_CASE(16) VEC_CASE(32)
    }
#undef VEC_CASE
}

// y = A x on stream s with the kernel chosen for the matrix (FP64 values
// and vectors; complex when vtype is TRLHIP_C128); partials in part.  With
// xq (vector-CSR kernel only) x is an unscaled source: the product uses
// fl(x * xscale), which is also stored to xq (and rounded to xq32).
int launch_spmv(const trlhip_csr *A, int vtype, const void *x, void *y, double *part, hipStream_t s,
                double xscale = 1.0, void *xq = nullptr, float *xq32 = nullptr)
{
    const bool cplx = vtype == TRLHIP_C128;
    if (A->lpr) {
        if (cplx) {
            if (xq) launch_vec<2, true>(A, x, y, part, s, xscale, xq, nullptr);
            else launch_vec<2, false>(A, x, y, part, s, xscale, xq, nullptr);
        } else {
            if (xq) launch_vec<1, true>(A, x, y, part, s, xscale, xq, xq32);
            else launch_vec<1, false>(A, x, y, part, s, xscale, xq, nullptr);
        }
        return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
    }
    const dim3 g(spmv_grid(A)), b(kBlock);
    const double *xs = (const double *)x;
    double *ys = (double *)y, *q = (double *)xq;
    if (A->pn_on && !cplx) {
        const dim3 g1(A->pn_nslot);
        if (xq) {
            KLAUNCH_FIRST((k_spmv_panel<true>), g1, b, 0, s, A->pn_perm, A->pn_cstart, A->pn_segbase,
                          A->pn_erow, A->pn_ecol, A->pn_eval, xs, xscale, A->pn_sp);
            KLAUNCH((k_spmv_panel_rows<true>), g, b, 0, s, A->n, A->pn_rsptr, A->pn_rsidx, A->pn_sp,
                    xs, ys, part, xscale, q, xq32);
        } else {
            KLAUNCH_FIRST((k_spmv_panel<false>), g1, b, 0, s, A->pn_perm, A->pn_cstart, A->pn_segbase,
                          A->pn_erow, A->pn_ecol, A->pn_eval, xs, xscale, A->pn_sp);
            KLAUNCH((k_spmv_panel_rows<false>), g, b, 0, s, A->n, A->pn_rsptr, A->pn_rsidx,
                    A->pn_sp, xs, ys, part, xscale, (double *)nullptr, (float *)nullptr);
        }
        return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
    }
    if (A->st_dim && !cplx) {
        if (A->nb <= kSpmvGrid) {  // one alpha partial per row block
            const dim3 gw(stencil_wgs(A));
            if (A->st_dim == 2)
                KLAUNCH((k_spmv_st_rb<2, kStRB>), gw, b, 0, s, A->n, A->nb, A->st_L, A->st_off, A->dg,
                        xs, ys, part, xq ? xscale : 1.0, q, xq ? xq32 : (float *)nullptr);
            else
                KLAUNCH((k_spmv_st_rb<3, kStRB>), gw, b, 0, s, A->n, A->nb, A->st_L, A->st_off, A->dg,
                        xs, ys, part, xq ? xscale : 1.0, q, xq ? xq32 : (float *)nullptr);
            return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
        }
#define ST_CASE(D, F)                                                                             \
    KLAUNCH((k_spmv_st<D, F>), g, b, 0, s, A->n, A->nb, A->st_L, A->st_off, A->dg, xs, ys, part, \
            xscale, q, F ? xq32 : (float *)nullptr)
        if (A->st_dim == 2) {
            if (xq) ST_CASE(2, true);
            else ST_CASE(2, false);
        } else {
            if (xq) ST_CASE(3, true);
            else ST_CASE(3, false);
        }
#undef ST_CASE
        return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
    }
    if (A->pk && !cplx) {
        if (xq)
            KLAUNCH((k_spmv_pk<true>), g, b, 0, s, A->nb, A->rb, A->rp, A->pk, A->dg, A->tab, xs, ys,
                    part, xscale, q, xq32);
        else
            KLAUNCH((k_spmv_pk<false>), g, b, 0, s, A->nb, A->rb, A->rp, A->pk, A->dg, A->tab, xs, ys,
                    part, xscale, q, (float *)nullptr);
        return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
    }
    if (cplx) {
        if (xq)
            KLAUNCH((k_spmv<double, double, 2, true>), g, b, 0, s, A->nb, A->rb, A->rp, A->ci,
                    A->val, xs, ys, part, xscale, q, (float *)nullptr);
        else
            KLAUNCH((k_spmv<double, double, 2, false>), g, b, 0, s, A->nb, A->rb, A->rp, A->ci,
                    A->val, xs, ys, part, xscale, q, (float *)nullptr);
    } else {
        if (xq)
            KLAUNCH((k_spmv<double, double, 1, true>), g, b, 0, s, A->nb, A->rb, A->rp, A->ci,
                    A->val, xs, ys, part, xscale, q, xq32);
        else
            KLAUNCH((k_spmv<double, double, 1, false>), g, b, 0, s, A->nb, A->rb, A->rp, A->ci,
                    A->val, xs, ys, part, xscale, q, (float *)nullptr);
    }
    return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
}

int launch_sum(const double *part, int np, double *out, hipStream_t s)
{
    KLAUNCH(k_sum, dim3(1), dim3(kBlock), 0, s, part, np, out);
    return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
}

bool csr_fits(const trlhip_ctx *c, const trlhip_csr *A)
{
    if (!A || A->magic != kCsrMagic || A->n != c->n) return false;
    if (c->vtype == TRLHIP_C128) return A->vtype == TRLHIP_C128;
    return A->vtype == TRLHIP_R64 || A->vtype == TRLHIP_R32;
}

// FP32 shadow of basis column col (mixed-precision contexts), else null.
float *shadow32(trlhip_ctx *c, int col)
{
    return c->V32 ? c->V32 + (size_t)col * c->ld : nullptr;
}

// y col = A x col; Re(x^H y) left as partials in part_alpha (c->np_alpha).
int run_spmv(trlhip_ctx *c, const trlhip_csr *A, int xcol, int ycol)
{
    c->np_alpha = spmv_grid(A);
    Timed t(TRLHIP_K_SPMV, c->s, spmv_bytes(A, c->E, 8, 8));
    return launch_spmv(A, c->vtype, c->V + (size_t)xcol * c->ld * c->esz,
                       c->V + (size_t)ycol * c->ld * c->esz, part_alpha(c), c->s);
}

// The deferred scale-copy into column xcol fused into the SpMV that uses
// it: y = A (s * src), and s * src stored to column xcol (and to its FP32
// shadow in a mixed-precision context); src is the scratch column or, for a
// deferred column copy (s = 1, exact), another basis column.
int run_spmv_fused(trlhip_ctx *c, const trlhip_csr *A, int xcol, int ycol)
{
    c->np_alpha = spmv_grid(A);
    const double s = c->pend_s;
    const int src = c->pend_src;
    c->pend_dst = -1;
    Timed t(TRLHIP_K_SPMV, c->s,
            spmv_bytes(A, c->E, 8, 8) + (double)c->esz * c->n + (c->V32 ? 4.0 * c->n : 0.0));
    return launch_spmv(A, c->vtype, c->V + (size_t)src * c->ld * c->esz,
                       c->V + (size_t)ycol * c->ld * c->esz, part_alpha(c), c->s, s,
                       c->V + (size_t)xcol * c->ld * c->esz, shadow32(c, xcol));
}

// Re(x^H y) of two columns as partials in part_alpha.
int run_dot(trlhip_ctx *c, int xcol, int ycol)
{
    const long ne = (long)c->n * c->E;
    const int g = grid_for(ne, kBlock * 8, kMaxGrid);
    c->np_alpha = g;
    Timed t(TRLHIP_K_MISC, c->s, 2.0 * (double)c->esz * c->n);
    KLAUNCH((k_dot<double>), dim3(g), dim3(kBlock), 0, c->s, ne, colp<double>(c, xcol),
            colp<double>(c, ycol), part_alpha(c));
    return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
}

template <typename VS, int E, int RU, bool NTV = false>
void launch_update(trlhip_ctx *c, int g, const VS *V, int c0, int m, const double *dh, double h0,
                   int aslot, int rin, int rout, double *o_hh, double *pdst)
{
    const size_t shm = sizeof(double) * (size_t)c->E * (m > 0 ? m : 1);
    const double *cp = aslot >= 0 ? part_alpha(c) : nullptr;
    KLAUNCH((k_update<VS, double, E, RU, NTV>), dim3(g), dim3(kBlock), shm, c->s, c->n, V, (long)c->ld,
            c0, m, dh, h0, cp, c->np_alpha, aslot, colp<double>(c, rin), colp<double>(c, rout),
            pdst, o_hh, c->dscal + S_ALPHA);
}

// rout = rin - V(:, c0:c0+m) dh; coefficient aslot (if >= 0) is the alpha
// left in part_alpha (also stored to dscal[S_ALPHA]); ||rout||^2 -> o_rr2
// and ||dh||^2 -> o_hh when requested.  dh == nullptr: coefficients
// (h0, 0, ...) apart from aslot.  The per-block partials of ||rout||^2 go
// to pdst (default part_update) and their count to *g_out when given.
// shadow: a Gram-Schmidt update, which reads the FP32 shadow of V in a
// mixed-precision context (r itself stays FP64).
int run_update(trlhip_ctx *c, int c0, int m, const double *dh, int aslot, int rin, int rout,
               double *o_rr2, double *o_hh, int kind, double h0 = 0.0, double *pdst = nullptr,
               int *g_out = nullptr, bool shadow = false)
{
    const int g = grid_for(c->n, kMinRowsUpd, kMaxGrid);
    if (!pdst) pdst = part_update(c);
    if (g_out) *g_out = g;
    const bool s32 = shadow && c->V32;
    {
        Timed t(kind, c->s, s32 ? (4.0 * m + 16.0) * c->n : (double)c->esz * c->n * (m + 2));
        if (s32)
            launch_update<float, 1, 4>(c, g, c->V32, c0, m, dh, h0, aslot, rin, rout, o_hh, pdst);
        else if (c->E == 1 && shadow)  // the plain path's Gram-Schmidt update
            launch_update<double, 1, 4>(c, g, colp<double>(c, 0), c0, m, dh, h0, aslot, rin, rout,
                                        o_hh, pdst);
        else if (c->E == 1)  // the recurrence (and other short updates): streamed columns
            launch_update<double, 1, 4, true>(c, g, colp<double>(c, 0), c0, m, dh, h0, aslot, rin,
                                              rout, o_hh, pdst);
        else if (shadow)  // complex Gram-Schmidt update
            launch_update<double, 2, 4>(c, g, colp<double>(c, 0), c0, m, dh, h0, aslot, rin, rout,
                                        o_hh, pdst);
        else  // complex recurrence: streamed columns
            launch_update<double, 2, 4, true>(c, g, colp<double>(c, 0), c0, m, dh, h0, aslot, rin,
                                              rout, o_hh, pdst);
        if (hipGetLastError() != hipSuccess) return TRLHIP_EHIP;
    }
    return o_rr2 ? launch_sum(pdst, g, o_rr2, c->s) : 0;
}

template <typename VS, int E, int CB, int RU>
void launch_proj(trlhip_ctx *c, int gx, const VS *V, int c0, int m, int rcol)
{
    const int gy = (m + CB - 1) / CB;
    KLAUNCH((k_proj<VS, double, E, CB, RU>), dim3(gx * gy), dim3(kBlock), 0, c->s, c->n, V,
            (long)c->ld, c0, m, colp<double>(c, rcol), part_proj(c), gx, gy);
}

// Columns per projection block (a chunk) for a real basis read in FP64
// ([0]) or through the FP32 shadow ([1]): 4, 8 (default, from the sweep
// below) or 16; set by trlhip_proj_autotune / trlhip_proj_tiles.  Single
// and batched solves read the same setting, so they stay bit-identical; a
// different width sums each coefficient's lanes in another order.
// Read once per launch (a solve's group, or one projection) so that a
// change from another thread never splits one launch between two widths.
std::atomic<int> g_proj_cb[2] = {8, 8};

// Row blocks of the projection: about kProjBlocks32/64/C128 blocks (by vsz,
// the bytes of a basis element read: 4, 8 or 16) over all column chunks, at
// most kProjGrid (the partial rows), a multiple of 8 (the XCD count) when
// there are at least 8.
int proj_gx_rows(int n, int m, int cb, int vsz)
{
    const int blocks = vsz == 4 ? kProjBlocks32 : vsz == 8 ? kProjBlocks64 : kProjBlocksC128;
    const int chunks = (m + cb - 1) / cb;
    const int want = (blocks + chunks - 1) / chunks;
    const int g = grid_for(n, kMinRowsUpd, want < kProjGrid ? want : kProjGrid);
    return g >= 8 ? g & ~7 : g;
}
int proj_gx(const trlhip_ctx *c, int m, int cb, int vsz) { return proj_gx_rows(c->n, m, cb, vsz); }

// hout = V(:, c0:c0+m)^H r (device, E*m doubles).  shadow: a Gram-Schmidt
// projection, which reads the FP32 shadow of V in a mixed-precision context.
// cb_force (4, 8, 16; 0: the process-wide setting) fixes the columns per
// block, as the autotuner's trials do.
int run_proj(trlhip_ctx *c, int c0, int m, int rcol, double *hout, bool shadow = false,
             int cb_force = 0)
{
    if (m <= 0) return 0;
    const bool s32 = shadow && c->V32;
    int g;
    {
        Timed t(TRLHIP_K_PROJ, c->s, s32 ? (4.0 * m + 8.0) * c->n : (double)c->esz * c->n * (m + 1));
        // chunk widths and row unrolling from a sweep at n = 2^21, m = 16..64
        // (about 5.7 TB/s for both element types)
        if (s32 || c->E == 1) {
            const int cb = cb_force ? cb_force : g_proj_cb[s32 ? 1 : 0].load(std::memory_order_relaxed);
            g = proj_gx(c, m, cb, s32 ? 4 : 8);
#define PROJ_CB(CB)                                                                      \
    case CB:                                                                             \
        if (s32)                                                                         \
            launch_proj<float, 1, CB, 4>(c, g, c->V32, c0, m, rcol);                     \
        else                                                                             \
            launch_proj<double, 1, CB, 4>(c, g, colp<double>(c, 0), c0, m, rcol);        \
        break;
            switch (cb) {
                PROJ_CB(4)
                PROJ_CB(16)
                default: PROJ_CB(8)
            }
#undef PROJ_CB
        } else {
            g = proj_gx(c, m, 4, 16);
            launch_proj<double, 2, 4, 4>(c, g, colp<double>(c, 0), c0, m, rcol);
        }
        if (hipGetLastError() != hipSuccess) return TRLHIP_EHIP;
    }
    const int np = m * c->E;
    KLAUNCH(k_colsum, dim3(np), dim3(kBlock), 0, c->s, part_proj(c), g, hout);
    return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
}

// dst = s * src (columns; src may be the scratch column); the FP32 shadow
// of dst follows in a mixed-precision context.
int run_scale_copy(trlhip_ctx *c, int dst, int src, double s)
{
    if (c->batch && c->E == 1) {  // recorded; the batch flush runs it
        BatchOp op{};
        op.c = c;
        op.type = OP_COPY;
        op.dst = dst;
        op.src = src;
        op.s = s;
        std::lock_guard<SpinLock> lk(c->batch->mu);
        op.phase = c->nrec++;
        c->batch->ops.push_back(op);
        return 0;
    }
    const long ne = (long)c->n * c->E;
    const int g = grid_for(ne, kBlock * 8, kMaxGrid);
    Timed t(TRLHIP_K_MISC, c->s, 2.0 * (double)c->esz * c->n);
    KLAUNCH((k_scale_copy<double>), dim3(g), dim3(kBlock), 0, c->s, ne, colp<double>(c, src),
            colp<double>(c, dst), s, shadow32(c, dst));
    return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
}

// Apply a deferred scale-copy (trlhip_col_scale_copy from the scratch
// column records it; the next step's SpMV usually absorbs it).
int flush_pending(trlhip_ctx *c)
{
    if (c->pend_dst < 0) return 0;
    const int dst = c->pend_dst;
    c->pend_dst = -1;
    return run_scale_copy(c, dst, c->pend_src, c->pend_s);
}

// Wait until the context's stream is idle: synchronize it, or, in a batch,
// yield to the batch scheduler, which synchronizes the shared stream once
// for all its coroutines (trlhip_batch_flush) before resuming this one.
int ctx_wait(trlhip_ctx *c)
{
    if (c->batch) {
        c->batch->yield(c->batch->user);
        return c->batch->err;
    }
    TRY(hipStreamSynchronize(c->s));
    stats_resolve();
    return 0;
}

// In a batch, the restart GEMMs and scale-copies a context records run at
// the next flush.  An operation that launches or copies at once applies a
// deferred scale-copy and, when the context has recorded operations,
// yields to that flush first, so the context's device work keeps its
// program order.
int batch_drain(trlhip_ctx *c) { return (c->batch && c->nrec) ? ctx_wait(c) : 0; }

// A chain follower the device ran that the solver has not taken yet: the
// solver's own decision for the step before it must be the common one the
// device made (chain_go), so any other device operation on the context
// now is a mismatch, reported as an error rather than run out of order.
bool chain_live(const trlhip_ctx *c)
{
    return c->ch_next > 0 && c->hres[hres_chain(c->ncols) + 7 * (size_t)(c->ch_next - 1) + 4] == 1.0;
}

int settle(trlhip_ctx *c)
{
    if (chain_live(c)) return TRLHIP_EINVAL;
    if (int rc = flush_pending(c)) return rc;
    return batch_drain(c);
}

// copy scalars back and wait
int fetch(trlhip_ctx *c, const double *dsrc, double *hdst, size_t cnt)
{
    TRY(hipMemcpyAsync(hdst, dsrc, cnt * sizeof(double), hipMemcpyDeviceToHost, c->s));
    return ctx_wait(c);
}

bool valid_col(const trlhip_ctx *c, int col) { return col >= 0 && col < c->ncols; }

bool valid_cols(const trlhip_ctx *c, int c0, int m)
{
    return c0 >= 0 && m >= 0 && c0 + m <= c->ncols;
}

// Row blocks for the CSR-stream SpMV: at most kRowsBlk rows and kNnzBlk
// nonzeros each; a longer row gets a block of its own.
std::vector<int> row_blocks(int n, const int *rp)
{
    std::vector<int> rb;
    rb.reserve((size_t)n / 16 + 2);
    int r = 0;
    rb.push_back(0);
    while (r < n) {
        int e = r + 1;
        while (e < n && e - r < kRowsBlk && rp[e + 1] - rp[r] <= kNnzBlk) ++e;
        rb.push_back(e);
        r = e;
    }
    return rb;
}

// ------------------------------------------------------------------ batch
//
// A batch runs many independent solves (one context each, all on the
// batch's stream).  Their host drivers are coroutines: where a context
// must wait for the device it yields instead of synchronizing, and a
// Lanczos step (trlhip_step_cgs) is only recorded.  The scheduler then
// calls trlhip_batch_flush, which runs every recorded step with one
// launch per phase for all problems (block -> problem through a prefix
// table), synchronizes once, and resumes the coroutines.

// Parameters of one recorded step (device copy, one per request).
struct BDesc {
    int n, ld, jnd, np_alpha, g_upd, gx, gy, rm, rc0, nb;
    const int *rb, *rp, *ci;
    const double *val;
    const uint32_t *pk;   // packed nonzeros (kb_spmv_pk), with dg and tab
    const double *dg, *tab;
    int st_L;             // stencil lattice side (kb_spmv_st), with st_off and dg
    double st_off;
    const double *x;      // SpMV input: the basis column q_a, or the unscaled scratch column
    double xscale;        // 1, or the deferred normalisation of the scratch column
    double *xq;           // fused store of q_a (null when not fused)
    float *xq32;          // its FP32 shadow (mixed precision)
    double *r;            // residual column
    double *scratch;      // speculative Gram-Schmidt output
    const double *V;      // FP64 basis
    const void *VG;       // basis the Gram-Schmidt pass reads (FP32 shadow or FP64)
    const double *rh;     // recurrence coefficients (first step after a restart) or null
    double rh0;           // beta_prev
    double *part_alpha, *part_recur, *part_update, *part_proj;
    double *dh, *dscal, *dres;
    // a chain follower reads its go flag, normalisation and beta_prev from
    // the device scalars the step before it left (null: xscale / rh0 hold
    // them and the step runs)
    const double *go, *dxscale, *drh0;
    double *hm;           // host-coherent mirror of this step's 4 results
    double *hm_next;      // decide: the follower's mirror slot (go, 1/beta, beta at 4..6)
    int decide, kept, ntot;
    double eps_cgs;       // the Gram-Schmidt pass test's epsilon (FP32 basis: FLT_EPSILON)
};

enum { BP_SPMV, BP_RECUR, BP_PROJ, BP_COLSUM, BP_UPDATE, BP_SUM2, BP_COUNT };

// Request p owning global block blk: the last p with pre[p] <= blk
// (pre[0] = 0, strictly increasing).  Each wave searches cooperatively: its
// 64 lanes test 64 evenly spaced candidates per step and a ballot narrows
// the range 64-fold, so 512 requests take 2 dependent loads instead of 9.
__device__ __forceinline__ int batch_find(const int *__restrict__ pre, int nreq, int blk)
{
    const int lane = threadIdx.x & 63;
    int lo = 0, hi = nreq - 1;  // pre[lo] <= blk, answer <= hi
    while (lo < hi) {
        const int step = (hi - lo + 64) / 64;  // ceil((hi - lo + 1) / 64)
        const int k = lo + lane * step;
        const bool le = k <= hi && pre[k] <= blk;
        const unsigned long long m = __ballot(le);  // lane 0 is always set
        const int last = 63 - __clzll(m);
        lo += last * step;
        hi = min(hi, lo + step - 1);
        if (step == 1) break;
    }
    return lo;
}

#define BATCH_LOCATE(phase)                                                   \
    const int *pre_ = pre + (size_t)(phase) * (nreq + 1);                     \
    const int p_ = batch_find(pre_, nreq, blockIdx.x);                        \
    const int bid_ = blockIdx.x - pre_[p_], nblk_ = pre_[p_ + 1] - pre_[p_];  \
    const BDesc &q = d[p_];                                                   \
    if (q.go && *q.go == 0.0) return; /* a chain follower that does not run */

// The batched SpMVs' locate: a chain follower's go flag and fused scale
// 1/beta are adjacent device scalars (S_GO, S_XS), read in one load.
static_assert(S_XS == S_GO + 1 && S_GO % 2 == 0, "go flag and scale adjacent, 16-byte aligned");
#define BATCH_LOCATE_SPMV                                                     \
    const int *pre_ = pre + (size_t)BP_SPMV * (nreq + 1);                     \
    const int p_ = batch_find(pre_, nreq, blockIdx.x);                        \
    const int bid_ = blockIdx.x - pre_[p_], nblk_ = pre_[p_ + 1] - pre_[p_];  \
    const BDesc &q = d[p_];                                                   \
    double xs_ = q.xscale;                                                    \
    if (q.go) {                                                               \
        const double2 gs_ = *reinterpret_cast<const double2 *>(q.go);         \
        if (gs_.x == 0.0) return; /* a chain follower that does not run */    \
        xs_ = gs_.y;                                                          \
    }

__global__ __launch_bounds__(kBlock) void kb_spmv_stream(const BDesc *__restrict__ d, const int *pre,
                                                        int nreq)
{
    BATCH_LOCATE_SPMV
    spmv_stream_body<double, double, 1, true>(q.nb, q.rb, q.rp, q.ci, q.val, q.x, q.r, q.part_alpha,
                                              xs_, q.xq, q.xq32, bid_, nblk_);
}

__global__ __launch_bounds__(kBlock) void kb_spmv_pk(const BDesc *__restrict__ d, const int *pre,
                                                    int nreq)
{
    BATCH_LOCATE_SPMV
    spmv_stream_body<double, double, 1, true, true>(q.nb, q.rb, q.rp, nullptr, nullptr, q.x, q.r,
                                                    q.part_alpha, xs_, q.xq, q.xq32, bid_, nblk_,
                                                    q.pk, q.dg, q.tab);
}

template <int D>
__global__ __launch_bounds__(kBlock) void kb_spmv_st(const BDesc *__restrict__ d, const int *pre,
                                                    int nreq)
{
    BATCH_LOCATE_SPMV
    if (q.nb <= kSpmvGrid)
        spmv_stencil_rb<D, kStRB>(q.n, q.nb, q.st_L, q.st_off, q.dg, q.x, q.r, q.part_alpha, xs_,
                                  q.xq, q.xq32, bid_);
    else
        spmv_stencil_body<D, true>(q.n, q.nb, q.st_L, q.st_off, q.dg, q.x, q.r, q.part_alpha, xs_,
                                   q.xq, q.xq32, bid_, nblk_);
}

template <int LPR>
__global__ __launch_bounds__(kBlock) void kb_spmv(const BDesc *__restrict__ d, const int *pre, int nreq)
{
    BATCH_LOCATE_SPMV
    spmv_vec_body<double, double, 1, LPR, 4, true>(q.n, q.rp, q.ci, q.val, q.x, q.r, q.part_alpha,
                                                   xs_, q.xq, q.xq32, bid_, nblk_);
}

__global__ __launch_bounds__(kBlock) void kb_recur(const BDesc *__restrict__ d, const int *pre, int nreq)
{
    BATCH_LOCATE(BP_RECUR)
    update_body<double, double, 1, 4, true>(q.n, q.V, q.ld, q.rc0, q.rm, q.rh, q.drh0 ? *q.drh0 : q.rh0,
                                      q.part_alpha, q.np_alpha, q.rm - 1, q.r, q.r, q.part_recur,
                                      nullptr, q.dscal + S_ALPHA, bid_, nblk_);
}

template <typename VS, int CB>
__global__ __launch_bounds__(kBlock) void kb_proj(const BDesc *__restrict__ d, const int *pre, int nreq)
{
    BATCH_LOCATE(BP_PROJ)
    proj_body<VS, double, 1, CB, 4>(q.n, (const VS *)q.VG, q.ld, 0, q.jnd, q.r, q.part_proj, q.gx,
                                    q.gy, bid_, nblk_);
}

template <int CB, typename VS>
__global__ __launch_bounds__(kBlock) void kb_recproj(const BDesc *__restrict__ d, const int *pre,
                                                    int nreq)
{
    const int *pre_ = pre + (size_t)BP_PROJ * (nreq + 1);
    const int p_ = batch_find(pre_, nreq, blockIdx.x);
    const int bid_ = blockIdx.x - pre_[p_];
    const BDesc &q = d[p_];
    double h0 = q.rh0;
    if (q.go) {  // a chain follower: go flag (S_GO) and beta_prev (S_BETA) loaded together
        static_assert(S_BETA == S_GO + 2, "go flag and beta in one 32-byte block");
        const double2 a = reinterpret_cast<const double2 *>(q.go)[0];
        const double2 b = reinterpret_cast<const double2 *>(q.go)[1];
        if (a.x == 0.0) return;  // a follower that does not run
        h0 = b.x;
    }
    recproj_body<CB, VS>(q.n, q.V, q.ld, q.rc0, q.rm, q.rh, h0, q.part_alpha,
                         q.np_alpha, q.r, q.part_recur, q.dscal + S_ALPHA, (const VS *)q.VG, q.jnd,
                         q.part_proj, q.gx, bid_);
}

// Four coefficients per block, one per wave: wave w sums coefficient
// 4 bid + w's gx partials exactly as ordered_sum does with a block (the
// four 64-lane wave totals of its strided sums, each a butterfly
// wave_sum, added to 0.0 in wave order), so dh equals k_colsum's bit for
// bit with a quarter of the blocks.
__global__ __launch_bounds__(kBlock) void kb_colsum(const BDesc *__restrict__ d, const int *pre, int nreq)
{
    BATCH_LOCATE(BP_COLSUM)
    (void)nblk_;
    const int lane = threadIdx.x & 63, c = 4 * bid_ + (threadIdx.x >> 6);
    if (c >= q.jnd) return;
    const double *pc = q.part_proj + (size_t)c * q.gx;
    double t = 0.0;
#pragma unroll
    for (int v = 0; v < kBlock / 64; ++v) {
        // a virtual wave past the partials holds zeros: its total is +0.0
        const double wv = v * 64 < q.gx ? wave_sum(strided_sum(pc, q.gx, 1, 0, v * 64 + lane, kBlock)) : 0.0;
        t += wv;
    }
    if (lane == 0) q.dh[c] = t;
}

template <typename VS>
__global__ __launch_bounds__(kBlock) void kb_update(const BDesc *__restrict__ d, const int *pre, int nreq)
{
    BATCH_LOCATE(BP_UPDATE)
    update_body<VS, double, 1, 4>(q.n, (const VS *)q.VG, q.ld, 0, q.jnd, q.dh, 0.0, nullptr, 0, -1,
                                  q.r, q.scratch, q.part_update, q.dscal + S_HH, nullptr, bid_,
                                  nblk_);
}

// Whether a chain follower runs: the solver's own decisions after a step
// (SRC/trlcore.c:1872-1898 and trl_cgs's first pass, 2913-3117, as
// lz_orth / lz_cgs make them), in the same FP64 operations, for the
// common case only -- the step is accepted after its one speculative
// Gram-Schmidt pass with beta = sqrt(||r - V h||^2) > eps |alpha|.  Any
// other outcome stops the chain and the solver takes over.
__device__ __forceinline__ bool chain_go(double alpha, double rr2, double hh, double rr2b,
                                         double bprev, int jnd, int kept, int ntot, double eps_cgs)
{
#pragma clang fp contract(off)
    if (!(rr2 >= 0.0) || !(rr2 <= DBL_MAX)) return false;
    const int jm1 = jnd - 1;
    if (jm1 <= kept) return false;  // a cycle's first step (not chained)
    if (jnd >= ntot) return false;
    double tmp = alpha * alpha;
    tmp += bprev * bprev;
    if (DBL_EPSILON * rr2 >= tmp) return false;  // local re-orthogonalisation
    if (!(eps_cgs * rr2b > hh)) return false;    // another Gram-Schmidt pass
    const double rnrm = sqrt(rr2b);
    if (!(rnrm > DBL_MIN)) return false;
    if (rnrm <= DBL_EPSILON * fabs(alpha)) return false;
    return true;
}

// Both norms of a step (before and after its speculative Gram-Schmidt
// pass), in one block per request: the same fixed-order sums k_sum2 forms,
// then the step's four results to its host-coherent mirror, and, when a
// follower comes next, the decision whether it runs.
__global__ __launch_bounds__(kBlock) void kb_sum2(const BDesc *__restrict__ d, const int *pre, int nreq)
{
    BATCH_LOCATE(BP_SUM2)
    (void)bid_;
    (void)nblk_;
    // wave 0 forms ||r||^2 and wave 1 ||r - V h||^2 from their partials
    // exactly as ordered_sum does with a block (k_sum2's bits)
    __shared__ double t1s;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double t = 0.0;
    if (w < 2) {
        const double *p = w == 0 ? q.part_recur : q.part_update;
#pragma unroll
        for (int v = 0; v < kBlock / 64; ++v) {
            const double wv = v * 64 < q.g_upd ? wave_sum(strided_sum(p, q.g_upd, 1, 0, v * 64 + lane, kBlock)) : 0.0;
            t += wv;
        }
        if (w == 1 && lane == 0) t1s = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double t0 = t, t1 = t1s;
        double *sc = q.dscal;
        sc[S_RR2] = t0;
        sc[S_RR2B] = t1;
        q.hm[S_ALPHA] = sc[S_ALPHA];
        q.hm[S_RR2] = t0;
        q.hm[S_HH] = sc[S_HH];
        q.hm[S_RR2B] = t1;
        if (q.decide) {
            const double bprev = q.drh0 ? *q.drh0 : q.rh0;
            const bool go = chain_go(sc[S_ALPHA], t0, sc[S_HH], t1, bprev, q.jnd, q.kept, q.ntot,
                                     q.eps_cgs);
            const double beta = go ? sqrt(t1) : 0.0;
            const double xs = go ? 1.0 / beta : 0.0;
            sc[S_GO] = go ? 1.0 : 0.0;
            sc[S_XS] = xs;
            sc[S_BETA] = beta;
            q.hm_next[4] = go ? 1.0 : 0.0;
            q.hm_next[5] = xs;
            q.hm_next[6] = beta;
        }
    }
}
#undef BATCH_LOCATE
#undef BATCH_LOCATE_SPMV

// Host staging of one group of recorded steps (same lanes per row, same
// basis precision): descriptors d[0:R) and BP_COUNT prefix tables of R + 1
// block offsets each; returns the phases' algorithmic bytes and the largest
// recurrence / projection widths.
struct GroupInfo {
    int R, lpr;  // lpr: the SpMV kernel (spmv_kind: lanes per row, 0: CSR-stream, -1: packed
                 // CSR-stream, -2 / -3: 2D / 3D stencil)
    bool mixed;
    bool fused;         // recurrence and first projection in one kernel (recproj_ok)
    int cb;             // projection columns per block, read once for the group
    size_t doff, poff;  // offsets into the staging arrays
    double bytes[BP_COUNT];
    int maxrm, maxj;
};

void batch_fill(const BatchStep *steps, int R, BDesc *desc, int *pre, GroupInfo &gi,
                const double *dcbuf)
{
    for (int ph = 0; ph < BP_COUNT; ++ph) pre[(size_t)ph * (R + 1)] = 0;
    for (double &x : gi.bytes) x = 0.0;
    gi.maxrm = gi.maxj = 1;
    for (int i = 0; i < R; ++i) {
        const BatchStep &st = steps[i];
        trlhip_ctx *c = st.c;
        const trlhip_csr *A = st.A;
        BDesc &d = desc[i];
        const int n = c->n, jnd = st.jnd;
        d.n = n;
        d.ld = c->ld;
        d.jnd = jnd;
        d.np_alpha = spmv_kind_batch(A) == spmv_kind(A) ? spmv_grid(A) : spmv_grid_stream(A);
        c->np_alpha = d.np_alpha;
        d.g_upd = grid_for(n, kMinRowsUpd, kMaxGrid);
        const int cb = gi.cb;
        d.gx = gi.fused ? d.g_upd : proj_gx(c, jnd, cb, gi.mixed ? 4 : 8);  // fused: one block per row block
        d.gy = gi.fused ? 1 : (jnd + cb - 1) / cb;
        if (st.first) {
            d.rm = jnd;
            d.rc0 = 0;
            d.rh = st.coff >= 0 ? dcbuf + st.coff : c->dcoef;
            d.rh0 = 0.0;
        } else if (st.qb >= 0) {
            d.rm = 2;
            d.rc0 = st.qb;
            d.rh = nullptr;
            d.rh0 = st.beta_prev;
        } else {
            d.rm = 1;
            d.rc0 = st.qa;
            d.rh = nullptr;
            d.rh0 = 0.0;
        }
        d.nb = A->nb;
        d.rb = A->rb;
        d.rp = A->rp;
        d.ci = A->ci;
        d.val = A->val;
        d.pk = A->pk;
        d.dg = A->dg;
        d.tab = A->tab;
        d.st_L = A->st_L;
        d.st_off = A->st_off;
        d.x = colp<double>(c, st.fuse ? st.fuse_src : st.qa);
        d.xscale = st.fuse ? st.fuse_s : 1.0;
        // a chain follower: its fused normalisation and beta_prev come from
        // the step before it, on the device, and it runs only if told to
        d.go = st.ci > 0 ? c->dscal + S_GO : nullptr;
        d.dxscale = st.ci > 0 ? c->dscal + S_XS : nullptr;
        d.drh0 = st.ci > 0 ? c->dscal + S_BETA : nullptr;
        d.hm = st.ci > 0 ? c->dres + hres_chain(c->ncols) + 7 * (size_t)(st.ci - 1) : c->dres;
        d.hm_next = c->dres + hres_chain(c->ncols) + 7 * (size_t)st.ci;
        d.decide = st.decide;
        d.kept = st.kept;
        d.ntot = st.ntot;
        d.eps_cgs = gi.mixed ? (double)FLT_EPSILON : DBL_EPSILON;
        d.xq = st.fuse ? colp<double>(c, st.qa) : nullptr;
        d.xq32 = st.fuse ? shadow32(c, st.qa) : nullptr;
        d.r = colp<double>(c, st.rcol);
        d.scratch = colp<double>(c, c->ncols);
        d.V = colp<double>(c, 0);
        d.VG = gi.mixed ? (const void *)c->V32 : (const void *)colp<double>(c, 0);
        d.part_alpha = part_alpha(c);
        d.part_recur = part_recur(c);
        d.part_update = part_update(c);
        d.part_proj = part_proj(c);
        d.dh = c->dh;
        d.dscal = c->dscal;
        d.dres = c->dres;
        const int kind = spmv_kind_batch(A);
        const int spmv_wg = (kind == -2 || kind == -3) ? stencil_wgs(A) : d.np_alpha;
        const int nb[BP_COUNT] = {spmv_wg, d.g_upd, d.gx * d.gy, (jnd + 3) / 4, d.g_upd, 1};
        for (int ph = 0; ph < BP_COUNT; ++ph) {
            int *pp = pre + (size_t)ph * (R + 1);
            pp[i + 1] = pp[i] + nb[ph];
        }
        const double vs = gi.mixed ? 4.0 : 8.0;
        const double sb = spmv_kind_batch(A) == spmv_kind(A)
                              ? spmv_bytes(A, 1, 8, 8)
                              : 12.0 * (double)A->nnz + 4.0 * (n + 1) + 16.0 * n;
        gi.bytes[BP_SPMV] += sb + (st.fuse ? (gi.mixed ? 12.0 : 8.0) * n : 0.0);
        if (gi.fused) {  // y, the recurrence columns, r written once; the shadow basis
            gi.bytes[BP_PROJ] += 8.0 * n * (d.rm + 2) + vs * jnd * n;
        } else {
            gi.bytes[BP_RECUR] += 8.0 * n * (d.rm + 2);
            gi.bytes[BP_PROJ] += (vs * jnd + 8.0) * n;
        }
        gi.bytes[BP_UPDATE] += (vs * jnd + 16.0) * n;
        gi.maxrm = std::max(gi.maxrm, d.rm);
        gi.maxj = std::max(gi.maxj, jnd);
    }
}

// The launches of one staged group: one kernel per phase over its R
// requests (descriptors dd, prefix tables dp on the device).
int batch_run(hipStream_t s, const GroupInfo &gi, const BDesc *dd, const int *dp, const int *hpre)
{
    const int R = gi.R;
    auto blocks = [&](int ph) { return dim3(hpre[(size_t)ph * (R + 1) + R]); };
    {
        Timed t(TRLHIP_K_BSPMV, s, gi.bytes[BP_SPMV]);
        switch (gi.lpr) {
        case -1: KLAUNCH(kb_spmv_pk, blocks(BP_SPMV), dim3(kBlock), 0, s, dd, dp, R); break;
        case -2: KLAUNCH((kb_spmv_st<2>), blocks(BP_SPMV), dim3(kBlock), 0, s, dd, dp, R); break;
        case -3: KLAUNCH((kb_spmv_st<3>), blocks(BP_SPMV), dim3(kBlock), 0, s, dd, dp, R); break;
        case 0: KLAUNCH(kb_spmv_stream, blocks(BP_SPMV), dim3(kBlock), 0, s, dd, dp, R); break;
        case 2: KLAUNCH((kb_spmv<2>), blocks(BP_SPMV), dim3(kBlock), 0, s, dd, dp, R); break;
        case 4: KLAUNCH((kb_spmv<4>), blocks(BP_SPMV), dim3(kBlock), 0, s, dd, dp, R); break;
        case 8: KLAUNCH((kb_spmv<8>), blocks(BP_SPMV), dim3(kBlock), 0, s, dd, dp, R); break;
        case 16: KLAUNCH((kb_spmv<16>), blocks(BP_SPMV), dim3(kBlock), 0, s, dd, dp, R); break;
        default: KLAUNCH((kb_spmv<32>), blocks(BP_SPMV), dim3(kBlock), 0, s, dd, dp, R); break;
        }
    }
    if (!gi.fused) {
        Timed t(TRLHIP_K_BRECUR, s, gi.bytes[BP_RECUR]);
        KLAUNCH(kb_recur, blocks(BP_RECUR), dim3(kBlock), sizeof(double) * gi.maxrm, s, dd, dp, R);
    }
    {
        Timed t(TRLHIP_K_BPROJ, s, gi.bytes[BP_PROJ]);
#define KB_PROJ_CB(CB)                                                                        \
    case CB:                                                                                  \
        if (gi.fused && gi.mixed)                                                             \
            KLAUNCH((kb_recproj<CB, float>), blocks(BP_PROJ), dim3(kBlock),                    \
                    sizeof(double) * gi.maxrm, s, dd, dp, R);                                 \
        else if (gi.fused)                                                                    \
            KLAUNCH((kb_recproj<CB, double>), blocks(BP_PROJ), dim3(kBlock),                   \
                    sizeof(double) * gi.maxrm, s, dd, dp, R);                                 \
        else if (gi.mixed)                                                                    \
            KLAUNCH((kb_proj<float, CB>), blocks(BP_PROJ), dim3(kBlock), 0, s, dd, dp, R);    \
        else                                                                                  \
            KLAUNCH((kb_proj<double, CB>), blocks(BP_PROJ), dim3(kBlock), 0, s, dd, dp, R);   \
        break;
        switch (gi.cb) {
            KB_PROJ_CB(4)
            KB_PROJ_CB(16)
            default: KB_PROJ_CB(8)
        }
#undef KB_PROJ_CB
    }
    KLAUNCH(kb_colsum, blocks(BP_COLSUM), dim3(kBlock), 0, s, dd, dp, R);
    {
        Timed t(TRLHIP_K_BUPDATE, s, gi.bytes[BP_UPDATE]);
        if (gi.mixed)
            KLAUNCH((kb_update<float>), blocks(BP_UPDATE), dim3(kBlock), sizeof(double) * gi.maxj, s,
                    dd, dp, R);
        else
            KLAUNCH((kb_update<double>), blocks(BP_UPDATE), dim3(kBlock), sizeof(double) * gi.maxj, s,
                    dd, dp, R);
    }
    KLAUNCH(kb_sum2, blocks(BP_SUM2), dim3(kBlock), 0, s, dd, dp, R);
    return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
}

// (Grouping the launches group-major instead, so one group's vectors stay
// in the last-level cache from phase to phase, measured 7-33 % slower on
// the batched 2D Anderson workload: the smaller launches lose more than
// the cache hits gain.)
// Staging plan of a batch's recorded steps: the steps grouped (stable) by
// (SpMV kernel: lanes per row, 0 for CSR-stream; basis precision), and one
// staging block of `total` bytes: descriptors | prefix tables of all
// groups | coefficients.
struct StepsPlan {
    std::vector<BatchStep> order;
    std::vector<GroupInfo> groups;
    size_t npre = 0, off_pre = 0, off_cb = 0, total = 0;
};

void steps_plan(const trlhip_batch *b, StepsPlan &P)
{
    auto &order = P.order;
    auto &groups = P.groups;
    order.reserve(b->steps.size());
    std::vector<char> taken(b->steps.size(), 0);
    size_t npre = 0;
    // chain position first: a context's follower runs after its step
    // before it, so all steps at position 0 launch first, then position 1 ...
    int maxci = 0;
    for (const BatchStep &st : b->steps) maxci = std::max(maxci, st.ci);
    for (int ci = 0; ci <= maxci; ++ci)
        for (size_t i = 0; i < b->steps.size(); ++i) {
            if (taken[i] || b->steps[i].ci != ci) continue;
            GroupInfo gi{};
            gi.lpr = spmv_kind_batch(b->steps[i].A);
            gi.mixed = b->steps[i].c->V32 != nullptr;
            gi.fused = recproj_ok(b->steps[i].c);
            gi.cb = g_proj_cb[gi.mixed ? 1 : 0].load(std::memory_order_relaxed);
            gi.doff = order.size();
            gi.poff = npre;
            for (size_t k = i; k < b->steps.size(); ++k) {
                const BatchStep &st = b->steps[k];
                if (!taken[k] && st.ci == ci && spmv_kind_batch(st.A) == gi.lpr &&
                    (st.c->V32 != nullptr) == gi.mixed && recproj_ok(st.c) == gi.fused) {
                    order.push_back(st);
                    taken[k] = 1;
                }
            }
            gi.R = (int)(order.size() - gi.doff);
            npre += (size_t)BP_COUNT * (gi.R + 1);
            groups.push_back(gi);
        }
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    P.npre = npre;
    P.off_pre = up(sizeof(BDesc) * order.size());
    P.off_cb = up(P.off_pre + sizeof(int) * npre);
    P.total = P.off_cb + sizeof(double) * b->cbuf.size();
}

// Writes the plan's descriptors, prefix tables and coefficients into the
// staging block (h: host, d: the device address it is copied to).
void steps_fill(const trlhip_batch *b, StepsPlan &P, char *h, const char *d)
{
    BDesc *hdesc = reinterpret_cast<BDesc *>(h);
    int *hpre = reinterpret_cast<int *>(h + P.off_pre);
    const double *dcbuf = reinterpret_cast<const double *>(d + P.off_cb);
    for (GroupInfo &gi : P.groups)
        batch_fill(P.order.data() + gi.doff, gi.R, hdesc + gi.doff, hpre + gi.poff, gi, dcbuf);
    if (!b->cbuf.empty()) memcpy(h + P.off_cb, b->cbuf.data(), sizeof(double) * b->cbuf.size());
}

// Launches every recorded step, in the order form_residual + run_proj +
// run_update + k_sum2 run them for a single context: each group's kernel
// set, from a staging block already copied to d.  Complex steps never get
// here (trlhip_step_cgs runs them directly).
int steps_launch(hipStream_t s, const StepsPlan &P, const char *h, const char *d)
{
    const BDesc *ddesc = reinterpret_cast<const BDesc *>(d);
    const int *dpre = reinterpret_cast<const int *>(d + P.off_pre);
    const int *hpre = reinterpret_cast<const int *>(h + P.off_pre);
    for (const GroupInfo &gi : P.groups)
        if (int rc = batch_run(s, gi, ddesc + gi.doff, dpre + gi.poff, hpre + gi.poff)) return rc;
    return 0;
}

// Device record of one recorded operation: a restart GEMM (V, ld, lck,
// jl, k, Y, ldy over n rows, FP32 shadow slab V32) or a scale-copy
// (V = dst column, src, s over n elements, V32 = the dst shadow column).
struct ODesc {
    long n, ld;
    double *V;
    const double *src;
    const float *src32;  // copy: FP32 source instead of src
    float *V32;
    const double *Y;
    double s;
    int lck, jl, k, ldy;
    double *part;        // norm: its g block partials, the sum to o1 and o2
    double *o1, *o2;
    int g;
};

template <int KT>
__global__ __launch_bounds__(kBlock) void kb_ritz(const ODesc *__restrict__ d, const int *pre, int nreq)
{
    extern __shared__ double Ys[];
    const int p = batch_find(pre, nreq, blockIdx.x);
    const ODesc &q = d[p];
    ritz_body<KT>(q.n, q.V, q.ld, q.lck, q.jl, q.k, q.Y, q.ldy, kRitzTiles, q.V32, Ys,
                  blockIdx.x - pre[p]);
}

__global__ __launch_bounds__(kBlock) void kb_copy(const ODesc *__restrict__ d, const int *pre, int nreq)
{
    const int p = batch_find(pre, nreq, blockIdx.x);
    const ODesc &q = d[p];
    const int bid = blockIdx.x - pre[p], nblk = pre[p + 1] - pre[p];
    for (long i = (long)bid * kBlock + threadIdx.x; i < q.n; i += (long)nblk * kBlock) {
        const double v = (q.src32 ? (double)q.src32[i] : q.src[i]) * q.s;
        q.V[i] = v;
        if (q.V32) q.V32[i] = (float)v;
    }
}

// ||V||^2 of one column per request, in place: exactly the single
// context's trlhip_update with no coefficients (k_update over the same g
// row blocks, then k_sum), so batched and single solves agree bit for bit.
__global__ __launch_bounds__(kBlock) void kb_norm(const ODesc *__restrict__ d, const int *pre, int nreq)
{
    const int p = batch_find(pre, nreq, blockIdx.x);
    const ODesc &q = d[p];
    update_body<double, double, 1, 4>((int)q.n, q.V, q.ld, 0, 0, nullptr, 0.0, nullptr, 0, -1, q.V,
                                      q.V, q.part, nullptr, nullptr, blockIdx.x - pre[p],
                                      pre[p + 1] - pre[p]);
}

// One block per request: the norm's fixed-order sum (k_sum's).
__global__ __launch_bounds__(kBlock) void kb_norm_fin(const ODesc *__restrict__ d)
{
    const ODesc &q = d[blockIdx.x];
    const double t = ordered_sum(q.part, q.g);
    if (threadIdx.x == 0) {
        q.o1[0] = t;
        q.o2[0] = t;
    }
}

// Rayleigh-quotient refinement of one Ritz vector v (a basis column) with
// the FP64 operator: t = A v into a free column, rho = v'Av / v'v,
// ||A v - rho v|| / ||v||, all on the device.  Partials at part:
// [0, kSpmvGrid) v'Av, then kMaxGrid each for v'v and ||Av - rho v||^2;
// coef = {rho, v'v}; out = {rho, residual, v'v} in host-coherent memory.
struct FDesc {
    int n, nb, npx, nd;
    const int *rb, *rp, *ci;
    const double *val;
    const double *v;
    double *t, *part, *coef, *out;
};
constexpr int kRefinePart = kSpmvGrid + 2 * kMaxGrid;

__global__ __launch_bounds__(kBlock) void kr_spmv(const FDesc *__restrict__ d, const int *pre, int nreq)
{
    const int p = batch_find(pre, nreq, blockIdx.x);
    const FDesc &q = d[p];
    spmv_stream_body<double, double, 1, false>(q.nb, q.rb, q.rp, q.ci, q.val, q.v, q.t, q.part, 1.0,
                                               nullptr, nullptr, blockIdx.x - pre[p],
                                               pre[p + 1] - pre[p]);
}

__global__ __launch_bounds__(kBlock) void kr_dot(const FDesc *__restrict__ d, const int *pre, int nreq)
{
    const int p = batch_find(pre, nreq, blockIdx.x);
    const FDesc &q = d[p];
    const int bid = blockIdx.x - pre[p], nblk = pre[p + 1] - pre[p];
    double acc = 0.0;
    for (long i = (long)bid * kBlock + threadIdx.x; i < q.n; i += (long)nblk * kBlock) {
        const double x = q.v[i];
        acc += x * x;
    }
    const double sb = block_sum(acc);
    if (threadIdx.x == 0) q.part[kSpmvGrid + bid] = sb;
}

__global__ __launch_bounds__(kBlock) void kr_rho(const FDesc *__restrict__ d)
{
    const FDesc &q = d[blockIdx.x];
    const double vav = ordered_sum(q.part, q.npx);
    const double vv = ordered_sum(q.part + kSpmvGrid, q.nd);
    if (threadIdx.x == 0) {
        q.coef[0] = vv > 0.0 ? vav / vv : 0.0;
        q.coef[1] = vv;
    }
}

__global__ __launch_bounds__(kBlock) void kr_res(const FDesc *__restrict__ d, const int *pre, int nreq)
{
    const int p = batch_find(pre, nreq, blockIdx.x);
    const FDesc &q = d[p];
    const int bid = blockIdx.x - pre[p], nblk = pre[p + 1] - pre[p];
    const double rho = q.coef[0];
    double acc = 0.0;
    for (long i = (long)bid * kBlock + threadIdx.x; i < q.n; i += (long)nblk * kBlock) {
        const double r = q.t[i] - rho * q.v[i];
        acc += r * r;
    }
    const double sb = block_sum(acc);
    if (threadIdx.x == 0) q.part[kSpmvGrid + kMaxGrid + bid] = sb;
}

__global__ __launch_bounds__(kBlock) void kr_fin(const FDesc *__restrict__ d)
{
    const FDesc &q = d[blockIdx.x];
    const double rr = ordered_sum(q.part + kSpmvGrid + kMaxGrid, q.nd);
    if (threadIdx.x == 0) {
        const double vv = q.coef[1];
        q.out[0] = q.coef[0];
        q.out[1] = vv > 0.0 ? sqrt(rr / vv) : 0.0;
        q.out[2] = vv;
    }
}

// Pairs a refinement handles at once: one free column each (the columns
// after the k Ritz vectors except `avoid`, and the scratch column) and
// kRefinePart partials each in the projection partials.
int refine_width(const trlhip_ctx *c, int k, int avoid)
{
    int nfree = c->ncols + 1 - k - ((avoid >= k && avoid <= c->ncols) ? 1 : 0);
    const int np = (int)(((size_t)kProjGrid * c->E * (c->ncols + 1)) / kRefinePart);
    return std::max(0, std::min(std::min(k, nfree), np));
}

int refine_free_col(const trlhip_ctx *c, int k, int avoid, int j)
{
    int col = k;
    for (;; ++col) {
        if (col == avoid) continue;
        if (j-- == 0) return col;
    }
}

// Staging plan of recorded operations: for each phase (see BatchOp) one
// kb_ritz launch per restart width KT present, one kb_copy launch, and the
// refinements' kernels (pairs in chunks of refine_width; chunk j of every
// context together).  Records, prefix tables and restart matrices share one
// staging block of `bytes` bytes, copied to the device at once.  Operations
// of one phase belong to different contexts, so their order inside the
// phase is free.
enum { L_COPY = 0, L_REFINE = 9, L_NORM = 10 };  // 1..8: restart GEMMs of that width
struct OpsPlan {
    struct Launch {
        int kind, R;
        size_t doff, poff;  // first record (ODesc or FDesc array), first prefix entry
        size_t shm;
        double bytes, flops;
    };
    struct Pair {
        const BatchOp *op;
        int i, j;  // Ritz vector, slot in its chunk
    };
    std::vector<const BatchOp *> order;
    std::vector<Pair> pairs;
    std::vector<Launch> ls;
    size_t o_f = 0, o_pre = 0, o_y = 0, bytes = 0;
};

void ops_plan(const std::vector<BatchOp> &ops, const std::vector<double> &ybuf, OpsPlan &P)
{
    using Launch = OpsPlan::Launch;
    using Pair = OpsPlan::Pair;
    auto &order = P.order;
    auto &pairs = P.pairs;
    auto &ls = P.ls;
    int nph = 0;
    for (const BatchOp &op : ops) nph = std::max(nph, op.phase + 1);
    order.reserve(ops.size());
    size_t npre = 0;
    for (int ph = 0; ph < nph; ++ph) {
        for (int kind = 0; kind <= 8; ++kind) {
            Launch l{kind, 0, order.size(), npre, 0, 0.0};
            for (const BatchOp &op : ops) {
                if (op.phase != ph || op.type == OP_REFINE || op.type == OP_NORM) continue;
                const int okind = op.type == OP_COPY ? L_COPY : (op.k + 15) / 16;
                if (okind == kind) order.push_back(&op);
            }
            l.R = (int)(order.size() - l.doff);
            if (l.R == 0) continue;
            npre += (size_t)l.R + 1;
            ls.push_back(l);
        }
        {
            Launch l{L_NORM, 0, order.size(), npre, 0, 0.0};
            for (const BatchOp &op : ops)
                if (op.phase == ph && op.type == OP_NORM) order.push_back(&op);
            l.R = (int)(order.size() - l.doff);
            if (l.R > 0) {
                npre += (size_t)l.R + 1;
                ls.push_back(l);
            }
        }
        for (int chunk = 0;; ++chunk) {
            Launch l{L_REFINE, 0, pairs.size(), npre, 0, 0.0};
            bool more = false;
            for (const BatchOp &op : ops) {
                if (op.phase != ph || op.type != OP_REFINE) continue;
                const int w = refine_width(op.c, op.k, op.avoid);
                const int i0 = chunk * w, i1 = std::min(op.k, i0 + w);
                for (int i = i0; i < i1; ++i) pairs.push_back(Pair{&op, i, i - i0});
                if (i1 < op.k) more = true;
            }
            l.R = (int)(pairs.size() - l.doff);
            if (l.R > 0) {
                npre += 2 * ((size_t)l.R + 1);
                ls.push_back(l);
            }
            if (!more) break;
        }
    }
    P.o_f = sizeof(ODesc) * order.size();
    P.o_pre = P.o_f + sizeof(FDesc) * pairs.size();
    P.o_y = (P.o_pre + sizeof(int) * npre + 255) & ~(size_t)255;
    P.bytes = P.o_y + sizeof(double) * ybuf.size();
}

// Writes the plan's records into the staging block (h: host, d: the
// device address it is copied to).
void ops_fill(OpsPlan &P, const std::vector<double> &ybuf, char *h, const char *d)
{
    using Launch = OpsPlan::Launch;
    using Pair = OpsPlan::Pair;
    const size_t o_f = P.o_f, o_pre = P.o_pre, o_y = P.o_y;
    auto &order = P.order;
    auto &pairs = P.pairs;
    ODesc *ho = reinterpret_cast<ODesc *>(h);
    FDesc *hf = reinterpret_cast<FDesc *>(h + o_f);
    int *hp = reinterpret_cast<int *>(h + o_pre);
    const double *dy = reinterpret_cast<const double *>(d + o_y);
    if (!ybuf.empty()) memcpy(h + o_y, ybuf.data(), sizeof(double) * ybuf.size());
    for (Launch &l : P.ls) {
        int *pp = hp + l.poff;
        pp[0] = 0;
        if (l.kind == L_REFINE) {
            int *pd = pp + l.R + 1;
            pd[0] = 0;
            for (int r = 0; r < l.R; ++r) {
                const Pair &pr = pairs[l.doff + r];
                trlhip_ctx *c = pr.op->c;
                const trlhip_csr *A = pr.op->A;
                FDesc &f = hf[l.doff + r];
                f = FDesc{};
                f.n = c->n;
                f.nb = A->nb;
                f.npx = A->nb < kSpmvGrid ? A->nb : kSpmvGrid;
                f.nd = grid_for(c->n, kBlock * 8, kMaxGrid);
                f.rb = A->rb;
                f.rp = A->rp;
                f.ci = A->ci;
                f.val = A->val;
                f.v = colp<double>(c, pr.i);
                f.t = colp<double>(c, refine_free_col(c, pr.op->k, pr.op->avoid, pr.j));
                f.part = part_proj(c) + (size_t)pr.j * kRefinePart;
                f.coef = c->dh + 2 * pr.j;
                f.out = c->dres + 8 + 3 * pr.i;
                pp[r + 1] = pp[r] + f.npx;
                pd[r + 1] = pd[r] + f.nd;
                l.bytes += spmv_bytes(A, 1, 8, 8);
            }
            continue;
        }
        for (int i = 0; i < l.R; ++i) {
            const BatchOp &op = *order[l.doff + i];
            trlhip_ctx *c = op.c;
            ODesc &o = ho[l.doff + i];
            o = ODesc{};
            int nb;
            if (op.type == OP_NORM) {
                o.n = c->n;
                o.ld = c->ld;
                o.V = colp<double>(c, op.dst);
                o.part = part_update(c);
                o.o1 = c->dscal + S_RR2;
                o.o2 = c->dres + kHresNorm;
                nb = grid_for(c->n, kMinRowsUpd, kMaxGrid);
                o.g = nb;
                l.bytes += 16.0 * c->n;
            } else if (op.type == OP_COPY) {
                o.n = c->n;
                o.V = colp<double>(c, op.dst);
                o.src = op.src32 ? nullptr : colp<double>(c, op.src);
                o.src32 = op.src32;
                o.V32 = shadow32(c, op.dst);
                o.s = op.s;
                nb = grid_for(c->n, kBlock * 8, kMaxGrid);
                l.bytes += (c->V32 ? 20.0 : 16.0) * c->n;
            } else {
                o.n = c->n;
                o.ld = c->ld;
                o.V = colp<double>(c, 0);
                o.V32 = c->V32;
                o.Y = dy + op.yoff;
                o.lck = op.lck;
                o.jl = op.jl;
                o.k = op.k;
                o.ldy = op.ldy;
                const long rpb = 16L * kRitzTiles * (kBlock / 64);
                nb = (int)((c->n + rpb - 1) / rpb);
                const size_t shm = sizeof(double) * (size_t)((op.jl + 3) & ~3) * 16 * l.kind;
                l.shm = std::max(l.shm, shm);
                l.bytes += 8.0 * c->n * (op.jl + op.k) + (c->V32 ? 4.0 * c->n * op.k : 0.0);
                l.flops += 2.0 * c->n * op.jl * op.k;
            }
            pp[i + 1] = pp[i] + nb;
        }
    }
}

// Launches the plan's kernels on s from a staging block already copied to d.
int ops_launch(const OpsPlan &P, const char *h, const char *d, hipStream_t s)
{
    using Launch = OpsPlan::Launch;
    const int *hp = reinterpret_cast<const int *>(h + P.o_pre);
    const size_t o_f = P.o_f, o_pre = P.o_pre;
    const ODesc *dod = reinterpret_cast<const ODesc *>(d);
    const FDesc *dfd = reinterpret_cast<const FDesc *>(d + o_f);
    const int *dp = reinterpret_cast<const int *>(d + o_pre);
    for (const Launch &l : P.ls) {
        const int *pp = hp + l.poff;
        if (l.kind == L_REFINE) {
            const FDesc *f = dfd + l.doff;
            const int *dps = dp + l.poff, *dpd = dps + l.R + 1;
            const dim3 gs(pp[l.R]), gd(pp[l.R + 1 + l.R]), g1(l.R);
            {
                Timed t(TRLHIP_K_MISC, s, l.bytes);
                KLAUNCH(kr_spmv, gs, dim3(kBlock), 0, s, f, dps, l.R);
            }
            KLAUNCH(kr_dot, gd, dim3(kBlock), 0, s, f, dpd, l.R);
            KLAUNCH(kr_rho, g1, dim3(kBlock), 0, s, f);
            KLAUNCH(kr_res, gd, dim3(kBlock), 0, s, f, dpd, l.R);
            KLAUNCH(kr_fin, g1, dim3(kBlock), 0, s, f);
            continue;
        }
        const dim3 g(pp[l.R]);
        const ODesc *o = dod + l.doff;
        const int *dpp = dp + l.poff;
        if (l.kind == L_COPY) {
            Timed t(TRLHIP_K_MISC, s, l.bytes);
            KLAUNCH(kb_copy, g, dim3(kBlock), 0, s, o, dpp, l.R);
            continue;
        }
        if (l.kind == L_NORM) {
            {
                Timed t(TRLHIP_K_MISC, s, l.bytes);
                KLAUNCH(kb_norm, g, dim3(kBlock), 0, s, o, dpp, l.R);
            }
            KLAUNCH(kb_norm_fin, dim3(l.R), dim3(kBlock), 0, s, o);
            continue;
        }
        Timed t(TRLHIP_K_BRITZ, s, l.bytes, l.flops);
#define KB_RITZ_CASE(K)                                                            \
    case K:                                                                        \
        KLAUNCH((kb_ritz<K>), g, dim3(kBlock), l.shm, s, o, dpp, l.R);             \
        break;
        switch (l.kind) {
            KB_RITZ_CASE(1) KB_RITZ_CASE(2) KB_RITZ_CASE(3) KB_RITZ_CASE(4)
            KB_RITZ_CASE(5) KB_RITZ_CASE(6) KB_RITZ_CASE(7) KB_RITZ_CASE(8)
        }
#undef KB_RITZ_CASE
    }
    return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
}


// Runs recorded operations outside a batch: stage (h, d: pinned / device,
// cap bytes, grown as needed), copy, launch.
int ops_run(const std::vector<BatchOp> &ops, const std::vector<double> &ybuf, char *&h, char *&d,
            size_t &cap, hipStream_t s)
{
    OpsPlan P;
    ops_plan(ops, ybuf, P);
    const size_t bytes = P.bytes;
    if (bytes > cap) {
        if (h) (void)hipHostFree(h);
        if (d) (void)hipFree(d);
        h = d = nullptr;
        cap = 0;
        const size_t ncap = bytes + bytes / 2 + 4096;
        if (hipHostMalloc((void **)&h, ncap, 0) != hipSuccess || hipMalloc((void **)&d, ncap) != hipSuccess)
            return TRLHIP_ENOMEM;
        cap = ncap;
    }
    ops_fill(P, ybuf, h, d);
    TRY(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
    return ops_launch(P, h, d, s);
}

}  // namespace

// ================================================================ C ABI

extern "C" {

const trlhip_csr *trlhip_csr_from_op(void (*op)(void), void *mvparam, int vtype)
{
    const trlhip_csr *A = (const trlhip_csr *)mvparam;
    if (!A) return nullptr;
    if (vtype == TRLHIP_C128) {
        if (op != (void (*)(void))trlhip_csr_zmatvec) return nullptr;
        return (A->magic == kCsrMagic && A->vtype == TRLHIP_C128) ? A : nullptr;
    }
    if (op != (void (*)(void))trlhip_csr_matvec) return nullptr;
    if (A->magic != kCsrMagic || A->vtype == TRLHIP_C128) return nullptr;
    return A;
}

int trlhip_csr_is_mixed(const trlhip_csr *A) { return A && A->vtype == TRLHIP_R32; }

int trlhip_csr_is_packed(const trlhip_csr *A) { return A && A->magic == kCsrMagic && A->pk ? 1 : 0; }

int trlhip_csr_is_paneled(const trlhip_csr *A)
{
    return A && A->magic == kCsrMagic && A->pn_on && !A->lpr ? 1 : 0;
}

int trlhip_csr_stencil(const trlhip_csr *A, int *L)
{
    if (!A || A->magic != kCsrMagic) return 0;
    if (L) *L = A->st_L;
    return A->st_dim;
}

int trlhip_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// SpMV kernel for a matrix: vector CSR with LPR = the power of two at or
// above the mean row length (2..32) when no row is longer than 8*LPR
// (balanced rows); otherwise CSR-stream, whose row blocks absorb uneven
// rows.  A choice made from the structure alone, so a matrix always runs
// the same kernel (and reduction order).
// Select the SpMV kernel: 0 = CSR-stream (row blocks), else the vector
// kernel with that many lanes per row (and its grid).
static void set_lpr(trlhip_csr *A, int lpr)
{
    A->lpr = lpr;
    if (!lpr) return;
    const int U = A->E == 1 ? 4 : 2;
    const long rows_per_block = (long)(64 / lpr) * U * (kBlock / 64);
    const long g = (A->n + rows_per_block - 1) / rows_per_block;
    A->vgrid = (int)std::min<long>(std::max<long>(g, 1), kMaxGrid);
}

static void choose_spmv(trlhip_csr *A, const int *row_ptr)
{
    const int n = A->n;
    int maxlen = 0;
    for (int i = 0; i < n; ++i) maxlen = std::max(maxlen, row_ptr[i + 1] - row_ptr[i]);
    const double avg = (double)A->nnz / n;
    int lpr = 2;
    while (lpr < 32 && lpr < avg) lpr <<= 1;
    // the row-block (CSR-stream) kernel is the default: measured faster on
    // regular stencils too (6.0 vs 4.0 TB/s on the 128^3 Laplacian);
    // trlhip_csr_autotune times the vector kernel against it per matrix
    A->lpr_vec = maxlen <= 8 * lpr ? lpr : 0;
    set_lpr(A, 0);
}

}  // extern "C"

namespace {

// Packs the nonzeros of a real CSR matrix for the CSR-stream kernel (see
// spmv_stream_body): false when it has more than 255 distinct off-diagonal
// values (compared bit for bit), a row with two diagonal entries, or a
// column more than 2^23 rows from its row block's first row.
bool pack_csr(int n, long nnz, const int *rp, const int *ci, const double *val,
              const std::vector<int> &rb, std::vector<uint32_t> &pk, std::vector<double> &dg,
              std::vector<double> &tab)
{
    pk.assign((size_t)nnz, 0u);
    dg.assign((size_t)n, 0.0);
    tab.clear();
    std::unordered_map<uint64_t, uint32_t> idx;
    for (size_t blk = 0; blk + 1 < rb.size(); ++blk) {
        const long r0 = rb[blk];
        for (int row = rb[blk]; row < rb[blk + 1]; ++row) {
            bool diag = false;
            for (long k = rp[row]; k < rp[row + 1]; ++k) {
                const long off = (long)ci[k] - r0;
                if (off < -(1L << 23) || off >= (1L << 23)) return false;
                uint32_t vi;
                if (ci[k] == row) {
                    if (diag) return false;
                    diag = true;
                    dg[(size_t)row] = val[k];
                    vi = 255u;
                } else {
                    uint64_t bits;
                    memcpy(&bits, &val[k], sizeof bits);
                    auto it = idx.find(bits);
                    if (it == idx.end()) {
                        if (tab.size() >= 255) return false;
                        vi = (uint32_t)tab.size();
                        idx.emplace(bits, vi);
                        tab.push_back(val[k]);
                    } else {
                        vi = it->second;
                    }
                }
                pk[(size_t)k] = (vi << 24) | ((uint32_t)off & 0xFFFFFFu);
            }
        }
    }
    return true;
}

// A periodic nearest-neighbour stencil on an L^D lattice (D = 2 or 3,
// L >= 3, site x + L y (+ L^2 z)): every row holds exactly its 2D lattice
// neighbours and itself, columns ascending, and every off-diagonal value is
// the same double (bit for bit), returned in *off.  Returns D, or 0.
int detect_stencil(int n, long nnz, const int *rp, const int *ci, const double *val, int *L_out,
                   double *off)
{
    for (int D = 2; D <= 3; ++D) {
        const int K = 2 * D + 1;
        int L = (int)std::lround(std::pow((double)n, 1.0 / D));
        while ((long)std::pow((double)L, D) > n) --L;
        while ((long)std::pow((double)(L + 1), D) <= n) ++L;
        if (L < 3 || (D == 2 ? (long)L * L : (long)L * L * L) != n || nnz != (long)K * n) continue;
        bool ok = true;
        bool have_off = false;
        uint64_t offbits = 0;
        for (int row = 0; row < n && ok; ++row) {
            if (rp[row] != (long)K * row) {
                ok = false;
                break;
            }
            const int xs = row % L, rest = row / L;
            const int ys = D == 2 ? rest : rest % L, zs = D == 2 ? 0 : rest / L;
            int c[7];
            c[0] = row;
            c[1] = row - xs + (xs + L - 1) % L;
            c[2] = row - xs + (xs + 1) % L;
            c[3] = row - ys * L + ((ys + L - 1) % L) * L;
            c[4] = row - ys * L + ((ys + 1) % L) * L;
            if (D == 3) {
                c[5] = row - zs * L * L + ((zs + L - 1) % L) * L * L;
                c[6] = row - zs * L * L + ((zs + 1) % L) * L * L;
            }
            std::sort(c, c + K);
            for (int k = 0; k < K && ok; ++k) {
                const long e = (long)rp[row] + k;
                if (ci[e] != c[k]) ok = false;
                if (c[k] == row) continue;
                uint64_t bits;
                memcpy(&bits, &val[e], sizeof bits);
                if (!have_off) {
                    have_off = true;
                    offbits = bits;
                    *off = val[e];
                } else if (bits != offbits) {
                    ok = false;
                }
            }
        }
        if (ok && have_off) {
            *L_out = L;
            return D;
        }
    }
    return 0;
}

// Host side of the column-panel SpMV (see k_spmv_panel): W columns per
// panel.  Fills the nonzeros per panel in row order, the chunks and the
// XCD map of blocks to chunks, the run (partial) numbering and each row's
// runs in panel order, and copies them to the device.  Returns 0, or an
// error (A's panel fields stay partly allocated; trlhip_csr_destroy frees
// them).
int build_panels(trlhip_csr *A, const int *rp, const int *ci, const double *val, int W)
{
    const int n = A->n;
    const long nnz = A->nnz;
    const int P = (n + W - 1) / W;
    std::vector<long> pe((size_t)P + 1, 0);
    for (long k = 0; k < nnz; ++k) ++pe[(size_t)ci[k] / W + 1];
    for (int p = 0; p < P; ++p) pe[p + 1] += pe[p];
    std::vector<int> erow((size_t)nnz), ecol((size_t)nnz);
    std::vector<double> ev((size_t)nnz);
    {
        std::vector<long> cur(pe.begin(), pe.end() - 1);
        for (int i = 0; i < n; ++i)
            for (long k = rp[i]; k < rp[i + 1]; ++k) {
                const long q = cur[(size_t)ci[k] / W]++;
                erow[q] = i;
                ecol[q] = ci[k];
                ev[q] = val[k];
            }
    }
    // chunks (panel-major), their runs, each row's runs in panel order
    std::vector<long> cstart;
    std::vector<int> cpanel, segbase;
    long nseg = 0;
    std::vector<int> rcount((size_t)n + 1, 0);
    for (int p = 0; p < P; ++p)
        for (long k0 = pe[p]; k0 < pe[p + 1]; k0 += kPanelChunk) {
            const long k1 = std::min(pe[p + 1], k0 + kPanelChunk);
            cstart.push_back(k0);
            cpanel.push_back(p);
            segbase.push_back((int)nseg);
            for (long k = k0; k < k1; ++k)
                if (k == k0 || erow[k] != erow[k - 1]) {
                    ++nseg;
                    ++rcount[(size_t)erow[k] + 1];
                }
        }
    const int nc = (int)cstart.size();
    cstart.push_back(nnz);
    if (nseg >= (1L << 31)) return TRLHIP_EINVAL;
    for (int i = 0; i < n; ++i) rcount[(size_t)i + 1] += rcount[i];
    std::vector<int> rsidx((size_t)std::max<long>(nseg, 1));
    {
        std::vector<int> cur(rcount.begin(), rcount.end() - 1);
        long sid = 0;
        for (int c = 0; c < nc; ++c)
            for (long k = cstart[c]; k < cstart[c + 1]; ++k)
                if (k == cstart[c] || erow[k] != erow[k - 1]) rsidx[(size_t)cur[erow[k]]++] = (int)sid++;
    }
    // XCD map: workgroup 8 j + x runs chunk j of XCD x's list, which holds
    // the chunks of panels x, x + 8, x + 16, ... in order (-1: idle)
    std::vector<std::vector<int>> lists(8);
    for (int c = 0; c < nc; ++c) lists[(size_t)cpanel[c] % 8].push_back(c);
    size_t m = 0;
    for (auto &l : lists) m = std::max(m, l.size());
    std::vector<int> perm(8 * std::max<size_t>(m, 1), -1);
    for (int x = 0; x < 8; ++x)
        for (size_t j = 0; j < lists[x].size(); ++j) perm[8 * j + x] = lists[x][j];
    A->pn_nchunk = nc;
    A->pn_nslot = (int)perm.size();
    A->pn_nseg = nseg;
    A->pn_grid = (int)std::min<long>(std::max<long>((n + kBlock - 1) / kBlock, 1), kSpmvGrid);
    auto put = [](void **d, const void *h, size_t bytes) -> int {
        if (hipMalloc(d, bytes ? bytes : 8) != hipSuccess) return TRLHIP_ENOMEM;
        if (bytes && hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) return TRLHIP_EHIP;
        return 0;
    };
    int rc = put((void **)&A->pn_perm, perm.data(), sizeof(int) * perm.size());
    if (!rc) rc = put((void **)&A->pn_cstart, cstart.data(), sizeof(long) * cstart.size());
    if (!rc) rc = put((void **)&A->pn_segbase, segbase.data(), sizeof(int) * segbase.size());
    if (!rc) rc = put((void **)&A->pn_erow, erow.data(), sizeof(int) * erow.size());
    if (!rc) rc = put((void **)&A->pn_ecol, ecol.data(), sizeof(int) * ecol.size());
    if (!rc) rc = put((void **)&A->pn_eval, ev.data(), sizeof(double) * ev.size());
    if (!rc) rc = put((void **)&A->pn_rsptr, rcount.data(), sizeof(int) * rcount.size());
    if (!rc) rc = put((void **)&A->pn_rsidx, rsidx.data(), sizeof(int) * rsidx.size());
    if (!rc && hipMalloc((void **)&A->pn_sp, sizeof(double) * (size_t)std::max<long>(nseg, 1)) != hipSuccess)
        rc = TRLHIP_ENOMEM;
    return rc;
}

// Whether the column-panel SpMV pays: x larger than a few L2s and most
// nonzeros reaching beyond half a panel's width from their row.
}  // namespace

extern "C" {

int trlhip_csr_create(trlhip_csr **out, int n, long nnz, const int *row_ptr, const int *col_idx,
                      const void *values, int vtype)
{
    if (!out || n <= 0 || nnz < 0 || !row_ptr || (nnz > 0 && (!col_idx || !values)))
        return TRLHIP_EINVAL;
    if (vtype != TRLHIP_R64 && vtype != TRLHIP_C128 && vtype != TRLHIP_R32) return TRLHIP_EINVAL;
    if (row_ptr[0] != 0 || row_ptr[n] != nnz) return TRLHIP_EINVAL;
    for (int i = 0; i < n; ++i)
        if (row_ptr[i + 1] < row_ptr[i]) return TRLHIP_EINVAL;
    for (long k = 0; k < nnz; ++k)
        if (col_idx[k] < 0 || col_idx[k] >= n) return TRLHIP_EINVAL;
    if (trlhip_device_count() <= 0) return TRLHIP_ENODEV;
    trlhip_csr *A = (trlhip_csr *)calloc(1, sizeof(trlhip_csr));
    if (!A) return TRLHIP_ENOMEM;
    A->magic = kCsrMagic;
    A->n = n;
    A->nnz = nnz;
    A->vtype = vtype;
    A->E = vtype == TRLHIP_C128 ? 2 : 1;
    const std::vector<int> rb = row_blocks(n, row_ptr);
    A->nb = (int)rb.size() - 1;
    choose_spmv(A, row_ptr);
    const size_t nz1 = (size_t)(nnz > 0 ? nnz : 1);
    int rc = 0;
    if (!rc && hipMalloc((void **)&A->rb, sizeof(int) * rb.size()) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMalloc((void **)&A->rp, sizeof(int) * (size_t)(n + 1)) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMalloc((void **)&A->ci, sizeof(int) * nz1) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMalloc((void **)&A->val, sizeof(double) * A->E * nz1) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMemcpy(A->rb, rb.data(), sizeof(int) * rb.size(), hipMemcpyHostToDevice) != hipSuccess) rc = TRLHIP_EHIP;
    if (!rc && hipMemcpy(A->rp, row_ptr, sizeof(int) * (size_t)(n + 1), hipMemcpyHostToDevice) != hipSuccess) rc = TRLHIP_EHIP;
    if (!rc && nnz > 0 && hipMemcpy(A->ci, col_idx, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess) rc = TRLHIP_EHIP;
    if (!rc && nnz > 0 && hipMemcpy(A->val, values, sizeof(double) * A->E * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess) rc = TRLHIP_EHIP;
    // packed nonzeros for the CSR-stream kernel when the values allow
    // (TRLHIP_CSR_PACK=0: never)
    const char *ep = getenv("TRLHIP_CSR_PACK");
    if (!rc && A->E == 1 && nnz > 0 && !(ep && ep[0] == '0')) {
        std::vector<uint32_t> pk;
        std::vector<double> dg, tab;
        if (pack_csr(n, nnz, row_ptr, col_idx, (const double *)values, rb, pk, dg, tab)) {
            A->ntab = (int)tab.size();
            if (tab.empty()) tab.push_back(0.0);
            if (hipMalloc((void **)&A->pk, sizeof(uint32_t) * pk.size()) != hipSuccess ||
                hipMalloc((void **)&A->dg, sizeof(double) * dg.size()) != hipSuccess ||
                hipMalloc((void **)&A->tab, sizeof(double) * tab.size()) != hipSuccess)
                rc = TRLHIP_ENOMEM;
            if (!rc && (hipMemcpy(A->pk, pk.data(), sizeof(uint32_t) * pk.size(), hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(A->dg, dg.data(), sizeof(double) * dg.size(), hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(A->tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) != hipSuccess))
                rc = TRLHIP_EHIP;
            // a periodic stencil needs neither the packed words nor the
            // row pointers (TRLHIP_CSR_STENCIL=0: never)
            const char *es = getenv("TRLHIP_CSR_STENCIL");
            if (!rc && !(es && es[0] == '0'))
                A->st_dim = detect_stencil(n, nnz, row_ptr, col_idx, (const double *)values, &A->st_L,
                                           &A->st_off);
            // the stencil kernels form row block b as rows [kRowsBlk b,
            // kRowsBlk (b + 1)) instead of reading rb (2D + 1 <= 7 nonzeros
            // a row always fill kRowsBlk rows a block)
            for (int b = 0; A->st_dim && b < A->nb; ++b)
                if (rb[b] != b * kRowsBlk || rb[b + 1] != std::min(n, (b + 1) * kRowsBlk)) A->st_dim = 0;
        }
    }
    // column panels on request (TRLHIP_CSR_PANEL=1; TRLHIP_CSR_PANEL_COLS:
    // the panel width, a multiple of 64).  Not by default: on a random
    // symmetric pattern with n = 2^23 and 17 nonzeros per row the two panel
    // kernels together take 2.94 ms a product against the CSR-stream
    // kernel's 2.62 ms.
    const char *epn = getenv("TRLHIP_CSR_PANEL");
    const char *epw = getenv("TRLHIP_CSR_PANEL_COLS");
    int W = kPanelCols;
    if (epw && atoi(epw) >= 64) W = atoi(epw) & ~63;
    if (!rc && A->E == 1 && nnz > 0 && !A->pk && epn && epn[0] == '1') {
        rc = build_panels(A, row_ptr, col_idx, (const double *)values, W);
        if (!rc) A->pn_on = 1;
    }
    if (rc) {
        trlhip_csr_destroy(A);
        return rc;
    }
    *out = A;
    return 0;
}

int trlhip_csr_autotune(trlhip_csr *A, int reps, int *lpr_out, float *ms_out)
{
    if (!A || A->magic != kCsrMagic || reps < 0) return TRLHIP_EINVAL;
    if (reps == 0) reps = 10;
    const size_t ne = (size_t)A->n * A->E;
    double *x = nullptr, *y = nullptr, *part = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (!A->s && hipStreamCreate(&A->s) != hipSuccess) return TRLHIP_EHIP;
    if (hipMalloc((void **)&x, sizeof(double) * ne) != hipSuccess ||
        hipMalloc((void **)&y, sizeof(double) * ne) != hipSuccess ||
        hipMalloc((void **)&part, sizeof(double) * kSpmvGrid) != hipSuccess)
        rc = TRLHIP_ENOMEM;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = TRLHIP_EHIP;
    if (!rc && hipMemsetAsync(x, 0, sizeof(double) * ne, A->s) != hipSuccess) rc = TRLHIP_EHIP;
    // candidates: 1 the column-panel kernel (when A has panels), 0 the
    // CSR-stream kernel (stencil or packed when A allows), 2..32 lanes per
    // row of the vector kernel
    const int prev = A->lpr ? A->lpr : (A->pn_on ? 1 : 0);
    int best = prev;
    float best_ms = 0.f;
    int cand[7], nc = 0;
    if (A->pn_nchunk) cand[nc++] = 1;
    for (int c : {0, 2, 4, 8, 16, 32}) cand[nc++] = c;
    auto use = [&](int c) {
        set_lpr(A, c == 1 ? 0 : c);
        A->pn_on = c == 1;
    };
    for (int i = 0; i < nc && !rc; ++i) {
        use(cand[i]);
        for (int w = 0; w < 2 && !rc; ++w) rc = launch_spmv(A, A->vtype, x, y, part, A->s);
        if (!rc && hipEventRecord(e0, A->s) != hipSuccess) rc = TRLHIP_EHIP;
        for (int r = 0; r < reps && !rc; ++r) rc = launch_spmv(A, A->vtype, x, y, part, A->s);
        if (!rc && hipEventRecord(e1, A->s) != hipSuccess) rc = TRLHIP_EHIP;
        float ms = 0.f;
        if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = TRLHIP_EHIP;
        if (!rc && (i == 0 || ms < best_ms)) {
            best_ms = ms;
            best = cand[i];
        }
    }
    use(rc ? prev : best);
    if (!rc && lpr_out) *lpr_out = best;
    if (!rc && ms_out) *ms_out = best_ms / (float)reps;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (x) (void)hipFree(x);
    if (y) (void)hipFree(y);
    if (part) (void)hipFree(part);
    return rc;
}

int trlhip_csr_set_spmv(trlhip_csr *A, int lpr)
{
    if (!A || A->magic != kCsrMagic) return TRLHIP_EINVAL;
    if (lpr == 1) {  // the column-panel kernel, when A has panels
        if (!A->pn_nchunk) return TRLHIP_EINVAL;
        set_lpr(A, 0);
        A->pn_on = 1;
        return 0;
    }
    if (lpr != 0 && lpr != 2 && lpr != 4 && lpr != 8 && lpr != 16 && lpr != 32) return TRLHIP_EINVAL;
    set_lpr(A, lpr);
    if (lpr == 0) A->pn_on = 0;
    return 0;
}

int trlhip_proj_tiles(int cb_fp64, int cb_fp32)
{
    auto ok = [](int cb) { return cb == 0 || cb == 4 || cb == 8 || cb == 16; };
    if (!ok(cb_fp64) || !ok(cb_fp32)) return TRLHIP_EINVAL;
    g_proj_cb[0] = cb_fp64 ? cb_fp64 : 8;
    g_proj_cb[1] = cb_fp32 ? cb_fp32 : 8;
    return 0;
}

int trlhip_proj_tiles_get(int *cb_fp64, int *cb_fp32)
{
    if (cb_fp64) *cb_fp64 = g_proj_cb[0];
    if (cb_fp32) *cb_fp32 = g_proj_cb[1];
    return 0;
}

int trlhip_proj_autotune(int n, int m, int reps, int *cb_out, float *ms_out)
{
    if (n < 1 || m < 1 || m > 4096 || reps < 0) return TRLHIP_EINVAL;
    if (reps == 0) reps = 10;
    trlhip_ctx *c = nullptr;
    if (int rc = trlhip_ctx_create(&c, n, m + 1, TRLHIP_R32)) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = TRLHIP_EHIP;
    const int cand[3] = {4, 8, 16};
    for (int prec = 0; prec < 2 && !rc; ++prec) {
        int best = 8;
        float best_ms = 0.f;
        for (int i = 0; i < 3 && !rc; ++i) {
            // trial widths go to the launches directly: solves running
            // meanwhile keep the process-wide setting
            for (int w = 0; w < 2 && !rc; ++w) rc = run_proj(c, 0, m, m, c->dh, prec == 1, cand[i]);
            if (!rc && hipEventRecord(e0, c->s) != hipSuccess) rc = TRLHIP_EHIP;
            for (int r = 0; r < reps && !rc; ++r)
                rc = run_proj(c, 0, m, m, c->dh, prec == 1, cand[i]);
            if (!rc && hipEventRecord(e1, c->s) != hipSuccess) rc = TRLHIP_EHIP;
            float ms = 0.f;
            if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = TRLHIP_EHIP;
            if (!rc && (i == 0 || ms < best_ms)) {
                best_ms = ms;
                best = cand[i];
            }
        }
        if (!rc) g_proj_cb[prec] = best;
        if (!rc && cb_out) cb_out[prec] = best;
        if (!rc && ms_out) ms_out[prec] = best_ms / (float)reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    trlhip_ctx_destroy(c);
    return rc;
}

int trlhip_csr_destroy(trlhip_csr *A)
{
    if (!A) return 0;
    if (A->s) (void)hipStreamSynchronize(A->s);
    if (A->rb) (void)hipFree(A->rb);
    if (A->rp) (void)hipFree(A->rp);
    if (A->ci) (void)hipFree(A->ci);
    if (A->val) (void)hipFree(A->val);
    if (A->pk) (void)hipFree(A->pk);
    if (A->dg) (void)hipFree(A->dg);
    if (A->tab) (void)hipFree(A->tab);
    for (void *q : {(void *)A->pn_perm, (void *)A->pn_cstart, (void *)A->pn_segbase,
                    (void *)A->pn_erow, (void *)A->pn_ecol, (void *)A->pn_eval, (void *)A->pn_rsptr,
                    (void *)A->pn_rsidx, (void *)A->pn_sp})
        if (q) (void)hipFree(q);
    if (A->dx) (void)hipFree(A->dx);
    if (A->dy) (void)hipFree(A->dy);
    if (A->part) (void)hipFree(A->part);
    if (A->dscal) (void)hipFree(A->dscal);
    if (A->s) (void)hipStreamDestroy(A->s);
    A->magic = 0;
    free(A);
    return 0;
}

// Host-callable operator: stage x in, y = A x (FP64 values) on the device,
// stage y out.
static int csr_matvec_host(trlhip_csr *A, int nrow, int ncol, const double *x, int ldx,
                           double *y, int ldy)
{
    const int E = A->E;
    const int vt = A->vtype == TRLHIP_C128 ? TRLHIP_C128 : TRLHIP_R64;
    if (!A->dx) {
        // staging for host callers, created on first use
        if (!A->s && hipStreamCreate(&A->s) != hipSuccess) return TRLHIP_EHIP;
        if (hipMalloc((void **)&A->dx, sizeof(double) * E * (size_t)A->n) != hipSuccess) return TRLHIP_ENOMEM;
        if (hipMalloc((void **)&A->dy, sizeof(double) * E * (size_t)A->n) != hipSuccess) return TRLHIP_ENOMEM;
        if (hipMalloc((void **)&A->part, sizeof(double) * (size_t)kSpmvGrid) != hipSuccess) return TRLHIP_ENOMEM;
        if (hipMalloc((void **)&A->dscal, sizeof(double) * 4) != hipSuccess) return TRLHIP_ENOMEM;
    }
    for (int j = 0; j < ncol; ++j) {
        TRY(hipMemcpyAsync(A->dx, x + (size_t)j * ldx * E, sizeof(double) * E * (size_t)nrow,
                           hipMemcpyHostToDevice, A->s));
        if (launch_spmv(A, vt, A->dx, A->dy, A->part, A->s)) return TRLHIP_EHIP;
        TRY(hipMemcpyAsync(y + (size_t)j * ldy * E, A->dy, sizeof(double) * E * (size_t)nrow,
                           hipMemcpyDeviceToHost, A->s));
    }
    TRY(hipStreamSynchronize(A->s));
    return 0;
}

void trlhip_csr_matvec(const int nrow, const int ncol, const double *x, const int ldx, double *y,
                       const int ldy, void *mvparam)
{
    trlhip_csr *A = (trlhip_csr *)mvparam;
    if (!A || A->magic != kCsrMagic || A->E != 1 || nrow != A->n ||
        csr_matvec_host(A, nrow, ncol, x, ldx, y, ldy) != 0) {
        for (int j = 0; j < ncol; ++j)
            for (int i = 0; i < nrow; ++i) y[(size_t)j * ldy + i] = __builtin_nan("");
    }
}

void trlhip_csr_zmatvec(const int nrow, const int ncol, const void *x, const int ldx, void *y,
                        const int ldy, void *mvparam)
{
    trlhip_csr *A = (trlhip_csr *)mvparam;
    double *yd = (double *)y;
    if (!A || A->magic != kCsrMagic || A->E != 2 || nrow != A->n ||
        csr_matvec_host(A, nrow, ncol, (const double *)x, ldx, yd, ldy) != 0) {
        for (int j = 0; j < ncol; ++j)
            for (int i = 0; i < 2 * nrow; ++i) yd[(size_t)j * ldy * 2 + i] = __builtin_nan("");
    }
}

int trlhip_ctx_destroy(trlhip_ctx *c)
{
    if (!c) return 0;
    if (c->arena) {  // the batch owns the memory and the stream
        free(c);
        return 0;
    }
    if (c->s) (void)hipStreamSynchronize(c->s);
    if (c->V) (void)hipFree(c->V);
    if (c->V32) (void)hipFree(c->V32);
    if (c->part) (void)hipFree(c->part);
    if (c->dscal) (void)hipFree(c->dscal);
    if (c->dcoef) (void)hipFree(c->dcoef);
    if (c->dh) (void)hipFree(c->dh);
    if (c->dY) (void)hipFree(c->dY);
    if (c->hbuf) (void)hipHostFree(c->hbuf);
    if (c->hres) (void)hipHostFree(c->hres);
    if (c->hops) (void)hipHostFree(c->hops);
    if (c->dops) (void)hipFree(c->dops);
    if (c->hY) (void)hipHostFree(c->hY);
    if (c->s && !c->batch) (void)hipStreamDestroy(c->s);
    free(c);
    return 0;
}

int trlhip_ctx_create(trlhip_ctx **out, int n, int ncols, int vtype)
{
    return trlhip_ctx_create_batched(out, n, ncols, vtype, nullptr);
}

int trlhip_ctx_create_batched(trlhip_ctx **out, int n, int ncols, int vtype, trlhip_batch *b)
{
    if (!out || n <= 0 || ncols <= 0) return TRLHIP_EINVAL;
    if (vtype != TRLHIP_R64 && vtype != TRLHIP_C128 && vtype != TRLHIP_R32) return TRLHIP_EINVAL;
    // (a batch already holds streams on the device: no runtime query per
    // context)
    if (!b && trlhip_device_count() <= 0) return TRLHIP_ENODEV;
    trlhip_ctx *c = (trlhip_ctx *)calloc(1, sizeof(trlhip_ctx));
    if (!c) return TRLHIP_ENOMEM;
    c->n = n;
    c->ld = (n + 63) & ~63;
    c->ncols = ncols;
    c->vtype = vtype;
    c->E = vtype == TRLHIP_C128 ? 2 : 1;
    c->esz = 8 * (size_t)c->E;
    const size_t ncoef = 2 * (size_t)(ncols + 2);
    const size_t vbytes = c->esz * (size_t)c->ld * (ncols + 1);  // + scratch column
    int rc = 0;
    c->pend_dst = -1;
    c->hcap = 16 + 2 * ncoef;
    if (b) {
        // batched: everything from the batch's arenas, nothing to synchronize
        std::lock_guard<std::mutex> lk(b->arena_mu);
        c->batch = b;
        c->s = b->side;
        c->arena = 1;
        const size_t np = part_need(c->E, ncols);
        // restart matrices, and the polish's two Gram matrices (trlhip_gram)
        const size_t ny = std::max((size_t)(ncols + 1) * (ncols + 1), 2 * (size_t)ncols * ncols);
        c->V = (char *)arena_get(b, false, vbytes);
        if (vtype == TRLHIP_R32)
            c->V32 = (float *)arena_get(b, false, sizeof(float) * (size_t)c->ld * (ncols + 1));
        c->part = (double *)arena_get(b, false, sizeof(double) * np);
        c->partcap = np;
        c->dscal = (double *)arena_get(b, false, sizeof(double) * 16);
        c->dcoef = (double *)arena_get(b, false, sizeof(double) * ncoef);
        c->dh = (double *)arena_get(b, false, sizeof(double) * ncoef);
        c->dY = (double *)arena_get(b, false, sizeof(double) * ny);
        c->dYcap = ny;
        c->hbuf = (double *)arena_get(b, true, sizeof(double) * c->hcap);
        c->hres = (double *)arena_get(b, true, sizeof(double) * hres_count(ncols), (void **)&c->dres);
        c->hY = (double *)arena_get(b, true, sizeof(double) * ny);
        c->hstage = (char *)arena_get(b, true, c->esz * (size_t)c->ld);
        if (!c->V || (vtype == TRLHIP_R32 && !c->V32) || !c->part || !c->dscal || !c->dcoef ||
            !c->dh || !c->dY || !c->hbuf || !c->hres || !c->hY || !c->hstage) {
            free(c);
            return TRLHIP_ENOMEM;
        }
        *out = c;
        return 0;
    }
    if (hipStreamCreate(&c->s) != hipSuccess) rc = TRLHIP_EHIP;
    if (!rc && hipMalloc((void **)&c->V, vbytes) != hipSuccess) rc = TRLHIP_ENOMEM;
    const size_t sbytes = sizeof(float) * (size_t)c->ld * (ncols + 1);
    if (!rc && vtype == TRLHIP_R32 && hipMalloc((void **)&c->V32, sbytes) != hipSuccess)
        rc = TRLHIP_ENOMEM;
    if (!rc && c->V32 && hipMemsetAsync(c->V32, 0, sbytes, c->s) != hipSuccess) rc = TRLHIP_EHIP;
    if (!rc) rc = ensure_part(c, part_need(c->E, ncols));
    if (!rc && hipMalloc((void **)&c->dscal, sizeof(double) * 16) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMalloc((void **)&c->dcoef, sizeof(double) * ncoef) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMalloc((void **)&c->dh, sizeof(double) * ncoef) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipHostMalloc((void **)&c->hbuf, sizeof(double) * c->hcap, 0) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipHostMalloc((void **)&c->hres, sizeof(double) * hres_count(ncols),
                             hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess)
        rc = TRLHIP_ENOMEM;
    if (!rc && hipHostGetDevicePointer((void **)&c->dres, c->hres, 0) != hipSuccess) rc = TRLHIP_EHIP;
    if (!rc && hipMemsetAsync(c->V, 0, vbytes, c->s) != hipSuccess) rc = TRLHIP_EHIP;
    if (!rc && hipMemsetAsync(c->dscal, 0, sizeof(double) * 16, c->s) != hipSuccess) rc = TRLHIP_EHIP;
    if (!rc && hipStreamSynchronize(c->s) != hipSuccess) rc = TRLHIP_EHIP;
    if (rc) {
        trlhip_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

int trlhip_ctx_ld(const trlhip_ctx *c) { return c ? c->ld : 0; }

int trlhip_col_set(trlhip_ctx *c, int col, const void *x)
{
    if (!c || !x || !valid_col(c, col)) return TRLHIP_EINVAL;
    if (int rc = settle(c)) return rc;
    // in a batch through the context's pinned staging (an asynchronous copy;
    // a copy from pageable memory would block this host thread)
    const void *src = x;
    if (c->batch && c->hstage) {
        memcpy(c->hstage, x, c->esz * (size_t)c->n);
        src = c->hstage;
    }
    TRY(hipMemcpyAsync(c->V + (size_t)col * c->ld * c->esz, src, c->esz * (size_t)c->n,
                       hipMemcpyHostToDevice, c->s));
    if (c->V32) {  // refresh the FP32 shadow
        if (int rc = run_scale_copy(c, col, col, 1.0)) return rc;
    }
    return ctx_wait(c);  // x may be released by the caller on return
}

int trlhip_col_get(trlhip_ctx *c, int col, void *x)
{
    if (!c || !x || !valid_col(c, col)) return TRLHIP_EINVAL;
    if (int rc = settle(c)) return rc;
    if (c->batch && c->hstage) {  // pinned staging, as in trlhip_col_set
        TRY(hipMemcpyAsync(c->hstage, c->V + (size_t)col * c->ld * c->esz, c->esz * (size_t)c->n,
                           hipMemcpyDeviceToHost, c->s));
        if (int rc = ctx_wait(c)) return rc;
        memcpy(x, c->hstage, c->esz * (size_t)c->n);
        return 0;
    }
    TRY(hipMemcpyAsync(x, c->V + (size_t)col * c->ld * c->esz, c->esz * (size_t)c->n,
                       hipMemcpyDeviceToHost, c->s));
    return ctx_wait(c);
}

int trlhip_cols_get(trlhip_ctx *c, int c0, int k, void *x, long ldx)
{
    if (!c || !x || k < 1 || !valid_cols(c, c0, k) || ldx < c->n) return TRLHIP_EINVAL;
    if (int rc = settle(c)) return rc;
    const size_t colb = c->esz * (size_t)c->ld;  // a device column with its padding
    const size_t bytes = colb * (size_t)(k - 1) + c->esz * (size_t)c->n;
    const char *src = c->V + (size_t)c0 * colb;
    if (!c->batch) {
        TRY(hipMemcpy2DAsync(x, c->esz * (size_t)ldx, src, colb, c->esz * (size_t)c->n, k,
                             hipMemcpyDeviceToHost, c->s));
        return ctx_wait(c);
    }
    // batched: one transfer into pinned staging (stage_take), a single
    // yield, then the columns out to x
    const StageBuf st = stage_take(bytes);
    if (!st.p) return TRLHIP_ENOMEM;
    int rc = 0;
    if (hipMemcpyAsync(st.p, src, bytes, hipMemcpyDeviceToHost, c->s) != hipSuccess)
        rc = TRLHIP_EHIP;
    if (!rc) rc = ctx_wait(c);
    if (!rc)
        for (int i = 0; i < k; ++i)
            memcpy((char *)x + (size_t)i * c->esz * (size_t)ldx, st.p + (size_t)i * colb,
                   c->esz * (size_t)c->n);
    stage_give(st);
    return rc;
}

int trlhip_col_set_device_f32(trlhip_ctx *c, int col, const float *dsrc)
{
    if (!c || !dsrc || !valid_col(c, col) || c->E != 1) return TRLHIP_EINVAL;
    if (int rc = flush_pending(c)) return rc;
    if (c->batch) {  // recorded: the flush copies it with the batch's other copies
        BatchOp op{};
        op.c = c;
        op.type = OP_COPY;
        op.dst = col;
        op.src = col;
        op.s = 1.0;
        op.src32 = dsrc;
        std::lock_guard<SpinLock> lk(c->batch->mu);
        op.phase = c->nrec++;
        c->batch->ops.push_back(op);
        return 0;
    }
    const int g = grid_for(c->n, kBlock * 8, kMaxGrid);
    KLAUNCH(k_copy_f32, dim3(g), dim3(kBlock), 0, c->s, (long)c->n, dsrc, colp<double>(c, col),
            shadow32(c, col));
    return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
}

int trlhip_col_copy(trlhip_ctx *c, int dst, int src)
{
    if (!c || !valid_col(c, dst) || !valid_col(c, src)) return TRLHIP_EINVAL;
    if (int rc = flush_pending(c)) return rc;
    if (dst == src) return 0;
    if (c->E == 1) {  // deferred like a scale-copy from the scratch column: the
                      // next SpMV of dst reads src and stores the copy (exact)
        c->pend_dst = dst;
        c->pend_s = 1.0;
        c->pend_src = src;
        return 0;
    }
    return run_scale_copy(c, dst, src, 1.0);
}

int trlhip_col_scale(trlhip_ctx *c, int col, double s)
{
    if (!c || !valid_col(c, col)) return TRLHIP_EINVAL;
    if (int rc = flush_pending(c)) return rc;
    return run_scale_copy(c, col, col, s);
}

int trlhip_col_scale_copy(trlhip_ctx *c, int dst, int src, double s)
{
    if (!c || !valid_col(c, dst) || src < 0 || src > c->ncols) return TRLHIP_EINVAL;
    if (chain_live(c)) {
        // the follower's fused SpMV stored exactly this normalisation of the
        // scratch column into dst (same 1/beta, bit for bit), so it is done
        const double *slot = c->hres + hres_chain(c->ncols) + 7 * (size_t)(c->ch_next - 1);
        const int k = c->ch_next;
        if (c->ch_copy_ok || src != c->ncols || dst != c->ch_qa[k] ||
            memcmp(&s, &slot[5], sizeof(double)) != 0)
            return TRLHIP_EINVAL;
        c->ch_copy_ok = 1;
        return 0;
    }
    if (int rc = flush_pending(c)) return rc;
    if (src == c->ncols) {  // deferred: the next SpMV of column dst applies it
        c->pend_dst = dst;
        c->pend_s = s;
        c->pend_src = src;
        return 0;
    }
    return run_scale_copy(c, dst, src, s);
}

int trlhip_spmv(trlhip_ctx *c, const trlhip_csr *A, int xcol, int ycol, double *xdoty)
{
    if (!c || !csr_fits(c, A)) return TRLHIP_EINVAL;
    if (!valid_col(c, xcol) || !valid_col(c, ycol) || xcol == ycol) return TRLHIP_EINVAL;
    int rc = settle(c);
    if (!rc) rc = run_spmv(c, A, xcol, ycol);
    if (!rc) rc = launch_sum(part_alpha(c), c->np_alpha, c->dscal + S_ALPHA, c->s);
    if (!rc) rc = fetch(c, c->dscal + S_ALPHA, c->hbuf, 1);
    if (rc) return rc;
    if (xdoty) *xdoty = c->hbuf[0];
    return 0;
}

// y = A q_a (if A) or the host-written column, alpha = Re(q_a^H y), and
// r = y - coef*V over columns [c0, c0+m) with alpha as the last
// coefficient (filled on the device from the partials).  ||r||^2 ->
// dscal[S_RR2] (defer: its partials stay in part_recur, np_recur of them),
// alpha -> dscal[S_ALPHA].
// The recurrence's columns and coefficients, for a caller that runs it
// fused with the projection (form_residual with `rec`: SpMV only).
struct RecArgs {
    int c0, m;
    const double *coef;
    double h0;
};

static int form_residual(trlhip_ctx *c, const trlhip_csr *A, int qa, int first, int qb,
                         double beta_prev, const double *beta, int jnd, int rcol,
                         bool defer = false, RecArgs *rec = nullptr)
{
    const int E = c->E;
    int c0, m;
    const double *coef = c->dcoef;
    double h0 = 0.0;
    if (first) {
        double *h = c->hbuf + 16;
        for (int i = 0; i < jnd - 1; ++i) {
            h[E * i] = beta[i];
            if (E == 2) h[E * i + 1] = 0.0;
        }
        if (jnd > 1)
            TRY(hipMemcpyAsync(c->dcoef, h, sizeof(double) * E * (size_t)(jnd - 1),
                               hipMemcpyHostToDevice, c->s));
        c0 = 0;
        m = jnd;
    } else if (qb >= 0) {
        coef = nullptr;  // (beta_prev, alpha) passed by value
        h0 = beta_prev;
        c0 = qb;
        m = 2;
    } else {
        c0 = qa;
        m = 1;
    }
    int rc;
    if (A && c->pend_dst == qa && rcol != c->pend_src) {
        rc = run_spmv_fused(c, A, qa, rcol);
    } else {
        rc = settle(c);
        if (!rc) rc = A ? run_spmv(c, A, qa, rcol) : run_dot(c, qa, rcol);
    }
    if (rc) return rc;
    if (rec) {
        *rec = RecArgs{c0, m, coef, h0};
        return 0;
    }
    if (defer)
        return run_update(c, c0, m, coef, m - 1, rcol, rcol, nullptr, nullptr, TRLHIP_K_RECUR, h0,
                          part_recur(c), &c->np_recur);
    return run_update(c, c0, m, coef, m - 1, rcol, rcol, c->dscal + S_RR2, nullptr,
                      TRLHIP_K_RECUR, h0);
}

int trlhip_lanczos_step(trlhip_ctx *c, const trlhip_csr *A, int qa, int qb, double beta_prev,
                        int rcol, double *alpha, double *rr2)
{
    if (!c || !valid_col(c, qa) || !valid_col(c, rcol) || qa == rcol) return TRLHIP_EINVAL;
    if (qb >= 0 && qb != qa - 1) return TRLHIP_EINVAL;
    if (A && !csr_fits(c, A)) return TRLHIP_EINVAL;
    if (int rc = settle(c)) return rc;
    int rc = form_residual(c, A, qa, 0, qb, beta_prev, nullptr, 0, rcol);
    if (rc) return rc;
    rc = fetch(c, c->dscal, c->hbuf + 4, 2);
    if (rc) return rc;
    if (alpha) *alpha = c->hbuf[4 + S_ALPHA];
    if (rr2) *rr2 = c->hbuf[4 + S_RR2];
    return 0;
}

int trlhip_restart_step(trlhip_ctx *c, const trlhip_csr *A, int qa, const double *beta, int jnd,
                        int rcol, double *alpha, double *rr2)
{
    if (!c || jnd < 1 || jnd > c->ncols || qa != jnd - 1 || !valid_col(c, rcol) || rcol == qa)
        return TRLHIP_EINVAL;
    if (A && !csr_fits(c, A)) return TRLHIP_EINVAL;
    if (int rc = settle(c)) return rc;
    int rc = form_residual(c, A, qa, 1, -1, 0.0, beta, jnd, rcol);
    if (rc) return rc;
    rc = fetch(c, c->dscal, c->hbuf + 4, 2);
    if (rc) return rc;
    if (alpha) *alpha = c->hbuf[4 + S_ALPHA];
    if (rr2) *rr2 = c->hbuf[4 + S_RR2];
    return 0;
}

int trlhip_step_cgs(trlhip_ctx *c, const trlhip_csr *A, int first, int qa, int qb,
                    double beta_prev, const double *beta, int jnd, int rcol, double out[4])
{
    if (!c || !out || !valid_col(c, qa) || !valid_col(c, rcol) || qa == rcol) return TRLHIP_EINVAL;
    if (first && (jnd < 1 || qa != jnd - 1 || !beta)) return TRLHIP_EINVAL;
    if (!first && qb >= 0 && qb != qa - 1) return TRLHIP_EINVAL;
    if (jnd < 1 || jnd > c->ncols || (rcol >= 0 && rcol < jnd)) return TRLHIP_EINVAL;
    if (A && !csr_fits(c, A)) return TRLHIP_EINVAL;
    if (c->ch_next > 0) {
        // the step a follower already ran (it must be exactly that step)
        const int k = c->ch_next;
        const double *slot = c->hres + hres_chain(c->ncols) + 7 * (size_t)(k - 1);
        if (slot[4] == 1.0) {
            if (!c->ch_copy_ok || first || qa != c->ch_qa[k] || qb != c->ch_qb[k] ||
                rcol != c->ch_rcol[k] || jnd != c->ch_jnd[k] ||
                memcmp(&beta_prev, &slot[6], sizeof(double)) != 0)
                return TRLHIP_EINVAL;
            out[0] = slot[S_ALPHA];
            out[1] = slot[S_RR2];
            out[2] = slot[S_HH];
            out[3] = slot[S_RR2B];
            c->ch_copy_ok = 0;
            c->ch_next = k + 1 < c->ch_n ? k + 1 : 0;
            return 0;
        }
        c->ch_next = 0;  // the chain stopped before this step: record it
    }
    if (c->batch && A && c->E == 1) {
        // batched: record the step; trlhip_batch_flush runs it with the
        // other contexts' steps
        BatchStep st{c, A, first, qa, qb, rcol, jnd, beta_prev, 0, 1.0, -1, 0, 0, 0, 0};
        // followers the solver allows (trlhip_step_chain): steps jnd + 1 ...
        // on consecutive columns, each run only if the device decides the
        // solver would accept the step before it as it is
        int K = first ? 1 : std::min(c->ch_budget, kChainMax);
        while (K > 1 && (!valid_col(c, rcol + K - 1) || jnd + K - 1 > c->ncols)) --K;
        c->ch_budget = 0;
        if (K > 1) {
            st.decide = 1;
            st.kept = c->ch_kept;
            st.ntot = c->ch_ntot;
        }
        if (c->pend_dst == qa && rcol != c->pend_src) {
            st.fuse = 1;
            st.fuse_s = c->pend_s;
            st.fuse_src = c->pend_src;
            c->pend_dst = -1;
        } else if (int rc = flush_pending(c)) {
            return rc;
        }
        {
            std::lock_guard<SpinLock> lk(c->batch->mu);
            if (first && jnd > 1) {  // coefficients go with the flush's staging, with
                                     // a slot for alpha (the recurrence reads jnd)
                st.coff = (long)c->batch->cbuf.size();
                c->batch->cbuf.insert(c->batch->cbuf.end(), beta, beta + (jnd - 1));
                c->batch->cbuf.push_back(0.0);
            }
            c->batch->steps.push_back(st);
            for (int k = 1; k < K; ++k) {
                BatchStep f{c, A, 0, rcol + k - 1, k >= 2 ? rcol + k - 2 : qa, rcol + k, jnd + k, 0.0,
                            1, 1.0, -1, c->ncols, k, k + 1 < K, c->ch_kept, c->ch_ntot};
                c->ch_qa[k] = f.qa;
                c->ch_qb[k] = f.qb;
                c->ch_rcol[k] = f.rcol;
                c->ch_jnd[k] = f.jnd;
                c->hres[hres_chain(c->ncols) + 7 * (size_t)(k - 1) + 4] = 0.0;  // set if it runs
                c->batch->steps.push_back(f);
            }
        }
        c->ch_n = K;
        c->ch_next = K > 1 ? 1 : 0;
        c->ch_copy_ok = 0;
        if (int rc = ctx_wait(c)) return rc;
        const volatile double *hr = c->hres;
        out[0] = hr[S_ALPHA];
        out[1] = hr[S_RR2];
        out[2] = hr[S_HH];
        out[3] = hr[S_RR2B];
        return 0;
    }
    if (int rc = batch_drain(c)) return rc;
    int rc;
    if (recproj_ok(c)) {
        // SpMV, then the recurrence fused with the first projection
        RecArgs ra{};
        rc = form_residual(c, A, qa, first, qb, beta_prev, beta, jnd, rcol, true, &ra);
        if (rc) return rc;
        const int gx = grid_for(c->n, kMinRowsUpd, kMaxGrid);
        {
            const bool s32 = c->V32 != nullptr;  // mixed: the FP32 shadow, else the FP64 basis
            Timed t(TRLHIP_K_PROJ, c->s, 8.0 * c->n * (ra.m + 2) + (s32 ? 4.0 : 8.0) * jnd * c->n);
            const int cb = g_proj_cb[s32 ? 1 : 0].load(std::memory_order_relaxed);
#define RECPROJ_LAUNCH(CB, VS, VG)                                                              \
    KLAUNCH((k_recproj<CB, VS>), dim3(gx), dim3(kBlock), sizeof(double) * ra.m, c->s, c->n,      \
            colp<double>(c, 0), (long)c->ld, ra.c0, ra.m, ra.coef, ra.h0, part_alpha(c),          \
            c->np_alpha, colp<double>(c, rcol), part_recur(c), c->dscal + S_ALPHA, VG, jnd,       \
            part_proj(c))
#define RECPROJ_CB(CB)                                                                          \
    case CB:                                                                                    \
        if (s32)                                                                                \
            RECPROJ_LAUNCH(CB, float, c->V32);                                                  \
        else                                                                                    \
            RECPROJ_LAUNCH(CB, double, colp<double>(c, 0));                                     \
        break;
            switch (cb) {
                RECPROJ_CB(4)
                RECPROJ_CB(16)
                default: RECPROJ_CB(8)
            }
#undef RECPROJ_CB
#undef RECPROJ_LAUNCH
            if (hipGetLastError() != hipSuccess) return TRLHIP_EHIP;
        }
        c->np_recur = gx;
        KLAUNCH(k_colsum, dim3(jnd), dim3(kBlock), 0, c->s, part_proj(c), gx, c->dh);
        if (hipGetLastError() != hipSuccess) return TRLHIP_EHIP;
    } else {
        rc = form_residual(c, A, qa, first, qb, beta_prev, beta, jnd, rcol, true);
        if (rc) return rc;
        // speculative first Gram-Schmidt pass into the scratch column
        rc = run_proj(c, 0, jnd, rcol, c->dh, true);
        if (rc) return rc;
    }
    int g2 = 0;
    rc = run_update(c, 0, jnd, c->dh, -1, rcol, c->ncols, nullptr, c->dscal + S_HH,
                    TRLHIP_K_UPDATE, 0.0, nullptr, &g2, true);
    if (rc) return rc;
    // both norms (before and after the pass) in one finalize launch, which
    // also writes the four results to host-coherent memory
    KLAUNCH(k_sum2, dim3(2), dim3(kBlock), 0, c->s, part_recur(c), c->np_recur,
            c->dscal + S_RR2, part_update(c), g2, c->dscal + S_RR2B, c->dscal, c->dres);
    if (hipGetLastError() != hipSuccess) return TRLHIP_EHIP;
    if (int rc2 = ctx_wait(c)) return rc2;
    const volatile double *hr = c->hres;
    out[0] = hr[S_ALPHA];
    out[1] = hr[S_RR2];
    out[2] = hr[S_HH];
    out[3] = hr[S_RR2B];
    return 0;
}

int trlhip_step_chain(trlhip_ctx *c, int steps, int kept, int ntot)
{
    if (!c) return TRLHIP_EINVAL;
    if (!c->batch || c->E != 1) return 0;  // chains are a batched-step device layout
    c->ch_budget = steps;
    c->ch_kept = kept;
    c->ch_ntot = ntot;
    return 0;
}

int trlhip_cgs_pass(trlhip_ctx *c, int c0, int m, int rcol, double *hh, double *rr2, double *h_out)
{
    if (!c || m < 1 || !valid_cols(c, c0, m) || !valid_col(c, rcol) || (rcol >= c0 && rcol < c0 + m))
        return TRLHIP_EINVAL;
    int rc = settle(c);
    if (!rc) rc = run_proj(c, c0, m, rcol, c->dh, true);
    if (rc) return rc;
    rc = run_update(c, c0, m, c->dh, -1, rcol, rcol, c->dscal + S_RR2, c->dscal + S_HH,
                    TRLHIP_K_UPDATE, 0.0, nullptr, nullptr, true);
    if (rc) return rc;
    const int E = c->E;
    if (h_out)
        TRY(hipMemcpyAsync(c->hbuf + 16, c->dh, sizeof(double) * E * (size_t)m, hipMemcpyDeviceToHost, c->s));
    rc = fetch(c, c->dscal, c->hbuf + 4, 3);
    if (rc) return rc;
    if (hh) *hh = c->hbuf[4 + S_HH];
    if (rr2) *rr2 = c->hbuf[4 + S_RR2];
    if (h_out) memcpy(h_out, c->hbuf + 16, sizeof(double) * E * (size_t)m);
    return 0;
}

int trlhip_proj(trlhip_ctx *c, int c0, int m, int rcol, double *h)
{
    if (!c || !h || m < 1 || !valid_cols(c, c0, m) || !valid_col(c, rcol)) return TRLHIP_EINVAL;
    int rc = settle(c);
    if (!rc) rc = run_proj(c, c0, m, rcol, c->dh);
    if (rc) return rc;
    rc = fetch(c, c->dh, c->hbuf + 16, (size_t)c->E * m);
    if (rc) return rc;
    memcpy(h, c->hbuf + 16, sizeof(double) * c->E * (size_t)m);
    return 0;
}

int trlhip_update(trlhip_ctx *c, int c0, int m, const double *h, int rcol, double *rr2)
{
    if (!c || !h || m < 0 || !valid_cols(c, c0, m) || !valid_col(c, rcol) ||
        (m > 0 && rcol >= c0 && rcol < c0 + m))
        return TRLHIP_EINVAL;
    if (c->batch && c->E == 1 && m == 0) {
        // a column norm in a batch (a start vector's): recorded after the
        // context's other records, run with every context's norms at the
        // flush (kb_norm, the same sums as below)
        if (chain_live(c)) return TRLHIP_EINVAL;
        if (int rc = flush_pending(c)) return rc;
        BatchOp op{};
        op.c = c;
        op.type = OP_NORM;
        op.dst = rcol;
        {
            std::lock_guard<SpinLock> lk(c->batch->mu);
            op.phase = c->nrec++;
            c->batch->ops.push_back(op);
        }
        if (int rc = ctx_wait(c)) return rc;
        c->hbuf[4] = c->hres[kHresNorm];
        if (rr2) *rr2 = c->hbuf[4];
        return 0;
    }
    if (int rc = settle(c)) return rc;
    const int E = c->E;
    memcpy(c->hbuf + 16, h, sizeof(double) * E * (size_t)m);
    if (m > 0)
        TRY(hipMemcpyAsync(c->dh, c->hbuf + 16, sizeof(double) * E * (size_t)m, hipMemcpyHostToDevice, c->s));
    int rc = run_update(c, c0, m, c->dh, -1, rcol, rcol, c->dscal + S_RR2, nullptr, TRLHIP_K_UPDATE);
    if (rc) return rc;
    rc = fetch(c, c->dscal + S_RR2, c->hbuf + 4, 1);
    if (rc) return rc;
    if (rr2) *rr2 = c->hbuf[4];
    return 0;
}

int trlhip_residuals(trlhip_ctx *c, const trlhip_csr *A, int k, const double *rho, double *rr2)
{
    if (!c || !rho || !rr2 || k < 1 || c->E != 1 || !csr_fits(c, A) || 2 * k > c->ncols)
        return TRLHIP_EINVAL;
    if (int rc = settle(c)) return rc;
    // rho staged once (hbuf holds 16 + 2 (ncols + 2) doubles); each norm's
    // fixed-order sum goes straight to the host-coherent results (the
    // refinement's slots, free once trlhip_refine has returned)
    memcpy(c->hbuf + 16, rho, sizeof(double) * (size_t)k);
    TRY(hipMemcpyAsync(c->dh, c->hbuf + 16, sizeof(double) * (size_t)k, hipMemcpyHostToDevice, c->s));
    for (int i = 0; i < k; ++i) {
        if (int rc = run_spmv(c, A, i, k + i)) return rc;
        if (int rc = run_update(c, i, 1, c->dh + i, -1, k + i, k + i, c->dres + 8 + i, nullptr,
                                TRLHIP_K_UPDATE))
            return rc;
    }
    if (int rc = ctx_wait(c)) return rc;
    const volatile double *hr = c->hres + 8;
    for (int i = 0; i < k; ++i) rr2[i] = hr[i];
    return 0;
}

// G = W^T W for the first m <= kGramMaxM FP64 basis columns W, in one
// pass: block b takes the 64-row tiles b, b + gridDim.x, ... (staged in
// LDS); thread t owns entries e = t + q * kBlock < m^2 (G(i, j) at
// e = j m + i) and accumulates w_i[r] * w_j[r] over its tiles' rows in
// order; partial of entry e at part[e * gridDim.x + b], summed in block
// order by k_colsum.  G(i, j) and G(j, i) are the same products in the same
// order, so G is exactly symmetric.
constexpr int kGramMaxM = 32, kGramRows = 64, kGramBlocks = 256;

__global__ __launch_bounds__(kBlock) void k_gram(int n, const double *__restrict__ V, long ld, int m,
                                                 double *__restrict__ part)
{
    constexpr int EPT = kGramMaxM * kGramMaxM / kBlock;
    __shared__ double w[kGramRows][kGramMaxM + 1];
    const int t = threadIdx.x, mm = m * m;
    double acc[EPT];
    int ei[EPT], ej[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = t + q * kBlock;
        acc[q] = 0.0;
        ei[q] = e < mm ? e % m : 0;
        ej[q] = e < mm ? e / m : 0;
    }
    for (int r0 = blockIdx.x * kGramRows; r0 < n; r0 += gridDim.x * kGramRows) {
        const int rows = min(kGramRows, n - r0);
        __syncthreads();  // the previous tile has been read
        for (int idx = t; idx < kGramRows * m; idx += kBlock) {
            const int col = idx / kGramRows, row = idx % kGramRows;
            if (row < rows) w[row][col] = V[(size_t)col * ld + r0 + row];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < EPT; ++q)
            if (t + q * kBlock < mm)
                for (int r = 0; r < rows; ++r) acc[q] = fma(w[r][ei[q]], w[r][ej[q]], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = t + q * kBlock;
        if (e < mm) part[(size_t)e * gridDim.x + blockIdx.x] = acc[q];
    }
}

int trlhip_gram(trlhip_ctx *c, const trlhip_csr *A, int m, int tmp, double *G, double *H)
{
    if (!c || !G || !H || m < 1 || c->E != 1 || !csr_fits(c, A) || !valid_cols(c, 0, m) ||
        !valid_col(c, tmp) || tmp < m)
        return TRLHIP_EINVAL;
    if (int rc = settle(c)) return rc;
    const size_t mm = (size_t)m * m;
    if (c->dYcap < 2 * mm) {
        if (c->arena) return TRLHIP_EINVAL;  // sized for 2 ncols^2 at creation
        if (c->dY) (void)hipFree(c->dY);
        if (c->hY) (void)hipHostFree(c->hY);
        c->dY = nullptr;
        c->hY = nullptr;
        c->dYcap = 0;
        if (hipMalloc((void **)&c->dY, sizeof(double) * 2 * mm) != hipSuccess) return TRLHIP_ENOMEM;
        if (hipHostMalloc((void **)&c->hY, sizeof(double) * 2 * mm, 0) != hipSuccess) return TRLHIP_ENOMEM;
        c->dYcap = 2 * mm;
    }
    // G in one pass over W (k_gram + k_colsum) when m <= kGramMaxM, else
    // column by column with the projection kernels; column j of H = W^T A w_j
    // by the projection kernels; both matrices back in one copy
    const size_t avail = c->partcap - ((size_t)kSpmvGrid + 2 * (size_t)kMaxGrid);
    const int gb = (int)std::min<size_t>({(size_t)(c->n + kGramRows - 1) / kGramRows,
                                          avail / mm, (size_t)kGramBlocks});
    const bool one_pass = m <= kGramMaxM && gb >= 1;
    if (one_pass) {
        KLAUNCH(k_gram, dim3(gb), dim3(kBlock), 0, c->s, c->n, colp<double>(c, 0), (long)c->ld, m,
                part_proj(c));
        KLAUNCH(k_colsum, dim3((unsigned)mm), dim3(kBlock), 0, c->s, part_proj(c), gb, c->dY);
        if (hipGetLastError() != hipSuccess) return TRLHIP_EHIP;
    }
    for (int j = 0; j < m; ++j) {
        if (!one_pass)
            if (int rc = run_proj(c, 0, m, j, c->dY + (size_t)j * m)) return rc;
        if (int rc = run_spmv(c, A, j, tmp)) return rc;
        if (int rc = run_proj(c, 0, m, tmp, c->dY + mm + (size_t)j * m)) return rc;
    }
    TRY(hipMemcpyAsync(c->hY, c->dY, sizeof(double) * 2 * mm, hipMemcpyDeviceToHost, c->s));
    if (int rc = ctx_wait(c)) return rc;
    memcpy(G, c->hY, sizeof(double) * mm);
    memcpy(H, c->hY + mm, sizeof(double) * mm);
    return 0;
}

int trlhip_ritz(trlhip_ctx *c, int lck, int jl, int k, const double *Y, int ldy)
{
    if (!c || !Y || lck < 0 || jl < 1 || k < 1 || k > jl || ldy < jl || lck + jl > c->ncols)
        return TRLHIP_EINVAL;
    if (int rc = flush_pending(c)) return rc;
    const int KT = (k + 15) / 16;
    if (KT > 8) return TRLHIP_EINVAL;  // k <= 128
    const size_t ny = (size_t)ldy * k;
    const int jl4 = (jl + 3) & ~3;
    const size_t shm = sizeof(double) * (size_t)jl4 * 16 * KT;
    if (shm > 160 * 1024) return TRLHIP_EINVAL;  // gfx950: 160 KB of LDS per CU
    if (c->batch && c->E == 1) {
        // batched: recorded with Y; the flush runs it with the other
        // contexts' restart GEMMs of the same phase
        BatchOp op{};
        op.c = c;
        op.type = OP_RITZ;
        op.lck = lck;
        op.jl = jl;
        op.k = k;
        op.ldy = ldy;
        std::lock_guard<SpinLock> lk(c->batch->mu);
        op.yoff = c->batch->ybuf.size();
        c->batch->ybuf.insert(c->batch->ybuf.end(), Y, Y + ny);
        op.phase = c->nrec++;
        c->batch->ops.push_back(op);
        return 0;
    }
    if (int rc = batch_drain(c)) return rc;
    if (c->dYcap < ny) {
        if (c->arena) return TRLHIP_EINVAL;  // sized for the largest restart at creation
        if (c->dY) (void)hipFree(c->dY);
        if (c->hY) (void)hipHostFree(c->hY);
        c->dY = nullptr;
        c->hY = nullptr;
        c->dYcap = 0;
        if (hipMalloc((void **)&c->dY, sizeof(double) * ny) != hipSuccess) return TRLHIP_ENOMEM;
        if (hipHostMalloc((void **)&c->hY, sizeof(double) * ny, 0) != hipSuccess) return TRLHIP_ENOMEM;
        c->dYcap = ny;
    }
    memcpy(c->hY, Y, sizeof(double) * ny);
    TRY(hipMemcpyAsync(c->dY, c->hY, sizeof(double) * ny, hipMemcpyHostToDevice, c->s));
    // complex columns multiply as 2n-row real matrices (Y is real)
    const long nrows = (long)c->n * c->E;
    const long ldr = (long)c->ld * c->E;
    // big single operators have parallelism to spare: 4 tiles per wave
    // amortise the LDS staging of Y (about 2.1 vs 1.9 TB/s at n = 2^21)
    const int tpw = nrows >= (1L << 20) ? 4 : kRitzTiles;
    const long rows_per_block = 16L * tpw * (kBlock / 64);
    const int g = (int)((nrows + rows_per_block - 1) / rows_per_block);
    Timed t(TRLHIP_K_RITZ, c->s,
            (double)c->esz * c->n * (jl + k) + (c->V32 ? 4.0 * c->n * k : 0.0),
            2.0 * (double)nrows * jl * k);
#define RITZ_CASE(K)                                                                          \
    case K:                                                                                   \
        KLAUNCH((k_ritz_mfma<K>), dim3(g), dim3(kBlock), shm, c->s, nrows,                     \
                reinterpret_cast<double *>(c->V), ldr, lck, jl, k, c->dY, ldy, tpw, c->V32);    \
        break;
    switch (KT) {
        RITZ_CASE(1) RITZ_CASE(2) RITZ_CASE(3) RITZ_CASE(4)
        RITZ_CASE(5) RITZ_CASE(6) RITZ_CASE(7) RITZ_CASE(8)
    }
#undef RITZ_CASE
    if (hipGetLastError() != hipSuccess) return TRLHIP_EHIP;
    // the pinned staging of Y is reused by the next call: wait for the copy
    return ctx_wait(c);
}

int trlhip_refine(trlhip_ctx *c, const trlhip_csr *A, int k, int avoid, double *rho, double *res)
{
    if (!c || !rho || !res || k < 1 || k > c->ncols || c->E != 1 || !csr_fits(c, A))
        return TRLHIP_EINVAL;
    if (refine_width(c, k, avoid) < 1) return TRLHIP_EINVAL;
    if (int rc = flush_pending(c)) return rc;
    BatchOp op{};
    op.c = c;
    op.type = OP_REFINE;
    op.k = k;
    op.A = A;
    op.avoid = avoid;
    if (c->batch) {  // recorded; the batch flush runs it with the other contexts'
        std::lock_guard<SpinLock> lk(c->batch->mu);
        op.phase = c->nrec++;
        c->batch->ops.push_back(op);
    } else {
        const std::vector<BatchOp> one{op};
        const std::vector<double> none;
        if (int rc = ops_run(one, none, c->hops, c->dops, c->ocap, c->s)) return rc;
    }
    if (int rc = ctx_wait(c)) return rc;
    const volatile double *hr = c->hres + 8;
    for (int i = 0; i < k; ++i) {
        const bool ok = hr[3 * i + 2] > 0.0;  // v'v
        rho[i] = ok ? hr[3 * i] : __builtin_nan("");
        res[i] = ok ? hr[3 * i + 1] : __builtin_nan("");
    }
    return 0;
}

void trlhip_stats_enable(int on)
{
    stats_resolve();
    g_stats.on = on > 0 ? on : 0;
    for (int k = 0; k < TRLHIP_K_COUNT; ++k) {
        g_stats.seen[k] = 0;
        g_stats.launches[k] = 0;
        g_stats.ms[k] = 0;
        g_stats.bytes[k] = 0;
        g_stats.flops[k] = 0;
    }
}

int trlhip_stats_flops(double *out, int nout)
{
    stats_resolve();
    for (int k = 0; k < TRLHIP_K_COUNT && k < nout; ++k) out[k] = g_stats.flops[k];
    return TRLHIP_K_COUNT;
}

int trlhip_stats_get(double *out, int nout)
{
    stats_resolve();
    for (int k = 0; k < TRLHIP_K_COUNT && 3 * k + 2 < nout; ++k) {
        out[3 * k] = g_stats.launches[k];
        out[3 * k + 1] = g_stats.ms[k];
        out[3 * k + 2] = g_stats.bytes[k];
    }
    return TRLHIP_K_COUNT;
}

int trlhip_batch_create(trlhip_batch **out) { return trlhip_batch_create_on(out, nullptr); }

int trlhip_batch_create_on(trlhip_batch **out, trlhip_batch *share)
{
    if (!out) return TRLHIP_EINVAL;
    if (trlhip_device_count() <= 0) return TRLHIP_ENODEV;
    trlhip_batch *b = new (std::nothrow) trlhip_batch();
    if (!b) return TRLHIP_ENOMEM;
    if (share) {
        b->s = share->s;
        b->own_stream = false;
    } else if (hipStreamCreate(&b->s) != hipSuccess) {
        delete b;
        return TRLHIP_EHIP;
    }
    if (hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->side_done, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreate(&b->side) != hipSuccess) {
        if (b->done) (void)hipEventDestroy(b->done);
        if (b->side_done) (void)hipEventDestroy(b->side_done);
        if (b->own_stream) (void)hipStreamDestroy(b->s);
        delete b;
        return TRLHIP_EHIP;
    }
    *out = b;
    return 0;
}

void trlhip_batch_destroy(trlhip_batch *b)
{
    if (!b) return;
    (void)hipStreamSynchronize(b->side);
    (void)hipStreamSynchronize(b->s);
    {
        std::lock_guard<std::mutex> lk(g_chunk_mu);
        if (b->hstg && b->dstg) g_stage_pairs.push_back(StagePair{b->hstg, b->dstg, b->scap});
        else {
            if (b->hstg) (void)hipHostFree(b->hstg);
            if (b->dstg) (void)hipFree(b->dstg);
        }
        for (auto &ch : b->dev) g_free_chunks[0].push_back(ch);
        for (auto &ch : b->host) g_free_chunks[1].push_back(ch);
    }
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->side_done) (void)hipEventDestroy(b->side_done);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->own_stream) (void)hipStreamDestroy(b->s);
    delete b;
}

void trlhip_batch_trim(void)
{
    {
        std::lock_guard<std::mutex> lk(g_chunk_mu);
        for (auto &ch : g_free_chunks[0]) (void)hipFree(ch.p);
        for (auto &ch : g_free_chunks[1]) (void)hipHostFree(ch.p);
        g_free_chunks[0].clear();
        g_free_chunks[1].clear();
        for (auto &sp : g_stage_pairs) {
            (void)hipHostFree(sp.h);
            (void)hipFree(sp.d);
        }
        g_stage_pairs.clear();
    }
    std::lock_guard<std::mutex> lk(g_stage_mu);
    for (auto &b : g_stage_free) (void)hipHostFree(b.p);
    g_stage_free.clear();
}

void trlhip_batch_set_yield(trlhip_batch *b, void (*yield)(void *), void *user)
{
    if (!b) return;
    b->yield = yield;
    b->user = user;
}

int trlhip_batch_launch(trlhip_batch *b)
{
    if (!b) return TRLHIP_EINVAL;
    int rc = b->err;
    // the contexts' immediate work of this round comes first (nothing to
    // order when the side stream has already drained, as it usually has)
    if (!rc) {
        const hipError_t q = hipStreamQuery(b->side);
        if (q == hipErrorNotReady) {
            if (hipEventRecord(b->side_done, b->side) != hipSuccess ||
                hipStreamWaitEvent(b->s, b->side_done, 0) != hipSuccess)
                rc = TRLHIP_EHIP;
        } else if (q != hipSuccess) {
            rc = TRLHIP_EHIP;
        }
    }
    // the recorded operations and steps share one staging block and one
    // transfer; the operations run first
    if (!rc && (!b->ops.empty() || !b->steps.empty())) {
        OpsPlan op;
        StepsPlan sp;
        size_t o_bytes = 0, total = 0;
        if (!b->ops.empty()) {
            ops_plan(b->ops, b->ybuf, op);
            o_bytes = (op.bytes + 255) & ~(size_t)255;
        }
        if (!b->steps.empty()) steps_plan(b, sp);
        total = o_bytes + sp.total;
        if (total > b->scap) {
            // a kept staging pair that fits (the largest), else a new one
            // (this lane's previous launch has finished: its staging is free)
            StagePair got{nullptr, nullptr, 0};
            {
                std::lock_guard<std::mutex> lk(g_chunk_mu);
                size_t best = g_stage_pairs.size();
                for (size_t i = 0; i < g_stage_pairs.size(); ++i)
                    if (g_stage_pairs[i].cap >= total &&
                        (best == g_stage_pairs.size() || g_stage_pairs[i].cap > g_stage_pairs[best].cap))
                        best = i;
                if (best < g_stage_pairs.size()) {
                    got = g_stage_pairs[best];
                    g_stage_pairs.erase(g_stage_pairs.begin() + (long)best);
                }
            }
            if (b->hstg) (void)hipHostFree(b->hstg);
            if (b->dstg) (void)hipFree(b->dstg);
            b->hstg = b->dstg = nullptr;
            b->scap = 0;
            if (got.h) {
                b->hstg = got.h;
                b->dstg = got.d;
                b->scap = got.cap;
            } else {
                const size_t cap = ((total + total / 2 + 65536) + 255) & ~(size_t)255;
                if (hipHostMalloc((void **)&b->hstg, cap, 0) != hipSuccess ||
                    hipMalloc((void **)&b->dstg, cap) != hipSuccess)
                    rc = TRLHIP_ENOMEM;
                else
                    b->scap = cap;
            }
        }
        if (!rc) {
            if (!b->ops.empty()) ops_fill(op, b->ybuf, b->hstg, b->dstg);
            if (!b->steps.empty()) steps_fill(b, sp, b->hstg + o_bytes, b->dstg + o_bytes);
            if (hipMemcpyAsync(b->dstg, b->hstg, total, hipMemcpyHostToDevice, b->s) != hipSuccess)
                rc = TRLHIP_EHIP;
        }
        if (!rc && !b->ops.empty()) rc = ops_launch(op, b->hstg, b->dstg, b->s);
        if (!rc && !b->steps.empty()) rc = steps_launch(b->s, sp, b->hstg + o_bytes, b->dstg + o_bytes);
    }
    // everything this batch queued so far (this launch and the contexts'
    // own copies) is behind this event
    if (!rc && hipEventRecord(b->done, b->s) != hipSuccess) rc = TRLHIP_EHIP;
    // the records are staged; contexts stay suspended until trlhip_batch_wait
    for (const BatchOp &op : b->ops) op.c->nrec = 0;
    b->ops.clear();
    b->ybuf.clear();
    b->steps.clear();
    b->cbuf.clear();
    if (rc) b->err = rc;
    return rc;
}

int trlhip_batch_after(trlhip_batch *b, trlhip_batch *prev)
{
    if (!b || !prev) return TRLHIP_EINVAL;
    return hipStreamWaitEvent(b->s, prev->done, 0) == hipSuccess ? 0 : TRLHIP_EHIP;
}

int trlhip_batch_wait(trlhip_batch *b)
{
    if (!b) return TRLHIP_EINVAL;
    int rc = b->err;
    if (!rc && hipEventSynchronize(b->done) != hipSuccess) rc = TRLHIP_EHIP;
    if (rc) {  // after an error, drain what was queued
        (void)hipStreamSynchronize(b->side);
        (void)hipStreamSynchronize(b->s);
    }
    stats_resolve();
    if (rc) b->err = rc;
    return rc;
}

int trlhip_batch_ready(trlhip_batch *b)
{
    if (!b || b->err) return 1;
    return hipEventQuery(b->done) == hipErrorNotReady ? 0 : 1;
}

int trlhip_batch_flush(trlhip_batch *b)
{
    const int rc = trlhip_batch_launch(b);
    const int rc2 = trlhip_batch_wait(b);
    return rc ? rc : rc2;
}

}  // extern "C"

// ------------------------------------------------ allocation-free pass
//
// trlhip_cgs_pass_noalloc: one classical Gram-Schmidt pass over caller
// device memory with no allocation at all.  Its reduction partials (the
// projection's partial rows, h, the update's ||r||^2 partials and the
// three result scalars) live in `scratch`, a fixed device array that every
// GPU the library runs on holds once, for the life of the process.  Host
// code never passes `scratch` itself to a kernel (in host code the name
// is a host-side stand-in, not a device address): the device address is
// taken once per device with hipGetSymbolAddress and kept in g_scratch_dev.
// Every kernel writes each scratch element it later reads (no part of it
// needs to start at zero); calls that overlap on one GPU share the array
// and race.

namespace {
constexpr int kNoAllocMaxCols = 256;  // TRLHIP_NOALLOC_MAX_COLS
constexpr int kMaxDevices = 64;
// partial rows: kProjGrid per coefficient, 2 coefficients per complex
// column; then h (2 per column), the update partials, 8 scalars
constexpr size_t kScrProj = 0;
constexpr size_t kScrH = kScrProj + (size_t)kProjGrid * 2 * kNoAllocMaxCols;
constexpr size_t kScrUpd = kScrH + 2 * (size_t)kNoAllocMaxCols;
constexpr size_t kScrScal = kScrUpd + kMaxGrid;
constexpr size_t kScrCount = kScrScal + 8;
}  // namespace

static __device__ __align__(256) double scratch[kScrCount];

namespace {
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (taken on the first
// call on that device), or nullptr.
double *scratch_dev(void)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        double *p = nullptr;
        if (hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_scratch_dev[dev] = p;
    }
    return g_scratch_dev[dev];
}

template <typename VS, int E>
int noalloc_launch(int n, int m, const VS *V, long ld, double *r, double *scr, hipStream_t s)
{
    double *pproj = scr + kScrProj, *dh = scr + kScrH, *pupd = scr + kScrUpd, *dscal = scr + kScrScal;
    // the same kernels, tiles and grids as trlhip_cgs_pass on a context of
    // n rows, so the two agree bit for bit
    const bool s32 = std::is_same<VS, float>::value;
    const int cb = E == 2 ? 4 : g_proj_cb[s32 ? 1 : 0].load(std::memory_order_relaxed);
    const int gx = proj_gx_rows(n, m, cb, s32 ? 4 : E == 2 ? 16 : 8);
    const int gy = (m + cb - 1) / cb;
#define NA_PROJ(CB)                                                                              \
    case CB:                                                                                     \
        KLAUNCH((k_proj<VS, double, E, CB, 4>), dim3(gx * gy), dim3(kBlock), 0, s, n, V, ld, 0, m, \
                (const double *)r, pproj, gx, gy);                                               \
        break;
    if (E == 2) {
        KLAUNCH((k_proj<VS, double, E, 4, 4>), dim3(gx * gy), dim3(kBlock), 0, s, n, V, ld, 0, m,
                (const double *)r, pproj, gx, gy);
    } else {
        switch (cb) {
            NA_PROJ(4)
            NA_PROJ(16)
            default: NA_PROJ(8)
        }
    }
#undef NA_PROJ
    if (hipGetLastError() != hipSuccess) return TRLHIP_EHIP;
    KLAUNCH(k_colsum, dim3(m * E), dim3(kBlock), 0, s, pproj, gx, dh);
    const int g = grid_for(n, kMinRowsUpd, kMaxGrid);
    const size_t shm = sizeof(double) * (size_t)E * m;
    KLAUNCH((k_update<VS, double, E, 4>), dim3(g), dim3(kBlock), shm, s, n, V, ld, 0, m,
            (const double *)dh, 0.0, (const double *)nullptr, 0, -1, (const double *)r, r, pupd,
            dscal + S_HH, dscal + S_ALPHA);
    if (hipGetLastError() != hipSuccess) return TRLHIP_EHIP;
    return launch_sum(pupd, g, dscal + S_RR2, s);
}
}  // namespace

extern "C" {

int trlhip_cgs_pass_noalloc(int n, int m, int vtype, const void *V, long ld, void *r, double *hh,
                            double *rr2, double *h_out, void *stream)
{
    if (m < 1 || m > kNoAllocMaxCols) return TRLHIP_ERANGE;
    if (n < 1 || !V || !r || ld < n) return TRLHIP_EINVAL;
    if (vtype != TRLHIP_R64 && vtype != TRLHIP_C128 && vtype != TRLHIP_R32) return TRLHIP_EINVAL;
    // 16-byte aligned columns: the real kernels move four rows per 16-byte
    // load or store
    if (((uintptr_t)V & 15) || ((uintptr_t)r & 15) || (vtype != TRLHIP_C128 && (ld & 3)))
        return TRLHIP_EINVAL;
    double *scr = scratch_dev();
    if (!scr) return TRLHIP_EHIP;
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if (vtype == TRLHIP_R64)
        rc = noalloc_launch<double, 1>(n, m, (const double *)V, ld, (double *)r, scr, s);
    else if (vtype == TRLHIP_R32)
        rc = noalloc_launch<float, 1>(n, m, (const float *)V, ld, (double *)r, scr, s);
    else
        rc = noalloc_launch<double, 2>(n, m, (const double *)V, ld, (double *)r, scr, s);
    if (rc) return rc;
    const int E = vtype == TRLHIP_C128 ? 2 : 1;
    double sc[3];
    TRY(hipMemcpyAsync(sc, scr + kScrScal, sizeof sc, hipMemcpyDeviceToHost, s));
    if (h_out)
        TRY(hipMemcpyAsync(h_out, scr + kScrH, sizeof(double) * E * (size_t)m, hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    if (hh) *hh = sc[S_HH];
    if (rr2) *rr2 = sc[S_RR2];
    return 0;
}

}  // extern "C"
