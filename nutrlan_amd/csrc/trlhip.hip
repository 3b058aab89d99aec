// trlhip.hip -- CDNA4 (gfx950) device layer of the thick-restart Lanczos
// solver: device CSR operator, Lanczos basis in HBM, and the hot-path
// kernels (SpMV + alpha, recurrence, classical Gram-Schmidt projection and
// update, FP64-MFMA restart GEMM).  C ABI in include/trlhip.h.
//
// Reductions (alpha, ||r||^2, V^H r) never leave the device until the final
// scalar: each block reduces its rows (wave shuffles, then LDS) into a
// partial slot, and the last block to finish (atomic ticket) sums the
// partials in block order, so results are deterministic run to run.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "trlhip.h"
#include "trlan_internal.h"

namespace {

constexpr int kBlock = 256;             // 4 waves of 64
constexpr int kMaxGrid = 1024;          // partial slots per reduction
constexpr uint64_t kCsrMagic = 0x7472686970637372ull;  // "trhipcsr"

// ------------------------------------------------------------------ utils

#define TRY(expr)                                                   \
    do {                                                            \
        hipError_t e_ = (expr);                                     \
        if (e_ != hipSuccess) return TRLHIP_EHIP;                   \
    } while (0)

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
            const double send = hi ? v[i] : v[i + cnt / 2];
            const double keep = hi ? v[i + cnt / 2] : v[i];
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

// Release this block's partial writes, take a ticket; true in every thread
// of the block that finished last.  That block resets the ticket.
__device__ __forceinline__ bool last_block(unsigned *ticket, unsigned nblocks)
{
    __shared__ int amlast;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned t = atomicAdd(ticket, 1u);
        amlast = (t == nblocks - 1) ? 1 : 0;
        if (amlast) atomicExch(ticket, 0u);
    }
    __syncthreads();
    if (amlast) __threadfence();
    return amlast != 0;
}

// Ordered sum of `nb` scalar partials by the last block; result in thread 0.
__device__ __forceinline__ double ordered_sum(const double *part, int nb)
{
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) s += part[b];
    return block_sum(s);
}

// ---------------------------------------------------------------- kernels

// y = A x (CSR), one SW-lane subgroup per row; epilogue: Re(x^H y) reduced
// to out0[0] (and out1[0], out1[1] = 0 for a complex coefficient slot).
// E = doubles per element (1 real, 2 complex).
template <int E, int SW>
__global__ __launch_bounds__(kBlock) void k_spmv(
    int n, const int *__restrict__ rp, const int *__restrict__ ci,
    const double *__restrict__ val, const double *__restrict__ x,
    double *__restrict__ y, double *__restrict__ part,
    unsigned *__restrict__ ticket, double *__restrict__ out0,
    double *__restrict__ out1)
{
    constexpr int subs = kBlock / SW;
    const int sl = threadIdx.x % SW;
    const int sub = threadIdx.x / SW;
    double acc = 0.0;
    for (int row = blockIdx.x * subs + sub; row < n; row += gridDim.x * subs) {
        const int kb = rp[row], ke = rp[row + 1];
        if constexpr (E == 1) {
            double s = 0.0;
            for (int k = kb + sl; k < ke; k += SW) s += val[k] * x[ci[k]];
#pragma unroll
            for (int o = SW / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
            if (sl == 0) {
                y[row] = s;
                acc += x[row] * s;
            }
        } else {
            double sr = 0.0, si = 0.0;
            for (int k = kb + sl; k < ke; k += SW) {
                const double ar = val[2 * k], ai = val[2 * k + 1];
                const int c = ci[k];
                const double xr = x[2 * c], xi = x[2 * c + 1];
                sr += ar * xr - ai * xi;
                si += ar * xi + ai * xr;
            }
#pragma unroll
            for (int o = SW / 2; o > 0; o >>= 1) {
                sr += __shfl_xor(sr, o);
                si += __shfl_xor(si, o);
            }
            if (sl == 0) {
                y[2 * row] = sr;
                y[2 * row + 1] = si;
                acc += x[2 * row] * sr + x[2 * row + 1] * si;
            }
        }
    }
    const double bs = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
    if (last_block(ticket, gridDim.x)) {
        const double t = ordered_sum(part, gridDim.x);
        if (threadIdx.x == 0) {
            out0[0] = t;
            if (out1) {
                out1[0] = t;
                if (E == 2) out1[1] = 0.0;
            }
        }
    }
}

// Re(x^H y) over n elements, same outputs as k_spmv's epilogue.
template <int E>
__global__ __launch_bounds__(kBlock) void k_dot(
    int n, const double *__restrict__ x, const double *__restrict__ y,
    double *__restrict__ part, unsigned *__restrict__ ticket,
    double *__restrict__ out0, double *__restrict__ out1)
{
    double acc = 0.0;
    const long ne = (long)n * E;
    for (long i = blockIdx.x * kBlock + threadIdx.x; i < ne; i += (long)gridDim.x * kBlock)
        acc += x[i] * y[i];
    const double bs = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
    if (last_block(ticket, gridDim.x)) {
        const double t = ordered_sum(part, gridDim.x);
        if (threadIdx.x == 0) {
            out0[0] = t;
            if (out1) {
                out1[0] = t;
                if (E == 2) out1[1] = 0.0;
            }
        }
    }
}

// h(c0+cb : c0+cb+CB) = V(:, ...)^H r for the column chunk blockIdx.y.
// Each thread streams rows (grid stride), keeps CB*E running sums in
// registers; the wave transposes-and-reduces them, waves combine in LDS,
// the block writes one partial row, the last block of the chunk reduces the
// partial rows in block order into hout.
template <int E, int CB>
__global__ __launch_bounds__(kBlock) void k_proj(
    int n, const double *__restrict__ V, long ld, int c0, int m,
    const double *__restrict__ r, double *__restrict__ part,
    unsigned *__restrict__ ticket, double *__restrict__ hout)
{
    constexpr int N = CB * E;
    static_assert(N <= 64, "chunk too wide");
    __shared__ double sm[kBlock / 64][N];
    const int chunk = blockIdx.y;
    const int cb = chunk * CB;
    const int ncol = min(CB, m - cb);
    const double *Vc = V + (size_t)(c0 + cb) * ld * E;
    double acc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = 0.0;
    for (int row = blockIdx.x * kBlock + threadIdx.x; row < n; row += gridDim.x * kBlock) {
        if constexpr (E == 1) {
            const double rv = r[row];
#pragma unroll
            for (int c = 0; c < CB; ++c)
                if (c < ncol) acc[c] += Vc[(size_t)c * ld + row] * rv;
        } else {
            const double rr = r[2 * row], ri = r[2 * row + 1];
#pragma unroll
            for (int c = 0; c < CB; ++c)
                if (c < ncol) {
                    const double vr = Vc[((size_t)c * ld + row) * 2];
                    const double vi = Vc[((size_t)c * ld + row) * 2 + 1];
                    acc[2 * c] += vr * rr + vi * ri;      // conj(v) * r
                    acc[2 * c + 1] += vr * ri - vi * rr;
                }
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double ws = wave_transpose_sum<N>(acc, lane);
    constexpr int grp = 64 / N;
    if (lane % grp == 0) sm[w][lane / grp] = ws;
    __syncthreads();
    const int G = gridDim.x;
    double *prow = part + ((size_t)chunk * G + blockIdx.x) * N;
    if (threadIdx.x < N) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < kBlock / 64; ++i) t += sm[i][threadIdx.x];
        prow[threadIdx.x] = t;
    }
    if (last_block(ticket + chunk, G)) {
        // 256 threads: column index p = t % N, block phase q = t / N
        constexpr int Q = kBlock / N;
        const int p = threadIdx.x % N, q = threadIdx.x / N;
        double s = 0.0;
        for (int b = q; b < G; b += Q) s += part[((size_t)chunk * G + b) * N + p];
        __shared__ double red[kBlock];
        red[threadIdx.x] = s;
        __syncthreads();
        if (threadIdx.x < N) {
            double t = 0.0;
            for (int i = 0; i < Q; ++i) t += red[i * N + threadIdx.x];
            if (threadIdx.x < ncol * E) hout[cb * E + threadIdx.x] = t;
        }
    }
}

// r -= V(:, c0:c0+m) h ; ||r||^2 -> out_rr2 ; ||h||^2 -> out_hh (block 0).
// h (E*m doubles) is staged in LDS.
template <int E>
__global__ __launch_bounds__(kBlock) void k_update(
    int n, const double *__restrict__ V, long ld, int c0, int m,
    const double *__restrict__ h, double *__restrict__ r,
    double *__restrict__ part, unsigned *__restrict__ ticket,
    double *__restrict__ out_rr2, double *__restrict__ out_hh)
{
    extern __shared__ double hs[];
    for (int i = threadIdx.x; i < E * m; i += kBlock) hs[i] = h[i];
    __syncthreads();
    const double *Vc = V + (size_t)c0 * ld * E;
    double acc = 0.0;
    for (int row = blockIdx.x * kBlock + threadIdx.x; row < n; row += gridDim.x * kBlock) {
        if constexpr (E == 1) {
            double s = 0.0;
            int c = 0;
            for (; c + 4 <= m; c += 4) {
                const double v0 = Vc[(size_t)c * ld + row];
                const double v1 = Vc[(size_t)(c + 1) * ld + row];
                const double v2 = Vc[(size_t)(c + 2) * ld + row];
                const double v3 = Vc[(size_t)(c + 3) * ld + row];
                s += v0 * hs[c];
                s += v1 * hs[c + 1];
                s += v2 * hs[c + 2];
                s += v3 * hs[c + 3];
            }
            for (; c < m; ++c) s += Vc[(size_t)c * ld + row] * hs[c];
            const double rn = r[row] - s;
            r[row] = rn;
            acc += rn * rn;
        } else {
            double sr = 0.0, si = 0.0;
            for (int c = 0; c < m; ++c) {
                const double vr = Vc[((size_t)c * ld + row) * 2];
                const double vi = Vc[((size_t)c * ld + row) * 2 + 1];
                const double hr = hs[2 * c], hi = hs[2 * c + 1];
                sr += vr * hr - vi * hi;
                si += vr * hi + vi * hr;
            }
            const double rr = r[2 * row] - sr, ri = r[2 * row + 1] - si;
            r[2 * row] = rr;
            r[2 * row + 1] = ri;
            acc += rr * rr + ri * ri;
        }
    }
    const double bs = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
    if (out_hh && blockIdx.x == 0 && threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < E * m; ++i) t += hs[i] * hs[i];
        out_hh[0] = t;
    }
    if (last_block(ticket, gridDim.x)) {
        const double t = ordered_sum(part, gridDim.x);
        if (threadIdx.x == 0 && out_rr2) out_rr2[0] = t;
    }
}

__global__ __launch_bounds__(kBlock) void k_scale(long ne, double *__restrict__ x, double s)
{
    for (long i = blockIdx.x * kBlock + threadIdx.x; i < ne; i += (long)gridDim.x * kBlock)
        x[i] *= s;
}

// In-place restart GEMM V(:, lck:lck+k) <- V(:, lck:lck+jl) * Y on FP64
// MFMA (v_mfma_f64_16x16x4_f64).  Each wave owns 16-row tiles (rows
// disjoint across waves), reads its rows over all jl input columns into
// the accumulators before writing any output column, so the update is safe
// in place.  Y (jl x k) is staged in LDS as Ys[kk * kp + col], padded with
// zeros to jl4 = ceil4(jl) rows and kp = 16*KT columns.
// MFMA f64 16x16x4 operand layout (lane l): A[l%16][l/16], B[l/16][l%16],
// D[l/16 + 4*r][l%16] for accumulator element r (measured on gfx950).
typedef double d4 __attribute__((ext_vector_type(4)));

template <int KT>
__global__ __launch_bounds__(kBlock) void k_ritz_mfma(
    long nrows, double *__restrict__ V, long ld, int lck, int jl, int k,
    const double *__restrict__ Y, int ldy, int tiles_per_wave)
{
    extern __shared__ double Ys[];
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
        const long row0 = ((long)(blockIdx.x * (kBlock / 64) + w) * tiles_per_wave + t) * 16;
        if (row0 >= nrows) break;
        d4 acc[KT];
#pragma unroll
        for (int q = 0; q < KT; ++q) acc[q] = d4{0.0, 0.0, 0.0, 0.0};
        const long arow = row0 + li;
        for (int kk = 0; kk < jl4; kk += 4) {
            const int col = kk + lq;
            const double a = (arow < nrows && col < jl) ? Vin[(size_t)col * ld + arow] : 0.0;
#pragma unroll
            for (int q = 0; q < KT; ++q) {
                const double b = Ys[(kk + lq) * kp + q * 16 + li];
                acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
            }
        }
#pragma unroll
        for (int q = 0; q < KT; ++q) {
            const int col = q * 16 + li;
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const long row = row0 + lq + 4 * rr;
                if (row < nrows && col < k) Vin[(size_t)col * ld + row] = acc[q][rr];
            }
        }
    }
}

// ---------------------------------------------------------------- stats

struct Stats {
    int on = 0;
    double launches[TRLHIP_K_COUNT] = {};
    double ms[TRLHIP_K_COUNT] = {};
    double bytes[TRLHIP_K_COUNT] = {};
    struct Pending {
        hipEvent_t a, b;
        int k;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
};
Stats g_stats;

hipEvent_t ev_get()
{
    if (!g_stats.pool.empty()) {
        hipEvent_t e = g_stats.pool.back();
        g_stats.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Resolve timings of completed launches (called after a stream sync).
void stats_resolve()
{
    for (auto &p : g_stats.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) g_stats.ms[p.k] += ms;
        g_stats.pool.push_back(p.a);
        g_stats.pool.push_back(p.b);
    }
    g_stats.pending.clear();
}

struct Timed {
    hipEvent_t a = nullptr, b = nullptr;
    int k;
    hipStream_t s;
    Timed(int kind, hipStream_t st, double bytes) : k(kind), s(st)
    {
        if (!g_stats.on) return;
        g_stats.launches[k] += 1.0;
        g_stats.bytes[k] += bytes;
        a = ev_get();
        b = ev_get();
        if (a) (void)hipEventRecord(a, s);
    }
    ~Timed()
    {
        if (!g_stats.on || !a || !b) return;
        (void)hipEventRecord(b, s);
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
    int esz;       // doubles per value (complex 2); float storage flagged by vtype
    int sw;        // subgroup width per row
    int *rp;
    int *ci;
    double *val;
    // staging for host-called matvec
    double *dx;
    double *dy;
    double *part;
    unsigned *tick;
    double *dscal;
    hipStream_t s;
};

struct trlhip_ctx {
    int n;
    int ld;
    int ncols;
    int vtype;
    int E;          // doubles per element
    double *V;
    double *part;   // kMaxGrid * (E*ncols + 64) partial slots
    unsigned *tick; // 64 tickets
    double *dscal;  // 16 device scalars
    double *dcoef;  // 2*(ncols+2) coefficients (recurrence / restart step)
    double *dh;     // 2*(ncols+2) CGS projection
    double *dY;     // restart matrix
    size_t dYcap;
    double *hbuf;   // pinned host mirror: 16 scalars + 2*(ncols+2) + ...
    size_t hcap;
    double *hY;     // pinned staging for Y
    size_t hYcap;
    hipStream_t s;
};

namespace {

enum { S_ALPHA = 0, S_RR2 = 1, S_HH = 2 };

int grid_for(long work, int per_block_min)
{
    long g = (work + per_block_min - 1) / per_block_min;
    if (g < 1) g = 1;
    if (g > kMaxGrid) g = kMaxGrid;
    return (int)g;
}

int pick_sw(long nnz, int n)
{
    const double avg = n > 0 ? (double)nnz / n : 1.0;
    if (avg <= 2.0) return 2;
    if (avg <= 4.5) return 4;
    if (avg <= 12.0) return 8;
    if (avg <= 24.0) return 16;
    if (avg <= 48.0) return 32;
    return 64;
}

template <int E>
hipError_t launch_spmv(int sw, int n, const trlhip_csr *A, const double *x,
                       double *y, double *part, unsigned *tick, double *o0,
                       double *o1, hipStream_t s)
{
    const int subs = kBlock / sw;
    const int g = grid_for(n, subs * 8);
#define SPMV_CASE(SW)                                                      \
    case SW:                                                               \
        hipLaunchKernelGGL((k_spmv<E, SW>), dim3(g), dim3(kBlock), 0, s, n, \
                           A->rp, A->ci, A->val, x, y, part, tick, o0, o1); \
        break;
    switch (sw) {
        SPMV_CASE(2)
        SPMV_CASE(4)
        SPMV_CASE(8)
        SPMV_CASE(16)
        SPMV_CASE(32)
        SPMV_CASE(64)
    default:
        return hipErrorInvalidValue;
    }
#undef SPMV_CASE
    return hipGetLastError();
}

double spmv_bytes(const trlhip_csr *A, int E)
{
    // values + col_idx + row_ptr + x read once + y written once
    return (double)A->nnz * (8.0 * E + 4.0) + 4.0 * (A->n + 1) + 2.0 * 8.0 * E * A->n;
}

int run_spmv(trlhip_ctx *c, const trlhip_csr *A, int xcol, int ycol, double *o1)
{
    const size_t cs = (size_t)c->ld * c->E;
    const double *x = c->V + (size_t)xcol * cs;
    double *y = c->V + (size_t)ycol * cs;
    Timed t(TRLHIP_K_SPMV, c->s, spmv_bytes(A, c->E));
    hipError_t e = (c->E == 1)
        ? launch_spmv<1>(A->sw, c->n, A, x, y, c->part, c->tick, c->dscal + S_ALPHA, o1, c->s)
        : launch_spmv<2>(A->sw, c->n, A, x, y, c->part, c->tick, c->dscal + S_ALPHA, o1, c->s);
    return e == hipSuccess ? 0 : TRLHIP_EHIP;
}

int run_dot(trlhip_ctx *c, int xcol, int ycol, double *o1)
{
    const size_t cs = (size_t)c->ld * c->E;
    const int g = grid_for((long)c->n * c->E, kBlock * 8);
    Timed t(TRLHIP_K_MISC, c->s, 2.0 * 8.0 * c->E * c->n);
    if (c->E == 1)
        hipLaunchKernelGGL((k_dot<1>), dim3(g), dim3(kBlock), 0, c->s, c->n,
                           c->V + (size_t)xcol * cs, c->V + (size_t)ycol * cs, c->part,
                           c->tick, c->dscal + S_ALPHA, o1);
    else
        hipLaunchKernelGGL((k_dot<2>), dim3(g), dim3(kBlock), 0, c->s, c->n,
                           c->V + (size_t)xcol * cs, c->V + (size_t)ycol * cs, c->part,
                           c->tick, c->dscal + S_ALPHA, o1);
    return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
}

int run_update(trlhip_ctx *c, int c0, int m, const double *dh, int rcol, double *o_rr2,
               double *o_hh, int kind)
{
    const size_t cs = (size_t)c->ld * c->E;
    const int g = grid_for(c->n, kBlock * 4);
    const size_t shm = sizeof(double) * (size_t)c->E * (m > 0 ? m : 1);
    Timed t(kind, c->s, 8.0 * c->E * (double)c->n * (m + 2));
    if (c->E == 1)
        hipLaunchKernelGGL((k_update<1>), dim3(g), dim3(kBlock), shm, c->s, c->n, c->V,
                           (long)c->ld, c0, m, dh, c->V + (size_t)rcol * cs, c->part,
                           c->tick + 63, o_rr2, o_hh);
    else
        hipLaunchKernelGGL((k_update<2>), dim3(g), dim3(kBlock), shm, c->s, c->n, c->V,
                           (long)c->ld, c0, m, dh, c->V + (size_t)rcol * cs, c->part,
                           c->tick + 63, o_rr2, o_hh);
    return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
}

template <int E, int CB>
void launch_proj(trlhip_ctx *c, int g, int c0, int m, const double *r, double *hout)
{
    const int nch = (m + CB - 1) / CB;
    hipLaunchKernelGGL((k_proj<E, CB>), dim3(g, nch), dim3(kBlock), 0, c->s, c->n, c->V,
                       (long)c->ld, c0, m, r, c->part, c->tick, hout);
}

int run_proj(trlhip_ctx *c, int c0, int m, int rcol, double *hout)
{
    if (m <= 0) return 0;
    if ((m + 7) / 8 > 60) return TRLHIP_EINVAL;  // one ticket per column chunk
    const size_t cs = (size_t)c->ld * c->E;
    const double *r = c->V + (size_t)rcol * cs;
    const int g = grid_for(c->n, kBlock * 4);
    Timed t(TRLHIP_K_PROJ, c->s, 8.0 * c->E * (double)c->n * (m + 1));
    if (c->E == 1) {
        if (m <= 8) launch_proj<1, 8>(c, g, c0, m, r, hout);
        else if (m <= 16) launch_proj<1, 16>(c, g, c0, m, r, hout);
        else launch_proj<1, 32>(c, g, c0, m, r, hout);
    } else {
        if (m <= 8) launch_proj<2, 8>(c, g, c0, m, r, hout);
        else launch_proj<2, 16>(c, g, c0, m, r, hout);
    }
    return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
}

// copy scalars back and wait
int fetch(trlhip_ctx *c, const double *dsrc, double *hdst, size_t cnt)
{
    TRY(hipMemcpyAsync(hdst, dsrc, cnt * sizeof(double), hipMemcpyDeviceToHost, c->s));
    TRY(hipStreamSynchronize(c->s));
    stats_resolve();
    return 0;
}

bool valid_cols(const trlhip_ctx *c, int c0, int m)
{
    return c0 >= 0 && m >= 0 && c0 + m <= c->ncols;
}

}  // namespace

// ================================================================ C ABI

extern "C" {

const trlhip_csr *trlhip_csr_from_op(void (*op)(void), void *mvparam, int vtype)
{
    const trlhip_csr *A = (const trlhip_csr *)mvparam;
    if (!A) return nullptr;
    if (vtype == TRLHIP_R64 && op != (void (*)(void))trlhip_csr_matvec) return nullptr;
    if (vtype == TRLHIP_C128 && op != (void (*)(void))trlhip_csr_zmatvec) return nullptr;
    if (A->magic != kCsrMagic || A->vtype != vtype) return nullptr;
    return A;
}

int trlhip_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int trlhip_csr_create(trlhip_csr **out, int n, long nnz, const int *row_ptr,
                      const int *col_idx, const void *values, int vtype)
{
    if (!out || n <= 0 || nnz < 0 || !row_ptr || (nnz > 0 && (!col_idx || !values)))
        return TRLHIP_EINVAL;
    if (vtype != TRLHIP_R64 && vtype != TRLHIP_C128) return TRLHIP_EINVAL;
    if (trlhip_device_count() <= 0) return TRLHIP_ENODEV;
    trlhip_csr *A = (trlhip_csr *)calloc(1, sizeof(trlhip_csr));
    if (!A) return TRLHIP_ENOMEM;
    A->magic = kCsrMagic;
    A->n = n;
    A->nnz = nnz;
    A->vtype = vtype;
    A->esz = vtype == TRLHIP_C128 ? 2 : 1;
    A->sw = pick_sw(nnz, n);
    int rc = 0;
    if (hipStreamCreate(&A->s) != hipSuccess) rc = TRLHIP_EHIP;
    if (!rc && hipMalloc((void **)&A->rp, sizeof(int) * (size_t)(n + 1)) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMalloc((void **)&A->ci, sizeof(int) * (size_t)(nnz > 0 ? nnz : 1)) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMalloc((void **)&A->val, sizeof(double) * A->esz * (size_t)(nnz > 0 ? nnz : 1)) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMemcpy(A->rp, row_ptr, sizeof(int) * (size_t)(n + 1), hipMemcpyHostToDevice) != hipSuccess) rc = TRLHIP_EHIP;
    if (!rc && nnz > 0 && hipMemcpy(A->ci, col_idx, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess) rc = TRLHIP_EHIP;
    if (!rc && nnz > 0 && hipMemcpy(A->val, values, sizeof(double) * A->esz * (size_t)nnz, hipMemcpyHostToDevice) != hipSuccess) rc = TRLHIP_EHIP;
    if (rc) {
        trlhip_csr_destroy(A);
        return rc;
    }
    *out = A;
    return 0;
}

int trlhip_csr_destroy(trlhip_csr *A)
{
    if (!A) return 0;
    if (A->rp) (void)hipFree(A->rp);
    if (A->ci) (void)hipFree(A->ci);
    if (A->val) (void)hipFree(A->val);
    if (A->dx) (void)hipFree(A->dx);
    if (A->dy) (void)hipFree(A->dy);
    if (A->part) (void)hipFree(A->part);
    if (A->tick) (void)hipFree(A->tick);
    if (A->dscal) (void)hipFree(A->dscal);
    if (A->s) (void)hipStreamDestroy(A->s);
    A->magic = 0;
    free(A);
    return 0;
}

// Host-callable operator: stage x in, y = A x on the device, stage y out.
static int csr_matvec_host(trlhip_csr *A, int nrow, int ncol, const double *x, int ldx,
                           double *y, int ldy)
{
    const int E = A->esz;
    if (!A->dx) {
        if (hipMalloc((void **)&A->dx, sizeof(double) * E * (size_t)A->n) != hipSuccess) return TRLHIP_ENOMEM;
        if (hipMalloc((void **)&A->dy, sizeof(double) * E * (size_t)A->n) != hipSuccess) return TRLHIP_ENOMEM;
        if (hipMalloc((void **)&A->part, sizeof(double) * kMaxGrid) != hipSuccess) return TRLHIP_ENOMEM;
        if (hipMalloc((void **)&A->tick, sizeof(unsigned) * 4) != hipSuccess) return TRLHIP_ENOMEM;
        if (hipMalloc((void **)&A->dscal, sizeof(double) * 4) != hipSuccess) return TRLHIP_ENOMEM;
        TRY(hipMemset(A->tick, 0, sizeof(unsigned) * 4));
    }
    for (int j = 0; j < ncol; ++j) {
        TRY(hipMemcpyAsync(A->dx, x + (size_t)j * ldx * E, sizeof(double) * E * (size_t)nrow,
                           hipMemcpyHostToDevice, A->s));
        hipError_t e = (E == 1)
            ? launch_spmv<1>(A->sw, nrow, A, A->dx, A->dy, A->part, A->tick, A->dscal, nullptr, A->s)
            : launch_spmv<2>(A->sw, nrow, A, A->dx, A->dy, A->part, A->tick, A->dscal, nullptr, A->s);
        if (e != hipSuccess) return TRLHIP_EHIP;
        TRY(hipMemcpyAsync(y + (size_t)j * ldy * E, A->dy, sizeof(double) * E * (size_t)nrow,
                           hipMemcpyDeviceToHost, A->s));
    }
    TRY(hipStreamSynchronize(A->s));
    return 0;
}

void trlhip_csr_matvec(const int nrow, const int ncol, const double *x, const int ldx,
                       double *y, const int ldy, void *mvparam)
{
    trlhip_csr *A = (trlhip_csr *)mvparam;
    if (!A || A->magic != kCsrMagic || A->esz != 1 || nrow != A->n) {
        for (int j = 0; j < ncol; ++j)
            for (int i = 0; i < nrow; ++i) y[(size_t)j * ldy + i] = __builtin_nan("");
        return;
    }
    if (csr_matvec_host(A, nrow, ncol, x, ldx, y, ldy) != 0) {
        for (int j = 0; j < ncol; ++j)
            for (int i = 0; i < nrow; ++i) y[(size_t)j * ldy + i] = __builtin_nan("");
    }
}

void trlhip_csr_zmatvec(const int nrow, const int ncol, const void *x, const int ldx,
                        void *y, const int ldy, void *mvparam)
{
    trlhip_csr *A = (trlhip_csr *)mvparam;
    double *yd = (double *)y;
    if (!A || A->magic != kCsrMagic || A->esz != 2 || nrow != A->n ||
        csr_matvec_host(A, nrow, ncol, (const double *)x, ldx, yd, ldy) != 0) {
        for (int j = 0; j < ncol; ++j)
            for (int i = 0; i < 2 * nrow; ++i) yd[(size_t)j * ldy * 2 + i] = __builtin_nan("");
    }
}

int trlhip_ctx_destroy(trlhip_ctx *c)
{
    if (!c) return 0;
    if (c->s) (void)hipStreamSynchronize(c->s);
    if (c->V) (void)hipFree(c->V);
    if (c->part) (void)hipFree(c->part);
    if (c->tick) (void)hipFree(c->tick);
    if (c->dscal) (void)hipFree(c->dscal);
    if (c->dcoef) (void)hipFree(c->dcoef);
    if (c->dh) (void)hipFree(c->dh);
    if (c->dY) (void)hipFree(c->dY);
    if (c->hbuf) (void)hipHostFree(c->hbuf);
    if (c->hY) (void)hipHostFree(c->hY);
    if (c->s) (void)hipStreamDestroy(c->s);
    free(c);
    return 0;
}

int trlhip_ctx_create(trlhip_ctx **out, int n, int ncols, int vtype)
{
    if (!out || n <= 0 || ncols <= 0) return TRLHIP_EINVAL;
    if (vtype != TRLHIP_R64 && vtype != TRLHIP_C128) return TRLHIP_EINVAL;
    if (trlhip_device_count() <= 0) return TRLHIP_ENODEV;
    trlhip_ctx *c = (trlhip_ctx *)calloc(1, sizeof(trlhip_ctx));
    if (!c) return TRLHIP_ENOMEM;
    c->n = n;
    c->ld = (n + 63) & ~63;
    c->ncols = ncols;
    c->vtype = vtype;
    c->E = vtype == TRLHIP_C128 ? 2 : 1;
    const size_t ncoef = 2 * (size_t)(ncols + 2);
    int rc = 0;
    if (hipStreamCreate(&c->s) != hipSuccess) rc = TRLHIP_EHIP;
    if (!rc && hipMalloc((void **)&c->V, sizeof(double) * c->E * (size_t)c->ld * ncols) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMalloc((void **)&c->part, sizeof(double) * kMaxGrid * ((size_t)c->E * ncols + 64)) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMalloc((void **)&c->tick, sizeof(unsigned) * 64) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMalloc((void **)&c->dscal, sizeof(double) * 16) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMalloc((void **)&c->dcoef, sizeof(double) * ncoef) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMalloc((void **)&c->dh, sizeof(double) * ncoef) != hipSuccess) rc = TRLHIP_ENOMEM;
    c->hcap = 16 + 2 * ncoef;
    if (!rc && hipHostMalloc((void **)&c->hbuf, sizeof(double) * c->hcap, 0) != hipSuccess) rc = TRLHIP_ENOMEM;
    if (!rc && hipMemsetAsync(c->tick, 0, sizeof(unsigned) * 64, c->s) != hipSuccess) rc = TRLHIP_EHIP;
    if (!rc && hipMemsetAsync(c->V, 0, sizeof(double) * c->E * (size_t)c->ld * ncols, c->s) != hipSuccess) rc = TRLHIP_EHIP;
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
    if (!c || !x || col < 0 || col >= c->ncols) return TRLHIP_EINVAL;
    const size_t cs = (size_t)c->ld * c->E;
    TRY(hipMemcpyAsync(c->V + (size_t)col * cs, x, sizeof(double) * c->E * (size_t)c->n,
                       hipMemcpyHostToDevice, c->s));
    TRY(hipStreamSynchronize(c->s));
    return 0;
}

int trlhip_col_get(trlhip_ctx *c, int col, void *x)
{
    if (!c || !x || col < 0 || col >= c->ncols) return TRLHIP_EINVAL;
    const size_t cs = (size_t)c->ld * c->E;
    TRY(hipMemcpyAsync(x, c->V + (size_t)col * cs, sizeof(double) * c->E * (size_t)c->n,
                       hipMemcpyDeviceToHost, c->s));
    TRY(hipStreamSynchronize(c->s));
    stats_resolve();
    return 0;
}

int trlhip_col_copy(trlhip_ctx *c, int dst, int src)
{
    if (!c || dst < 0 || src < 0 || dst >= c->ncols || src >= c->ncols) return TRLHIP_EINVAL;
    if (dst == src) return 0;
    const size_t cs = (size_t)c->ld * c->E;
    Timed t(TRLHIP_K_MISC, c->s, 2.0 * 8.0 * c->E * c->n);
    TRY(hipMemcpyAsync(c->V + (size_t)dst * cs, c->V + (size_t)src * cs,
                       sizeof(double) * c->E * (size_t)c->n, hipMemcpyDeviceToDevice, c->s));
    return 0;
}

int trlhip_col_scale(trlhip_ctx *c, int col, double s)
{
    if (!c || col < 0 || col >= c->ncols) return TRLHIP_EINVAL;
    const size_t cs = (size_t)c->ld * c->E;
    const long ne = (long)c->n * c->E;
    const int g = grid_for(ne, kBlock * 8);
    Timed t(TRLHIP_K_MISC, c->s, 2.0 * 8.0 * ne);
    hipLaunchKernelGGL(k_scale, dim3(g), dim3(kBlock), 0, c->s, ne, c->V + (size_t)col * cs, s);
    return hipGetLastError() == hipSuccess ? 0 : TRLHIP_EHIP;
}

int trlhip_spmv(trlhip_ctx *c, const trlhip_csr *A, int xcol, int ycol, double *xdoty)
{
    if (!c || !A || A->magic != kCsrMagic || A->n != c->n || A->esz != c->E) return TRLHIP_EINVAL;
    if (xcol < 0 || ycol < 0 || xcol >= c->ncols || ycol >= c->ncols || xcol == ycol) return TRLHIP_EINVAL;
    int rc = run_spmv(c, A, xcol, ycol, nullptr);
    if (rc) return rc;
    rc = fetch(c, c->dscal + S_ALPHA, c->hbuf, 1);
    if (rc) return rc;
    if (xdoty) *xdoty = c->hbuf[0];
    return 0;
}

int trlhip_lanczos_step(trlhip_ctx *c, const trlhip_csr *A, int qa, int qb, double beta_prev,
                        int rcol, double *alpha, double *rr2)
{
    if (!c || qa < 0 || qa >= c->ncols || rcol < 0 || rcol >= c->ncols || qa == rcol) return TRLHIP_EINVAL;
    if (qb >= 0 && (qb != qa - 1 || qb >= c->ncols)) return TRLHIP_EINVAL;
    if (A && (A->magic != kCsrMagic || A->n != c->n || A->esz != c->E)) return TRLHIP_EINVAL;
    const int E = c->E;
    // coefficient slots: [beta_prev, alpha] for columns (qb, qa)
    double *slot_alpha = c->dcoef + E;
    if (qb >= 0) {
        c->hbuf[0] = beta_prev;
        c->hbuf[1] = 0.0;
        TRY(hipMemcpyAsync(c->dcoef, c->hbuf, sizeof(double) * E, hipMemcpyHostToDevice, c->s));
    }
    int rc = A ? run_spmv(c, A, qa, rcol, slot_alpha) : run_dot(c, qa, rcol, slot_alpha);
    if (rc) return rc;
    rc = (qb >= 0) ? run_update(c, qb, 2, c->dcoef, rcol, c->dscal + S_RR2, nullptr, TRLHIP_K_RECUR)
                   : run_update(c, qa, 1, slot_alpha, rcol, c->dscal + S_RR2, nullptr, TRLHIP_K_RECUR);
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
    if (!c || jnd < 1 || jnd > c->ncols || qa != jnd - 1 || rcol < 0 || rcol >= c->ncols || rcol == qa)
        return TRLHIP_EINVAL;
    if (A && (A->magic != kCsrMagic || A->n != c->n || A->esz != c->E)) return TRLHIP_EINVAL;
    const int E = c->E;
    double *h = c->hbuf + 16;
    for (int i = 0; i < jnd - 1; ++i) {
        h[E * i] = beta[i];
        if (E == 2) h[E * i + 1] = 0.0;
    }
    if (jnd > 1)
        TRY(hipMemcpyAsync(c->dcoef, h, sizeof(double) * E * (size_t)(jnd - 1), hipMemcpyHostToDevice, c->s));
    double *slot_alpha = c->dcoef + (size_t)E * (jnd - 1);
    int rc = A ? run_spmv(c, A, qa, rcol, slot_alpha) : run_dot(c, qa, rcol, slot_alpha);
    if (rc) return rc;
    rc = run_update(c, 0, jnd, c->dcoef, rcol, c->dscal + S_RR2, nullptr, TRLHIP_K_RECUR);
    if (rc) return rc;
    rc = fetch(c, c->dscal, c->hbuf + 4, 2);
    if (rc) return rc;
    if (alpha) *alpha = c->hbuf[4 + S_ALPHA];
    if (rr2) *rr2 = c->hbuf[4 + S_RR2];
    return 0;
}

int trlhip_cgs_pass(trlhip_ctx *c, int c0, int m, int rcol, double *hh, double *rr2, double *h_out)
{
    if (!c || m < 1 || !valid_cols(c, c0, m) || rcol < 0 || rcol >= c->ncols || (rcol >= c0 && rcol < c0 + m))
        return TRLHIP_EINVAL;
    int rc = run_proj(c, c0, m, rcol, c->dh);
    if (rc) return rc;
    rc = run_update(c, c0, m, c->dh, rcol, c->dscal + S_RR2, c->dscal + S_HH, TRLHIP_K_UPDATE);
    if (rc) return rc;
    const int E = c->E;
    if (h_out) {
        TRY(hipMemcpyAsync(c->hbuf + 16, c->dh, sizeof(double) * E * (size_t)m, hipMemcpyDeviceToHost, c->s));
    }
    rc = fetch(c, c->dscal, c->hbuf + 4, 3);
    if (rc) return rc;
    if (hh) *hh = c->hbuf[4 + S_HH];
    if (rr2) *rr2 = c->hbuf[4 + S_RR2];
    if (h_out) memcpy(h_out, c->hbuf + 16, sizeof(double) * E * (size_t)m);
    return 0;
}

int trlhip_proj(trlhip_ctx *c, int c0, int m, int rcol, double *h)
{
    if (!c || !h || m < 1 || !valid_cols(c, c0, m) || rcol < 0 || rcol >= c->ncols) return TRLHIP_EINVAL;
    int rc = run_proj(c, c0, m, rcol, c->dh);
    if (rc) return rc;
    rc = fetch(c, c->dh, c->hbuf + 16, (size_t)c->E * m);
    if (rc) return rc;
    memcpy(h, c->hbuf + 16, sizeof(double) * c->E * (size_t)m);
    return 0;
}

int trlhip_update(trlhip_ctx *c, int c0, int m, const double *h, int rcol, double *rr2)
{
    if (!c || !h || m < 0 || !valid_cols(c, c0, m) || rcol < 0 || rcol >= c->ncols ||
        (m > 0 && rcol >= c0 && rcol < c0 + m))
        return TRLHIP_EINVAL;
    const int E = c->E;
    memcpy(c->hbuf + 16, h, sizeof(double) * E * (size_t)m);
    if (m > 0)
        TRY(hipMemcpyAsync(c->dh, c->hbuf + 16, sizeof(double) * E * (size_t)m, hipMemcpyHostToDevice, c->s));
    int rc = run_update(c, c0, m, c->dh, rcol, c->dscal + S_RR2, nullptr, TRLHIP_K_UPDATE);
    if (rc) return rc;
    rc = fetch(c, c->dscal + S_RR2, c->hbuf + 4, 1);
    if (rc) return rc;
    if (rr2) *rr2 = c->hbuf[4];
    return 0;
}

int trlhip_ritz(trlhip_ctx *c, int lck, int jl, int k, const double *Y, int ldy)
{
    if (!c || !Y || lck < 0 || jl < 1 || k < 1 || k > jl || ldy < jl || lck + jl > c->ncols)
        return TRLHIP_EINVAL;
    const int KT = (k + 15) / 16;
    if (KT > 8) return TRLHIP_EINVAL;  // k <= 128
    const size_t ny = (size_t)ldy * k;
    if (c->dYcap < ny) {
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
    const int tpw = 4;  // 16-row tiles per wave
    const long rows_per_block = 16L * tpw * (kBlock / 64);
    const int g = (int)((nrows + rows_per_block - 1) / rows_per_block);
    const int jl4 = (jl + 3) & ~3;
    const size_t shm = sizeof(double) * (size_t)jl4 * 16 * KT;
    if (shm > 160 * 1024) return TRLHIP_EINVAL;
    Timed t(TRLHIP_K_RITZ, c->s, 8.0 * (double)nrows * (jl + k));
    double *Vr = c->V;
    switch (KT) {
#define RITZ_CASE(K)                                                                        \
    case K:                                                                                 \
        hipLaunchKernelGGL((k_ritz_mfma<K>), dim3(g), dim3(kBlock), shm, c->s, nrows, Vr, \
                           ldr, lck, jl, k, c->dY, ldy, tpw);                               \
        break;
        RITZ_CASE(1)
        RITZ_CASE(2)
        RITZ_CASE(3)
        RITZ_CASE(4)
        RITZ_CASE(5)
        RITZ_CASE(6)
        RITZ_CASE(7)
        RITZ_CASE(8)
#undef RITZ_CASE
    }
    if (hipGetLastError() != hipSuccess) return TRLHIP_EHIP;
    TRY(hipStreamSynchronize(c->s));
    stats_resolve();
    return 0;
}

void trlhip_stats_enable(int on)
{
    stats_resolve();
    g_stats.on = on ? 1 : 0;
    for (int k = 0; k < TRLHIP_K_COUNT; ++k) {
        g_stats.launches[k] = 0;
        g_stats.ms[k] = 0;
        g_stats.bytes[k] = 0;
    }
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

}  // extern "C"
