// trlnet.hip -- the learned start vector: a 4-layer 3x3 convolutional net
// (1 -> 32 -> 32 -> 32 -> 1 channels, ReLU between layers, periodic
// padding) that maps a lattice's on-site potential to a start vector for
// trlan().  Forward and backward run on BF16 MFMA (v_mfma_f32_16x16x32_bf16)
// with FP32 accumulation; master weights and Adam state are FP32.  No
// counterpart in the reference (it takes a start vector through
// info->guess = 1 and evec(:, nec)).
//
// Layout: activations NHWC, pixels flattened over (image, y, x), 32
// channels of BF16 (64 B) per pixel; the input potential and the output are
// one FP32 value per pixel.  A convolution is an implicit GEMM: M = 16
// pixels per MFMA tile, N = 16 output channels, K = 32 = one tap's 32 input
// channels (or, for the one-channel input layer, the 9 taps padded to 32).
//
// MFMA f32 16x16x32 bf16 operand layout (lane l): A row l%16 with K
// elements 8*(l/16) .. +7, B column l%16 with the same K elements, result
// D[4*(l/16) + i][l%16] in accumulator element i.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "trlnet.h"

namespace {

constexpr int C = 32;          // hidden channels
constexpr int T = 9;           // 3x3 taps
constexpr int NB = 256;        // threads per block
constexpr int PIX_BLK = 64;    // pixels per block (4 waves x 16)

// parameter offsets in the flat FP32 vector
constexpr int OW1 = 0, OB1 = OW1 + C * T;             // W1[co][t]
constexpr int OW2 = OB1 + C, OB2 = OW2 + C * T * C;   // W2[co][t][ci]
constexpr int OW3 = OB2 + C, OB3 = OW3 + C * T * C;   // W3[co][t][ci]
constexpr int OW4 = OB3 + C, OB4 = OW4 + T * C;       // W4[t][ci]
constexpr int NPARAM = OB4 + 1;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16_t;  // storage

__host__ __device__ inline bf16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(bf16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ inline __bf16 as_bf(bf16_t h) { return __builtin_bit_cast(__bf16, h); }

struct Geo {
    int nimg, h, w;
    long np;  // pixels
};

// pixel p shifted by tap t (dy = t/3 - 1, dx = t%3 - 1; sign -1 inverts), periodic
__device__ inline long nbr(const Geo &g, long p, int t, int sign)
{
    const int dy = sign * (t / 3 - 1), dx = sign * (t % 3 - 1);
    const long hw = (long)g.h * g.w;
    const long img = p / hw;
    const int r = (int)(p - img * hw);
    int y = r / g.w + dy, x = r % g.w + dx;
    y = y < 0 ? y + g.h : (y >= g.h ? y - g.h : y);
    x = x < 0 ? x + g.w : (x >= g.w ? x - g.w : x);
    return img * hw + (long)y * g.w + x;
}

// The 9 neighbours of pixel p (tap t = (dy+1)*3 + dx+1, periodic), with
// one division per pixel instead of one per tap: nb[t] equals nbr(g, p, t, 1).
__device__ inline void nbr9(const Geo &g, long p, long nb[9])
{
    const long hw = (long)g.h * g.w;
    const long img = p / hw;
    const int r = (int)(p - img * hw);
    const int y = r / g.w, x = r - y * g.w;
    const int ym = y == 0 ? g.h - 1 : y - 1, yp = y == g.h - 1 ? 0 : y + 1;
    const int xm = x == 0 ? g.w - 1 : x - 1, xp = x == g.w - 1 ? 0 : x + 1;
    const long base = img * hw;
    const int ys[3] = {ym, y, yp}, xs[3] = {xm, x, xp};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) nb[a * 3 + b] = base + (long)ys[a] * g.w + xs[b];
}

__device__ inline bf16x8 load8(const bf16_t *p)
{
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    return __builtin_bit_cast(bf16x8, v);
}

__device__ inline bf16x8 zero8() { return __builtin_bit_cast(bf16x8, uint4{0, 0, 0, 0}); }

// ---------------------------------------------------------------- forward

// Layer 1: one input channel (FP32 potential, rounded to BF16), K = the 9
// taps padded to 32.  out = relu(W1 * x + b1), BF16, 32 channels.
__device__ __forceinline__ void fwd1_body(const Geo &g, long bid, const float *__restrict__ x,
                                          const bf16_t *__restrict__ w1, const float *__restrict__ b1,
                                          bf16_t *__restrict__ out)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long p0 = (bid * 4 + wv) * 16;
    if (p0 >= g.np) return;
    const int m = lane & 15, kq = lane >> 4;
    const long p = p0 + m;
    bf16x8 a = zero8();
    if (p < g.np) {
        long nb[9];
        nbr9(g, p, nb);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int t = 8 * kq + j;
            if (t < T) a[j] = as_bf(f2bf(x[nb[t]]));
        }
    }
    for (int nt = 0; nt < 2; ++nt) {
        const int co = 16 * nt + m;
        bf16x8 b = zero8();
        for (int j = 0; j < 8; ++j) {
            const int t = 8 * kq + j;
            if (t < T) b[j] = as_bf(w1[co * T + t]);
        }
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        for (int i = 0; i < 4; ++i) {
            const long q = p0 + 4 * kq + i;
            if (q < g.np) out[q * C + co] = f2bf(fmaxf(acc[i] + b1[co], 0.f));
        }
    }
}

__global__ __launch_bounds__(NB) void k_fwd1(Geo g, const float *__restrict__ x,
                                             const bf16_t *__restrict__ w1, const float *__restrict__ b1,
                                             bf16_t *__restrict__ out)
{
    fwd1_body(g, blockIdx.x, x, w1, b1, out);
}

// Layers 2 and 3: 32 -> 32 channels, K = 9 taps x 32 channels.
__device__ __forceinline__ void fwd32_body(const Geo &g, long bid, const bf16_t *__restrict__ in,
                                           const bf16_t *__restrict__ w, const float *__restrict__ bias,
                                           bf16_t *__restrict__ out)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long p0 = (bid * 4 + wv) * 16;
    if (p0 >= g.np) return;
    const int m = lane & 15, kq = lane >> 4;
    const long p = p0 + m;
    f4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
    long nb[9];
    if (p < g.np) nbr9(g, p, nb);
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const bf16x8 a = p < g.np ? load8(in + nb[t] * C + 8 * kq) : zero8();
        for (int nt = 0; nt < 2; ++nt) {
            const int co = 16 * nt + m;
            const bf16x8 b = load8(w + ((long)co * T + t) * C + 8 * kq);
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[nt], 0, 0, 0);
        }
    }
    for (int nt = 0; nt < 2; ++nt) {
        const int co = 16 * nt + m;
        for (int i = 0; i < 4; ++i) {
            const long q = p0 + 4 * kq + i;
            if (q < g.np) out[q * C + co] = f2bf(fmaxf(acc[nt][i] + bias[co], 0.f));
        }
    }
}

__global__ __launch_bounds__(NB) void k_fwd32(Geo g, const bf16_t *__restrict__ in,
                                              const bf16_t *__restrict__ w, const float *__restrict__ bias,
                                              bf16_t *__restrict__ out)
{
    fwd32_body(g, blockIdx.x, in, w, bias, out);
}

// Layer 4: 32 -> 1 channel (N padded to 16, column 0 real), FP32 output.
__device__ __forceinline__ void fwd4_body(const Geo &g, long bid, const bf16_t *__restrict__ in,
                                          const bf16_t *__restrict__ w4, const float *__restrict__ b4,
                                          float *__restrict__ out)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long p0 = (bid * 4 + wv) * 16;
    if (p0 >= g.np) return;
    const int m = lane & 15, kq = lane >> 4;
    const long p = p0 + m;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    long nb[9];
    if (p < g.np) nbr9(g, p, nb);
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const bf16x8 a = p < g.np ? load8(in + nb[t] * C + 8 * kq) : zero8();
        const bf16x8 b = m == 0 ? load8(w4 + (long)t * C + 8 * kq) : zero8();
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    if (m == 0)
        for (int i = 0; i < 4; ++i) {
            const long q = p0 + 4 * kq + i;
            if (q < g.np) out[q] = acc[i] + b4[0];
        }
}

__global__ __launch_bounds__(NB) void k_fwd4(Geo g, const bf16_t *__restrict__ in,
                                             const bf16_t *__restrict__ w4, const float *__restrict__ b4,
                                             float *__restrict__ out)
{
    fwd4_body(g, blockIdx.x, in, w4, b4, out);
}

// Groups of images of different sizes in one launch per layer: group i
// (geometry gg[i].g, first pixel gg[i].off) owns blocks [pre[i], pre[i+1]).
struct GroupGeo {
    Geo g;
    long off;
};

// The group owning block blk (pre increasing, pre[0] = 0): each wave
// narrows the range 64-fold per step with a ballot over 64 candidates.
__device__ __forceinline__ int group_of(const int *__restrict__ pre, int ng, int blk)
{
    const int lane = threadIdx.x & 63;
    int lo = 0, hi = ng - 1;
    while (lo < hi) {
        const int step = (hi - lo + 64) / 64;
        const int k = lo + lane * step;
        const unsigned long long m = __ballot(k <= hi && pre[k] <= blk);
        lo += (63 - __clzll(m)) * step;
        hi = min(hi, lo + step - 1);
        if (step == 1) break;
    }
    return lo;
}

__global__ __launch_bounds__(NB) void k_fwd1_many(const GroupGeo *__restrict__ gg, const int *__restrict__ pre,
                                                  int ng, const float *__restrict__ x,
                                                  const bf16_t *__restrict__ w1, const float *__restrict__ b1,
                                                  bf16_t *__restrict__ out)
{
    const int i = group_of(pre, ng, blockIdx.x);
    const GroupGeo q = gg[i];
    fwd1_body(q.g, blockIdx.x - pre[i], x + q.off, w1, b1, out + q.off * C);
}

__global__ __launch_bounds__(NB) void k_fwd32_many(const GroupGeo *__restrict__ gg, const int *__restrict__ pre,
                                                   int ng, const bf16_t *__restrict__ in,
                                                   const bf16_t *__restrict__ w, const float *__restrict__ bias,
                                                   bf16_t *__restrict__ out)
{
    const int i = group_of(pre, ng, blockIdx.x);
    const GroupGeo q = gg[i];
    fwd32_body(q.g, blockIdx.x - pre[i], in + q.off * C, w, bias, out + q.off * C);
}

__global__ __launch_bounds__(NB) void k_fwd4_many(const GroupGeo *__restrict__ gg, const int *__restrict__ pre,
                                                  int ng, const bf16_t *__restrict__ in,
                                                  const bf16_t *__restrict__ w4, const float *__restrict__ b4,
                                                  float *__restrict__ out)
{
    const int i = group_of(pre, ng, blockIdx.x);
    const GroupGeo q = gg[i];
    fwd4_body(q.g, blockIdx.x - pre[i], in + q.off * C, w4, b4, out + q.off);
}

// ------------------------------------------------------------ loss, grads

// Per image: c = (o.t)^2 / (|o|^2 |t|^2), loss += (1 - c) / nimg and
// dL/do = (2 / nimg) (c o / |o|^2 - (o.t) t / (|o|^2 |t|^2)).  One block
// per image; loss_part[img] = (1 - c) / nimg.
__global__ __launch_bounds__(NB) void k_loss(Geo g, const float *__restrict__ o,
                                             const float *__restrict__ tg, float *__restrict__ dout,
                                             double *__restrict__ loss_part)
{
    __shared__ double red[3][NB / 64];
    const long hw = (long)g.h * g.w;
    const long base = (long)blockIdx.x * hw;
    double ot = 0.0, oo = 0.0, tt = 0.0;
    for (long i = threadIdx.x; i < hw; i += NB) {
        const double ov = o[base + i], tv = tg[base + i];
        ot += ov * tv;
        oo += ov * ov;
        tt += tv * tv;
    }
    double v[3] = {ot, oo, tt};
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = 0; k < 3; ++k) {
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
        if (lane == 0) red[k][wv] = v[k];
    }
    __syncthreads();
    double s[3];
    for (int k = 0; k < 3; ++k) s[k] = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
    const double oo_ = s[1] > 1e-300 ? s[1] : 1e-300, tt_ = s[2] > 1e-300 ? s[2] : 1e-300;
    const double c = s[0] * s[0] / (oo_ * tt_);
    const double sc = 2.0 / g.nimg;
    for (long i = threadIdx.x; i < hw; i += NB)
        dout[base + i] = (float)(sc * (c * o[base + i] / oo_ - s[0] * tg[base + i] / (oo_ * tt_)));
    if (threadIdx.x == 0) loss_part[blockIdx.x] = (1.0 - c) / g.nimg;
}

// dX for layer 4's input: dA3[p][ci] = relu'(a3[p][ci]) * sum_t dO[nbr(p,-t)] W4[t][ci].
// K = the 9 taps padded to 32 (one output channel), N = 32 input channels.
__global__ __launch_bounds__(NB) void k_bwd4(Geo g, const float *__restrict__ dout,
                                             const bf16_t *__restrict__ w4, const bf16_t *__restrict__ a3,
                                             bf16_t *__restrict__ da3)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long p0 = ((long)blockIdx.x * 4 + wv) * 16;
    if (p0 >= g.np) return;
    const int m = lane & 15, kq = lane >> 4;
    const long p = p0 + m;
    bf16x8 a = zero8();
    for (int j = 0; j < 8; ++j) {
        const int t = 8 * kq + j;
        if (t < T && p < g.np) a[j] = as_bf(f2bf(dout[nbr(g, p, t, -1)]));
    }
    for (int nt = 0; nt < 2; ++nt) {
        const int ci = 16 * nt + m;
        bf16x8 b = zero8();
        for (int j = 0; j < 8; ++j) {
            const int t = 8 * kq + j;
            if (t < T) b[j] = as_bf(w4[t * C + ci]);
        }
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        for (int i = 0; i < 4; ++i) {
            const long q = p0 + 4 * kq + i;
            if (q < g.np) {
                const float mask = bf2f(a3[q * C + ci]) > 0.f ? 1.f : 0.f;
                da3[q * C + ci] = f2bf(acc[i] * mask);
            }
        }
    }
}

// dX for a 32 -> 32 layer: dA[p][ci] = relu'(a[p][ci]) * sum_t sum_co
// dY[nbr(p,-t)][co] W[co][t][ci], with wt = W transposed to [ci][t][co].
__global__ __launch_bounds__(NB) void k_bwd32(Geo g, const bf16_t *__restrict__ dy,
                                              const bf16_t *__restrict__ wt, const bf16_t *__restrict__ a,
                                              bf16_t *__restrict__ da)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long p0 = ((long)blockIdx.x * 4 + wv) * 16;
    if (p0 >= g.np) return;
    const int m = lane & 15, kq = lane >> 4;
    const long p = p0 + m;
    f4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
    for (int t = 0; t < T; ++t) {
        const bf16x8 av = p < g.np ? load8(dy + nbr(g, p, t, -1) * C + 8 * kq) : zero8();
        for (int nt = 0; nt < 2; ++nt) {
            const int ci = 16 * nt + m;
            const bf16x8 b = load8(wt + ((long)ci * T + t) * C + 8 * kq);
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b, acc[nt], 0, 0, 0);
        }
    }
    for (int nt = 0; nt < 2; ++nt) {
        const int ci = 16 * nt + m;
        for (int i = 0; i < 4; ++i) {
            const long q = p0 + 4 * kq + i;
            if (q < g.np) {
                const float mask = bf2f(a[q * C + ci]) > 0.f ? 1.f : 0.f;
                da[q * C + ci] = f2bf(acc[nt][i] * mask);
            }
        }
    }
}

// Weight gradient dW[co][t][ci] = sum_p dY[p][co] X[nbr(p,t)][ci] for a
// 32 -> 32 layer, K = pixels in chunks of 32 (8 per lane group), M = co,
// N = ci.  Block b takes pixel chunks b, b + nblk, ...; its 4 waves take
// alternate chunks; the block's partial (co, t, ci) and bias partial sums
// go to part[b][.] and are reduced by k_reduce in block order.
__global__ __launch_bounds__(NB) void k_dw32(Geo g, const bf16_t *__restrict__ dy,
                                             const bf16_t *__restrict__ x, float *__restrict__ part)
{
    constexpr int NW = C * T * C;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    f4 acc[T][2][2];
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[t][i][j] = f4{0.f, 0.f, 0.f, 0.f};
    float bsum[2] = {0.f, 0.f};
    const long nchunk = (g.np + 31) / 32;
    for (long ch = (long)blockIdx.x * 4 + wv; ch < nchunk; ch += (long)gridDim.x * 4) {
        const long pc = ch * 32 + 8 * kq;
        bf16x8 av[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            const int co = 16 * mt + m;
            av[mt] = zero8();
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (pc + j < g.np) {
                    const bf16_t h = dy[(pc + j) * C + co];
                    av[mt][j] = as_bf(h);
                    bsum[mt] += bf2f(h);
                }
        }
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const int ci = 16 * nt + m;
                bf16x8 bv = zero8();
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (pc + j < g.np) bv[j] = as_bf(x[nbr(g, pc + j, t, 1) * C + ci]);
#pragma unroll
                for (int mt = 0; mt < 2; ++mt)
                    acc[t][mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[mt], bv, acc[t][mt][nt], 0, 0, 0);
            }
    }
    // combine the 4 waves through the partial buffer rows (wave w writes
    // slot 4*b + w; the reduce kernel sums all slots in order)
    float *out = part + ((long)blockIdx.x * 4 + wv) * (NW + C);
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int co = 16 * mt + 4 * kq + i, ci = 16 * nt + m;
                    out[((long)co * T + t) * C + ci] = acc[t][mt][nt][i];
                }
    // bias: lanes of one co (same m) across the 4 lane groups
    for (int mt = 0; mt < 2; ++mt) {
        float v = bsum[mt];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        if (kq == 0) out[NW + 16 * mt + m] = v;
    }
}

// Weight gradients of layer 1 (x: FP32 potential) and layer 4 (dy: FP32
// output gradient, x = a3): plain FP32 reductions (288 + 32 values each).
// Layer 1: dW1[co][t] = sum_p dA1[p][co] x[nbr(p,t)], db1[co] = sum_p dA1[p][co].
// Layer 4: dW4[t][ci] = sum_p dO[p] a3[nbr(p,t)][ci], db4 = sum_p dO[p].
// Thread j < 320 of each block owns one output value over the block's
// pixel range; partials part[b][320] reduced by k_reduce.
__global__ __launch_bounds__(NB) void k_dw_small(Geo g, int layer, const bf16_t *__restrict__ dy32,
                                                 const float *__restrict__ x1, const float *__restrict__ dy1,
                                                 const bf16_t *__restrict__ x32, float *__restrict__ part)
{
    constexpr int NV = C * T + C;  // 320
    const long per = (g.np + gridDim.x - 1) / gridDim.x;
    const long pa = (long)blockIdx.x * per, pb = pa + per < g.np ? pa + per : g.np;
    for (int j = threadIdx.x; j < NV; j += NB) {
        float s = 0.f;
        if (layer == 1) {
            if (j < C * T) {
                const int co = j / T, t = j % T;
                for (long p = pa; p < pb; ++p) s += bf2f(dy32[p * C + co]) * bf2f(f2bf(x1[nbr(g, p, t, 1)]));
            } else {
                const int co = j - C * T;
                for (long p = pa; p < pb; ++p) s += bf2f(dy32[p * C + co]);
            }
        } else {
            if (j < C * T) {
                const int t = j / C, ci = j % C;
                for (long p = pa; p < pb; ++p) s += bf2f(f2bf(dy1[p])) * bf2f(x32[nbr(g, p, t, 1) * C + ci]);
            } else if (j == C * T) {
                for (long p = pa; p < pb; ++p) s += dy1[p];
            }
        }
        part[(long)blockIdx.x * NV + j] = s;
    }
}

// out[j] = sum_b part[b][j] over nb rows of stride len, in row order.
__global__ __launch_bounds__(NB) void k_reduce(const float *__restrict__ part, int nb, int len,
                                               float *__restrict__ out)
{
    for (int j = blockIdx.x * NB + threadIdx.x; j < len; j += gridDim.x * NB) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += part[(long)b * len + j];
        out[j] = (float)s;
    }
}

// Adam on the FP32 master parameters; refreshes the BF16 copies (W and
// the transposed W of layers 2/3 used by the backward data pass).
__global__ __launch_bounds__(NB) void k_adam(float *__restrict__ prm, const float *__restrict__ grad,
                                             float *__restrict__ m1, float *__restrict__ m2, int n,
                                             float lr, float b1t, float b2t)
{
    const int i = blockIdx.x * NB + threadIdx.x;
    if (i >= n) return;
    const float gr = grad[i];
    const float a = 0.9f * m1[i] + 0.1f * gr;
    const float b = 0.999f * m2[i] + 0.001f * gr * gr;
    m1[i] = a;
    m2[i] = b;
    prm[i] -= lr * (a / b1t) / (sqrtf(b / b2t) + 1e-8f);
}

__global__ __launch_bounds__(NB) void k_bf16_params(const float *__restrict__ prm, bf16_t *__restrict__ wb,
                                                    bf16_t *__restrict__ wt2, bf16_t *__restrict__ wt3)
{
    const int i = blockIdx.x * NB + threadIdx.x;
    if (i < NPARAM) wb[i] = f2bf(prm[i]);
    if (i < C * T * C) {
        // i = (co*T + t)*C + ci  ->  transposed (ci*T + t)*C + co
        const int co = i / (T * C), t = (i / C) % T, ci = i % C;
        const int j = (ci * T + t) * C + co;
        wt2[j] = f2bf(prm[OW2 + i]);
        wt3[j] = f2bf(prm[OW3 + i]);
    }
}

int grid_px(long np) { return (int)((np + PIX_BLK - 1) / PIX_BLK); }

}  // namespace

struct trlnet {
    float *prm = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    bf16_t *wb = nullptr, *wt2 = nullptr, *wt3 = nullptr;
    int step = 0;
    // workspace
    long cap = 0;  // pixels
    float *x = nullptr, *tg = nullptr, *o = nullptr, *dO = nullptr;
    bf16_t *a1 = nullptr, *a2 = nullptr, *a3 = nullptr, *d1 = nullptr, *d2 = nullptr;
    float *part = nullptr;
    long partcap = 0;
    double *lpart = nullptr;
    int lcap = 0;
    float *hstage = nullptr;  // pinned staging of trlnet_forward_many (in, then out)
    long hcap = 0;            // floats
    char *hgeo = nullptr, *dgeo = nullptr;  // group table of the many-group launches (pinned / device)
    long geocap = 0;                        // bytes
    hipStream_t s = nullptr;
};

namespace {

constexpr int kDwBlocks = 256;  // blocks of the weight-gradient reductions

#define NTRY(e)                                  \
    do {                                         \
        if ((e) != hipSuccess) return TRLNET_EHIP; \
    } while (0)

int ensure(trlnet *t, long np, int nimg)
{
    if (np > t->cap) {
        (void)hipStreamSynchronize(t->s);
        for (void *q : {(void *)t->x, (void *)t->tg, (void *)t->o, (void *)t->dO, (void *)t->a1,
                        (void *)t->a2, (void *)t->a3, (void *)t->d1, (void *)t->d2})
            if (q) (void)hipFree(q);
        t->cap = 0;
        NTRY(hipMalloc((void **)&t->x, sizeof(float) * np));
        NTRY(hipMalloc((void **)&t->tg, sizeof(float) * np));
        NTRY(hipMalloc((void **)&t->o, sizeof(float) * np));
        NTRY(hipMalloc((void **)&t->dO, sizeof(float) * np));
        NTRY(hipMalloc((void **)&t->a1, sizeof(bf16_t) * C * np));
        NTRY(hipMalloc((void **)&t->a2, sizeof(bf16_t) * C * np));
        NTRY(hipMalloc((void **)&t->a3, sizeof(bf16_t) * C * np));
        NTRY(hipMalloc((void **)&t->d1, sizeof(bf16_t) * C * np));
        NTRY(hipMalloc((void **)&t->d2, sizeof(bf16_t) * C * np));
        t->cap = np;
    }
    const long need = (long)kDwBlocks * 4 * (C * T * C + C);
    if (need > t->partcap) {
        if (t->part) (void)hipFree(t->part);
        t->partcap = 0;
        NTRY(hipMalloc((void **)&t->part, sizeof(float) * need));
        t->partcap = need;
    }
    if (nimg > t->lcap) {
        if (t->lpart) (void)hipFree(t->lpart);
        t->lcap = 0;
        NTRY(hipMalloc((void **)&t->lpart, sizeof(double) * nimg));
        t->lcap = nimg;
    }
    return 0;
}

int refresh_bf16(trlnet *t)
{
    hipLaunchKernelGGL(k_bf16_params, dim3((C * T * C + NB - 1) / NB > (NPARAM + NB - 1) / NB
                                               ? (C * T * C + NB - 1) / NB
                                               : (NPARAM + NB - 1) / NB),
                       dim3(NB), 0, t->s, t->prm, t->wb, t->wt2, t->wt3);
    return hipGetLastError() == hipSuccess ? 0 : TRLNET_EHIP;
}

// Forward pass of the images g at pixel offset off of the workspace
// (activations are channel-last, pixel p at p * C).
int forward(trlnet *t, const Geo &g, long off = 0)
{
    const int gb = grid_px(g.np);
    const long oc = off * C;
    hipLaunchKernelGGL(k_fwd1, dim3(gb), dim3(NB), 0, t->s, g, t->x + off, t->wb + OW1, t->prm + OB1,
                       t->a1 + oc);
    hipLaunchKernelGGL(k_fwd32, dim3(gb), dim3(NB), 0, t->s, g, t->a1 + oc, t->wb + OW2, t->prm + OB2,
                       t->a2 + oc);
    hipLaunchKernelGGL(k_fwd32, dim3(gb), dim3(NB), 0, t->s, g, t->a2 + oc, t->wb + OW3, t->prm + OB3,
                       t->a3 + oc);
    hipLaunchKernelGGL(k_fwd4, dim3(gb), dim3(NB), 0, t->s, g, t->a3 + oc, t->wb + OW4, t->prm + OB4,
                       t->o + off);
    return hipGetLastError() == hipSuccess ? 0 : TRLNET_EHIP;
}

// The forward pass over groups back to back (inputs at t->x, outputs at
// t->o): one launch per layer for all groups, the same per-pixel
// arithmetic as forward() group by group.
int forward_groups(trlnet *t, int ngroups, const int *nimg, const int *h, const int *w)
{
    const size_t tab = sizeof(GroupGeo) * (size_t)ngroups;
    const size_t off_pre = (tab + 255) & ~(size_t)255;
    const size_t bytes = off_pre + sizeof(int) * (size_t)(ngroups + 1);
    if ((long)bytes > t->geocap) {
        if (t->hgeo) (void)hipHostFree(t->hgeo);
        if (t->dgeo) (void)hipFree(t->dgeo);
        t->hgeo = t->dgeo = nullptr;
        t->geocap = 0;
        NTRY(hipHostMalloc((void **)&t->hgeo, bytes, 0));
        NTRY(hipMalloc((void **)&t->dgeo, bytes));
        t->geocap = (long)bytes;
    }
    GroupGeo *gg = reinterpret_cast<GroupGeo *>(t->hgeo);
    int *pre = reinterpret_cast<int *>(t->hgeo + off_pre);
    long off = 0;
    pre[0] = 0;
    for (int i = 0; i < ngroups; ++i) {
        const Geo g{nimg[i], h[i], w[i], (long)nimg[i] * h[i] * w[i]};
        gg[i] = GroupGeo{g, off};
        const long nb = (long)pre[i] + grid_px(g.np);
        if (nb > 0x7fffffffL) return TRLNET_EINVAL;
        pre[i + 1] = (int)nb;
        off += g.np;
    }
    NTRY(hipMemcpyAsync(t->dgeo, t->hgeo, bytes, hipMemcpyHostToDevice, t->s));
    const GroupGeo *dg = reinterpret_cast<const GroupGeo *>(t->dgeo);
    const int *dp = reinterpret_cast<const int *>(t->dgeo + off_pre);
    const dim3 gb(pre[ngroups]);
    hipLaunchKernelGGL(k_fwd1_many, gb, dim3(NB), 0, t->s, dg, dp, ngroups, t->x, t->wb + OW1,
                       t->prm + OB1, t->a1);
    hipLaunchKernelGGL(k_fwd32_many, gb, dim3(NB), 0, t->s, dg, dp, ngroups, t->a1, t->wb + OW2,
                       t->prm + OB2, t->a2);
    hipLaunchKernelGGL(k_fwd32_many, gb, dim3(NB), 0, t->s, dg, dp, ngroups, t->a2, t->wb + OW3,
                       t->prm + OB3, t->a3);
    hipLaunchKernelGGL(k_fwd4_many, gb, dim3(NB), 0, t->s, dg, dp, ngroups, t->a3, t->wb + OW4,
                       t->prm + OB4, t->o);
    return hipGetLastError() == hipSuccess ? 0 : TRLNET_EHIP;
}

// loss and gradient into t->grad (device); *loss on the host
int backward(trlnet *t, const Geo &g, double *loss)
{
    const int gb = grid_px(g.np);
    hipLaunchKernelGGL(k_loss, dim3(g.nimg), dim3(NB), 0, t->s, g, t->o, t->tg, t->dO, t->lpart);
    // layer 4
    hipLaunchKernelGGL(k_dw_small, dim3(kDwBlocks), dim3(NB), 0, t->s, g, 4, (const bf16_t *)nullptr,
                       (const float *)nullptr, t->dO, t->a3, t->part);
    hipLaunchKernelGGL(k_reduce, dim3(2), dim3(NB), 0, t->s, t->part, kDwBlocks, C * T + C, t->part + (long)kDwBlocks * (C * T + C));
    NTRY(hipMemcpyAsync(t->grad + OW4, t->part + (long)kDwBlocks * (C * T + C), sizeof(float) * (C * T + 1),
                        hipMemcpyDeviceToDevice, t->s));
    hipLaunchKernelGGL(k_bwd4, dim3(gb), dim3(NB), 0, t->s, g, t->dO, t->wb + OW4, t->a3, t->d2);
    // layer 3: dY = d2 (for a3), X = a2
    const int NW = C * T * C + C;
    hipLaunchKernelGGL(k_dw32, dim3(kDwBlocks), dim3(NB), 0, t->s, g, t->d2, t->a2, t->part);
    hipLaunchKernelGGL(k_reduce, dim3((NW + NB - 1) / NB), dim3(NB), 0, t->s, t->part, kDwBlocks * 4, NW, t->grad + OW3);
    hipLaunchKernelGGL(k_bwd32, dim3(gb), dim3(NB), 0, t->s, g, t->d2, t->wt3, t->a2, t->d1);
    // layer 2: dY = d1 (for a2), X = a1
    hipLaunchKernelGGL(k_dw32, dim3(kDwBlocks), dim3(NB), 0, t->s, g, t->d1, t->a1, t->part);
    hipLaunchKernelGGL(k_reduce, dim3((NW + NB - 1) / NB), dim3(NB), 0, t->s, t->part, kDwBlocks * 4, NW, t->grad + OW2);
    hipLaunchKernelGGL(k_bwd32, dim3(gb), dim3(NB), 0, t->s, g, t->d1, t->wt2, t->a1, t->d2);
    // layer 1: dY = d2 (for a1), X = potential
    hipLaunchKernelGGL(k_dw_small, dim3(kDwBlocks), dim3(NB), 0, t->s, g, 1, t->d2, t->x,
                       (const float *)nullptr, (const bf16_t *)nullptr, t->part);
    hipLaunchKernelGGL(k_reduce, dim3(2), dim3(NB), 0, t->s, t->part, kDwBlocks, C * T + C, t->grad + OW1);
    if (hipGetLastError() != hipSuccess) return TRLNET_EHIP;
    std::vector<double> lp((size_t)g.nimg);
    NTRY(hipMemcpyAsync(lp.data(), t->lpart, sizeof(double) * g.nimg, hipMemcpyDeviceToHost, t->s));
    NTRY(hipStreamSynchronize(t->s));
    double l = 0.0;
    for (double v : lp) l += v;
    *loss = l;
    return 0;
}

int upload(trlnet *t, const Geo &g, const float *in, const float *target)
{
    if (int rc = ensure(t, g.np, g.nimg)) return rc;
    NTRY(hipMemcpyAsync(t->x, in, sizeof(float) * g.np, hipMemcpyHostToDevice, t->s));
    if (target) NTRY(hipMemcpyAsync(t->tg, target, sizeof(float) * g.np, hipMemcpyHostToDevice, t->s));
    return 0;
}

bool geo_ok(int nimg, int h, int w) { return nimg > 0 && h >= 3 && w >= 3 && (long)nimg * h * w < (1L << 40); }

}  // namespace

extern "C" {

int trlnet_param_count(void) { return NPARAM; }

int trlnet_create(trlnet **out, unsigned long long seed)
{
    if (!out) return TRLNET_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TRLNET_ENODEV;
    trlnet *t = new (std::nothrow) trlnet();
    if (!t) return TRLNET_ENOMEM;
    int rc = 0;
    if (hipStreamCreate(&t->s) != hipSuccess) rc = TRLNET_EHIP;
    const size_t pb = sizeof(float) * NPARAM;
    if (!rc && (hipMalloc((void **)&t->prm, pb) != hipSuccess || hipMalloc((void **)&t->grad, pb) != hipSuccess ||
                hipMalloc((void **)&t->m1, pb) != hipSuccess || hipMalloc((void **)&t->m2, pb) != hipSuccess ||
                hipMalloc((void **)&t->wb, sizeof(bf16_t) * NPARAM) != hipSuccess ||
                hipMalloc((void **)&t->wt2, sizeof(bf16_t) * C * T * C) != hipSuccess ||
                hipMalloc((void **)&t->wt3, sizeof(bf16_t) * C * T * C) != hipSuccess))
        rc = TRLNET_ENOMEM;
    if (!rc) {
        // He-normal weights from a 64-bit xorshift stream (Box-Muller), zero biases
        std::vector<float> h(NPARAM, 0.f);
        uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
        auto uni = [&]() {
            st ^= st >> 12;
            st ^= st << 25;
            st ^= st >> 27;
            return ((st * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
        };
        auto gauss = [&]() {
            const double u1 = uni() + 1e-300, u2 = uni();
            return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        };
        const struct { int off, n, fan_in; } L[4] = {
            {OW1, C * T, T}, {OW2, C * T * C, T * C}, {OW3, C * T * C, T * C}, {OW4, T * C, T * C}};
        for (const auto &l : L)
            for (int i = 0; i < l.n; ++i) h[(size_t)l.off + i] = (float)(gauss() * std::sqrt(2.0 / l.fan_in));
        if (hipMemcpy(t->prm, h.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(t->m1, 0, pb) != hipSuccess || hipMemset(t->m2, 0, pb) != hipSuccess ||
            hipMemset(t->grad, 0, pb) != hipSuccess)
            rc = TRLNET_EHIP;
    }
    if (!rc) rc = refresh_bf16(t);
    if (!rc && hipStreamSynchronize(t->s) != hipSuccess) rc = TRLNET_EHIP;
    if (rc) {
        trlnet_destroy(t);
        return rc;
    }
    *out = t;
    return 0;
}

int trlnet_destroy(trlnet *t)
{
    if (!t) return 0;
    if (t->s) (void)hipStreamSynchronize(t->s);
    for (void *p : {(void *)t->prm, (void *)t->grad, (void *)t->m1, (void *)t->m2, (void *)t->wb,
                    (void *)t->wt2, (void *)t->wt3, (void *)t->x, (void *)t->tg, (void *)t->o,
                    (void *)t->dO, (void *)t->a1, (void *)t->a2, (void *)t->a3, (void *)t->d1,
                    (void *)t->d2, (void *)t->part, (void *)t->lpart})
        if (p) (void)hipFree(p);
    if (t->hstage) (void)hipHostFree(t->hstage);
    if (t->hgeo) (void)hipHostFree(t->hgeo);
    if (t->dgeo) (void)hipFree(t->dgeo);
    if (t->s) (void)hipStreamDestroy(t->s);
    delete t;
    return 0;
}

int trlnet_get_params(trlnet *t, float *out, int n)
{
    if (!t || !out || n < NPARAM) return TRLNET_EINVAL;
    NTRY(hipMemcpyAsync(out, t->prm, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, t->s));
    NTRY(hipStreamSynchronize(t->s));
    return 0;
}

int trlnet_set_params(trlnet *t, const float *in, int n)
{
    if (!t || !in || n < NPARAM) return TRLNET_EINVAL;
    NTRY(hipMemcpyAsync(t->prm, in, sizeof(float) * NPARAM, hipMemcpyHostToDevice, t->s));
    if (int rc = refresh_bf16(t)) return rc;
    NTRY(hipStreamSynchronize(t->s));
    return 0;
}

int trlnet_forward(trlnet *t, int nimg, int h, int w, const float *in, float *out)
{
    if (!t || !in || !out || !geo_ok(nimg, h, w)) return TRLNET_EINVAL;
    const Geo g{nimg, h, w, (long)nimg * h * w};
    if (int rc = upload(t, g, in, nullptr)) return rc;
    if (int rc = forward(t, g)) return rc;
    NTRY(hipMemcpyAsync(out, t->o, sizeof(float) * g.np, hipMemcpyDeviceToHost, t->s));
    NTRY(hipStreamSynchronize(t->s));
    return 0;
}

int trlnet_forward_many(trlnet *t, int ngroups, const int *nimg, const int *h, const int *w,
                        const float *in, float *out)
{
    if (!t || ngroups < 1 || !nimg || !h || !w || !in || !out) return TRLNET_EINVAL;
    long total = 0;
    for (int i = 0; i < ngroups; ++i) {
        if (!geo_ok(nimg[i], h[i], w[i])) return TRLNET_EINVAL;
        total += (long)nimg[i] * h[i] * w[i];
    }
    if (int rc = ensure(t, total, 1)) return rc;
    if (total > t->hcap) {  // pinned staging: the copies run as DMA, asynchronously
        if (t->hstage) (void)hipHostFree(t->hstage);
        t->hstage = nullptr;
        t->hcap = 0;
        NTRY(hipHostMalloc((void **)&t->hstage, sizeof(float) * total, 0));
        t->hcap = total;
    }
    memcpy(t->hstage, in, sizeof(float) * total);
    NTRY(hipMemcpyAsync(t->x, t->hstage, sizeof(float) * total, hipMemcpyHostToDevice, t->s));
    if (int rc = forward_groups(t, ngroups, nimg, h, w)) return rc;
    NTRY(hipMemcpyAsync(t->hstage, t->o, sizeof(float) * total, hipMemcpyDeviceToHost, t->s));
    NTRY(hipStreamSynchronize(t->s));
    memcpy(out, t->hstage, sizeof(float) * total);
    return 0;
}

}  // extern "C"

namespace {

// Validates the groups, sizes the device buffers and the pinned staging
// for `total` input floats.
int many_prepare(trlnet *t, int ngroups, const int *nimg, const int *h, const int *w, long *total)
{
    long tot = 0;
    for (int i = 0; i < ngroups; ++i) {
        if (!geo_ok(nimg[i], h[i], w[i])) return TRLNET_EINVAL;
        tot += (long)nimg[i] * h[i] * w[i];
    }
    if (int rc = ensure(t, tot, 1)) return rc;
    if (tot > t->hcap) {
        if (t->hstage) (void)hipHostFree(t->hstage);
        t->hstage = nullptr;
        t->hcap = 0;
        NTRY(hipHostMalloc((void **)&t->hstage, sizeof(float) * tot, 0));
        t->hcap = tot;
    }
    *total = tot;
    return 0;
}

// Uploads the staged inputs and runs every group; outputs stay on the device.
int many_run(trlnet *t, int ngroups, const int *nimg, const int *h, const int *w, long total,
             const float **out_device)
{
    NTRY(hipMemcpyAsync(t->x, t->hstage, sizeof(float) * total, hipMemcpyHostToDevice, t->s));
    if (int rc = forward_groups(t, ngroups, nimg, h, w)) return rc;
    NTRY(hipStreamSynchronize(t->s));
    *out_device = t->o;
    return 0;
}

}  // namespace

extern "C" {

int trlnet_forward_many_device(trlnet *t, int ngroups, const int *nimg, const int *h, const int *w,
                               const float *in, const float **out_device)
{
    if (!t || ngroups < 1 || !nimg || !h || !w || !in || !out_device) return TRLNET_EINVAL;
    long total = 0;
    if (int rc = many_prepare(t, ngroups, nimg, h, w, &total)) return rc;
    memcpy(t->hstage, in, sizeof(float) * total);
    return many_run(t, ngroups, nimg, h, w, total, out_device);
}

int trlnet_forward_many_device_f64(trlnet *t, int ngroups, const int *nimg, const int *h,
                                   const int *w, const double *const *img, double scale,
                                   const float **out_device)
{
    if (!t || ngroups < 1 || !nimg || !h || !w || !img || !out_device) return TRLNET_EINVAL;
    long total = 0;
    if (int rc = many_prepare(t, ngroups, nimg, h, w, &total)) return rc;
    // image i: its pointer, size and offset in the staging buffer
    std::vector<const double *> src;
    std::vector<long> len, off;
    long o = 0;
    for (int g = 0; g < ngroups; ++g)
        for (int k = 0; k < nimg[g]; ++k) {
            const long np = (long)h[g] * w[g];
            if (!img[src.size()]) return TRLNET_EINVAL;
            src.push_back(img[src.size()]);
            len.push_back(np);
            off.push_back(o);
            o += np;
        }
    // fl32(x * scale) into the staging, images dealt to up to 16 threads
    // by a shared counter
    const int nimgs = (int)src.size();
    unsigned hw = std::thread::hardware_concurrency();
    const int nthr = (int)std::min<long>(std::min<unsigned>(hw ? hw : 1, 16), std::max(1, nimgs));
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int i; (i = next.fetch_add(1, std::memory_order_relaxed)) < nimgs;) {
            const double *p = src[i];
            float *q = t->hstage + off[i];
            for (long j = 0; j < len[i]; ++j) q[j] = (float)(p[j] * scale);
        }
    };
    std::vector<std::thread> th;
    for (int k = 1; k < nthr; ++k) th.emplace_back(work);
    work();
    for (auto &x : th) x.join();
    return many_run(t, ngroups, nimg, h, w, total, out_device);
}

int trlnet_loss_grad(trlnet *t, int nimg, int h, int w, const float *in, const float *target,
                     float *loss, float *grad)
{
    if (!t || !in || !target || !geo_ok(nimg, h, w)) return TRLNET_EINVAL;
    const Geo g{nimg, h, w, (long)nimg * h * w};
    if (int rc = upload(t, g, in, target)) return rc;
    if (int rc = forward(t, g)) return rc;
    double l = 0.0;
    if (int rc = backward(t, g, &l)) return rc;
    if (loss) *loss = (float)l;
    if (grad) {
        NTRY(hipMemcpyAsync(grad, t->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, t->s));
        NTRY(hipStreamSynchronize(t->s));
    }
    return 0;
}

int trlnet_train_step(trlnet *t, int nimg, int h, int w, const float *in, const float *target,
                      float lr, float *loss)
{
    float l = 0.f;
    if (int rc = trlnet_loss_grad(t, nimg, h, w, in, target, &l, nullptr)) return rc;
    ++t->step;
    const float b1t = 1.f - std::pow(0.9f, (float)t->step), b2t = 1.f - std::pow(0.999f, (float)t->step);
    hipLaunchKernelGGL(k_adam, dim3((NPARAM + NB - 1) / NB), dim3(NB), 0, t->s, t->prm, t->grad, t->m1,
                       t->m2, NPARAM, lr, b1t, b2t);
    if (int rc = refresh_bf16(t)) return rc;
    NTRY(hipStreamSynchronize(t->s));
    if (loss) *loss = l;
    return 0;
}

}  // extern "C"
