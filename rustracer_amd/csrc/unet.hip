// unet.hip — the learned denoiser: a small convolutional U-Net on the displayed
// (accumulated) image, with hand-written BF16 MFMA convolution kernels for the
// forward pass and both backward passes, FP32 accumulation, and Adam, for gfx950.
//
// This part has no counterpart in the reference renderer; its inputs are the
// renderer's own buffers (accumulation, G-buffer) plus a first-hit albedo pass.
//
// Network (NHWC bf16 activations, channel counts multiples of 16/32):
//   F  (16: log1p rgb, albedo, normal, depth, sample count, 5 zero)
//   A1 = relu(conv3x3(F, 16->32))           A2 = relu(conv3x3(A1, 32->32)) -> CAT[0:32]
//   P  = avgpool2(A2)                       B1 = relu(conv3x3(P, 32->64))  B2 = relu(conv3x3(B1, 64->64))
//   CAT[32:96] = upsample2(B2)              D1 = relu(conv3x3(CAT, 96->32)) D2 = relu(conv3x3(D1, 32->32))
//   pred = log1p(noisy) + conv1x1(D2, 32->3) + b   (FP32), image = expm1(pred)
// Loss: mean over pixels and rgb of (pred - log1p(target))^2.
//
// Convolution as implicit GEMM on v_mfma_f32_32x32x16_bf16: for the forward
// and backward-data passes M = 32 output channels, N = 32 consecutive pixels of
// one row, K = 9 taps x input channels in steps of 16. Lane layouts (checked on
// the device): A lane l holds row l%32, k = (l/32)*8 .. +7; B lane l holds
// column l%32, the same k; accumulator element e of lane l is row (e/4)*8 +
// (l/32)*4 + e%4, column l%32.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/rustracer_hip.h"
#include "capi_util.h"
#include "unet.h"

namespace rtu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef uint16_t bf16;  // storage type

__device__ __forceinline__ bf16 f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ bf16x8 load8(const bf16 *p) {
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u)); }

// Elementwise layers work on 8-channel groups (16-byte loads and stores), one
// thread per (pixel, group); grid.y = image row (n * rows + y), grid.x covers
// the row's pixels x groups. All index math is 32-bit.
__device__ __forceinline__ void unpack8(uint4 v, float f[8]) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        f[2 * i] = __uint_as_float(w[i] << 16);
        f[2 * i + 1] = __uint_as_float(w[i] & 0xFFFF0000u);
    }
}
__device__ __forceinline__ uint4 pack8(const float f[8]) {
    uint32_t w[4];
#pragma unroll
    for (int i = 0; i < 4; i++) w[i] = (uint32_t)f2bf(f[2 * i]) | ((uint32_t)f2bf(f[2 * i + 1]) << 16);
    return make_uint4(w[0], w[1], w[2], w[3]);
}
__device__ __forceinline__ uint4 ld16(const bf16 *p) { return *reinterpret_cast<const uint4 *>(p); }
__device__ __forceinline__ void st16(bf16 *p, uint4 v) { *reinterpret_cast<uint4 *>(p) = v; }


// One 3x3 convolution (stride 1, zero padding), optionally with bias and ReLU,
// or a mask (x > 0) of another tensor at the output position (the ReLU
// derivative in the backward-data pass). X: [N][H][W][xs] with CIN channels at
// offset 0; Wt: [9][cout][CIN]; Y: [N][H][W][ys], cout = 32 * NCO channels at
// offset yoff.
//
// A workgroup of four wavefronts takes 32 pixels of kConvRows consecutive rows;
// the input rows y0-1 .. y0+kConvRows, pixels x0-1 .. x0+32 are staged once in
// LDS (pixel stride CIN + 8 to spread banks) and serve all 9 taps; the layer's
// weights are staged in LDS beside them (row stride CIN + 8), so the MFMA loop
// waits on no global load. Both stagings issue up to eight 16-byte loads per
// thread before writing any of them. Each wavefront computes
// kConvRowsPerWave rows for all output channels. The weights are the MFMA's A
// operand (row = output channel) and the pixels its B operand (column =
// pixel), so every weight fragment, read once per tap and input-channel step,
// feeds kConvRowsPerWave MFMAs, and each lane ends with 4
// consecutive output channels of one pixel per accumulator group: the
// epilogue stores 8 bytes at a time. The workgroups are persistent (as many as
// are resident at once): each stages the weights once, then takes the tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... of the ceil(W/32) x
// ceil(H/kConvRows) x N tiles.
//
// The inference pass folds the pool and the upsample into the staging of the
// layers that read them (the training pass keeps P and CAT for its backward):
// SRC_POOL reads X at twice the output resolution and stages the 2x2 average
// of pool_kernel; SRC_CONCAT stages channels [0, c1) from X and [c1, CIN) from
// X2 (half resolution, stride x2s) at (y/2, x/2), as upsample_kernel writes
// them into CAT. The staged values are bit-identical to those kernels' outputs.
constexpr int kConvRowsPerWave = 2;
constexpr int kConvRows = 4 * kConvRowsPerWave;
enum ConvSrc { SRC_PLAIN = 0, SRC_POOL = 1, SRC_CONCAT = 2 };
template <int CIN, int NCO, int SRC = SRC_PLAIN>
__global__ __launch_bounds__(256) void conv3x3_kernel(const bf16 *__restrict__ X, int xs,
                                                      const bf16 *__restrict__ Wt, const float *__restrict__ bias,
                                                      bf16 *__restrict__ Y, int ys, int yoff, int relu,
                                                      const bf16 *__restrict__ mask, int ms, int H, int W,
                                                      const bf16 *__restrict__ X2 = nullptr, int x2s = 0,
                                                      int c1 = 0, int N = 1) {
    constexpr int R = kConvRowsPerWave;
    constexpr int PS = CIN + 8;  // LDS pixel stride (elements)
    constexpr int GROUPS = CIN / 8;
    constexpr int COUT = 32 * NCO;
    constexpr int WPS = CIN + 8;  // LDS weight row stride (elements)
    // dynamic LDS: the input tile [kConvRows + 2][34][PS], then the layer's
    // weights [9][COUT][WPS] (staged once, so the MFMA loop reads only LDS)
    extern __shared__ __attribute__((aligned(16))) bf16 tile[];
    bf16 *wl = tile + (kConvRows + 2) * 34 * PS;
    const int l = threadIdx.x % 64, wv = threadIdx.x / 64;
    const int tiles_x = (W + 31) / 32, tiles_y = (H + kConvRows - 1) / kConvRows;
    const int total = tiles_x * tiles_y * N;
    {
        constexpr int WITEMS = 9 * COUT * GROUPS;
        constexpr int WPER = (WITEMS + 255) / 256;
        constexpr int WB = 8;
#pragma unroll
        for (int k0 = 0; k0 < WPER; k0 += WB) {
            uint4 v[WB];
#pragma unroll
            for (int k = 0; k < WB; k++) {
                const int i = threadIdx.x + (k0 + k) * 256;
                if (k0 + k < WPER && i < WITEMS) v[k] = ld16(Wt + (size_t)(i / GROUPS) * CIN + (i % GROUPS) * 8);
            }
#pragma unroll
            for (int k = 0; k < WB; k++) {
                const int i = threadIdx.x + (k0 + k) * 256;
                if (k0 + k < WPER && i < WITEMS)
                    *reinterpret_cast<uint4 *>(wl + (i / GROUPS) * WPS + (i % GROUPS) * 8) = v[k];
            }
        }
    }
    // Tile staging: thread items threadIdx.x + 256 k. The loads of the next
    // tile are issued into registers before the current tile's MFMAs and
    // written to LDS after them, so their latency overlaps the compute.
    constexpr int ITEMS = (kConvRows + 2) * 34 * GROUPS;
    constexpr int PER = (ITEMS + 255) / 256;
    constexpr int NL = SRC == SRC_POOL ? 4 : 1;
    uint4 raw[PER][NL];
    auto fetch = [&](int t) {
        const int bx = t % tiles_x, rest = t / tiles_x;
        const int x0 = bx * 32, y0 = (rest % tiles_y) * kConvRows, n = rest / tiles_y;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = threadIdx.x + k * 256;
            const int g = i % GROUPS, q = i / GROUPS;
            const int px = q % 34, row = q / 34;
            const int yy = y0 - 1 + row, xx = x0 - 1 + px;
#pragma unroll
            for (int e = 0; e < NL; e++) raw[k][e] = make_uint4(0u, 0u, 0u, 0u);
            if (i < ITEMS && yy >= 0 && yy < H && xx >= 0 && xx < W) {
                if (SRC == SRC_POOL) {
                    const size_t p00 = ((size_t)(n * 2 * H + 2 * yy) * (2 * W) + 2 * xx) * xs + g * 8;
                    const size_t down = (size_t)(2 * W) * xs;
                    raw[k][0] = ld16(X + p00);
                    raw[k][NL > 1 ? 1 : 0] = ld16(X + p00 + xs);
                    raw[k][NL > 2 ? 2 : 0] = ld16(X + p00 + down);
                    raw[k][NL > 3 ? 3 : 0] = ld16(X + p00 + down + xs);
                } else if (SRC == SRC_CONCAT && g * 8 >= c1) {
                    raw[k][0] = ld16(X2 + ((size_t)(n * (H / 2) + yy / 2) * (W / 2) + xx / 2) * x2s + (g * 8 - c1));
                } else {
                    raw[k][0] = ld16(X + ((size_t)(n * H + yy) * W + xx) * xs + g * 8);
                }
            }
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = threadIdx.x + k * 256;
            if (i < ITEMS) {
                const int g = i % GROUPS, q = i / GROUPS;
                const int px = q % 34, row = q / 34;
                uint4 v = raw[k][0];
                if (SRC == SRC_POOL) {
                    // zero loads (outside the image) give the zero that the pool
                    // of a padded position is: 0.25 * (0 + 0 + 0 + 0) = 0
                    float a[8], b[8], d[8], e[8], o[8];
                    unpack8(raw[k][0], a);
                    unpack8(raw[k][NL > 1 ? 1 : 0], b);
                    unpack8(raw[k][NL > 2 ? 2 : 0], d);
                    unpack8(raw[k][NL > 3 ? 3 : 0], e);
#pragma unroll
                    for (int c8 = 0; c8 < 8; c8++) o[c8] = 0.25f * (((a[c8] + b[c8]) + d[c8]) + e[c8]);
                    v = pack8(o);
                }
                *reinterpret_cast<uint4 *>(tile + (row * 34 + px) * PS + g * 8) = v;
            }
        }
    };
    int t = blockIdx.x;
    if (t < total) fetch(t);
    for (; t < total; t += gridDim.x) {
        __syncthreads();  // every wavefront is done with the previous tile (and the weights are in)
        commit();
        __syncthreads();
        if (t + (int)gridDim.x < total) fetch(t + (int)gridDim.x);
        const int bx = t % tiles_x, rest = t / tiles_x;
        const int x0 = bx * 32, y0 = (rest % tiles_y) * kConvRows, n = rest / tiles_y;
        const int c = l % 32, kg = (l / 32) * 8;
        floatx16 acc[R][NCO];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int j = 0; j < NCO; j++) acc[r][j] = floatx16{};
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const bf16 *wp = wl + (tap * COUT + c) * WPS + kg;
#pragma unroll
            for (int c0 = 0; c0 < CIN; c0 += 16) {
                bf16x8 a[NCO], b[R];
#pragma unroll
                for (int j = 0; j < NCO; j++) a[j] = load8(wp + j * 32 * WPS + c0);
#pragma unroll
                for (int r = 0; r < R; r++)
                    b[r] = load8(tile + ((wv * R + r + tap / 3) * 34 + c + tap % 3) * PS + kg + c0);
#pragma unroll
                for (int r = 0; r < R; r++)
#pragma unroll
                    for (int j = 0; j < NCO; j++)
                        acc[r][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[j], b[r], acc[r][j], 0, 0, 0);
            }
        }
        const int x = x0 + c;
        if (x < W) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int y = y0 + wv * R + r;
                if (y >= H) break;
                const size_t p = (size_t)(n * H + y) * W + x;
#pragma unroll
                for (int j = 0; j < NCO; j++) {
#pragma unroll
                    for (int g = 0; g < 4; g++) {
                        // accumulator elements 4g .. 4g+3: output channels co .. co+3 of pixel x
                        const int co = j * 32 + g * 8 + (l / 32) * 4;
                        float v[4];
#pragma unroll
                        for (int e = 0; e < 4; e++) v[e] = acc[r][j][4 * g + e];
                        if (bias) {
                            const float4 bv = *reinterpret_cast<const float4 *>(bias + co);
                            v[0] += bv.x;
                            v[1] += bv.y;
                            v[2] += bv.z;
                            v[3] += bv.w;
                        }
                        if (relu)
#pragma unroll
                            for (int e = 0; e < 4; e++) v[e] = fmaxf(v[e], 0.0f);
                        if (mask) {
                            const uint2 mv = *reinterpret_cast<const uint2 *>(mask + p * ms + co);
                            if (!(__uint_as_float(mv.x << 16) > 0.0f)) v[0] = 0.0f;
                            if (!(__uint_as_float(mv.x & 0xFFFF0000u) > 0.0f)) v[1] = 0.0f;
                            if (!(__uint_as_float(mv.y << 16) > 0.0f)) v[2] = 0.0f;
                            if (!(__uint_as_float(mv.y & 0xFFFF0000u) > 0.0f)) v[3] = 0.0f;
                        }
                        const uint2 o = make_uint2((uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16),
                                                   (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16));
                        *reinterpret_cast<uint2 *>(Y + p * ys + yoff + co) = o;
                    }
                }
            }
        }
    }
}

// Weight gradient of a 3x3 convolution: gW[tap][co][ci] += sum over pixels of
// G[p][co] * X[p + tap][ci]. Each workgroup takes one (tap, 32 co, 32 ci) block
// and a strided set of 32-pixel row segments; G and X segments are staged
// transposed in LDS ([channel][pixel]) so MFMA fragments are 16-byte reads.
// Grid: (segment groups, 9, cout/32 * ceil(cin/32)).
__global__ __launch_bounds__(64) void conv3x3_wgrad_kernel(const bf16 *__restrict__ G, int gs, int cout,
                                                           const bf16 *__restrict__ X, int xs, int cin,
                                                           float *__restrict__ gW, int N, int H, int W) {
    __shared__ __attribute__((aligned(16))) bf16 gt[32][32];  // [co][pixel]
    __shared__ __attribute__((aligned(16))) bf16 xt[32][32];  // [ci][pixel]
    const int l = threadIdx.x;
    const int tap = blockIdx.y;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int nci = (cin + 31) / 32;
    const int co0 = (blockIdx.z / nci) * 32, ci0 = (blockIdx.z % nci) * 32;
    const int segs_x = (W + 31) / 32;
    const int total = N * H * segs_x;
    const int lp = l / 2, lh = (l % 2) * 16;  // staging: pixel, 16-channel half
    floatx16 acc = {};
    for (int s = blockIdx.x; s < total; s += gridDim.x) {
        const int sx = s % segs_x, ny = s / segs_x;
        const int y = ny % H, n = ny / H;
        const int x = sx * 32 + lp;
        {
            uint4 v0 = make_uint4(0u, 0u, 0u, 0u), v1 = v0;
            if (x < W) {
                const bf16 *gp = G + ((size_t)(n * H + y) * W + x) * gs + co0 + lh;
                v0 = *reinterpret_cast<const uint4 *>(gp);
                v1 = *reinterpret_cast<const uint4 *>(gp + 8);
            }
            const bf16 *e0 = reinterpret_cast<const bf16 *>(&v0), *e1 = reinterpret_cast<const bf16 *>(&v1);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                gt[lh + j][lp] = e0[j];
                gt[lh + 8 + j][lp] = e1[j];
            }
        }
        {
            const int yy = y + dy, xx = x + dx;
            uint4 v0 = make_uint4(0u, 0u, 0u, 0u), v1 = v0;
            if (x < W && yy >= 0 && yy < H && xx >= 0 && xx < W) {
                const bf16 *xp = X + ((size_t)(n * H + yy) * W + xx) * xs;
                if (ci0 + lh < cin) v0 = *reinterpret_cast<const uint4 *>(xp + ci0 + lh);
                if (ci0 + lh + 8 < cin) v1 = *reinterpret_cast<const uint4 *>(xp + ci0 + lh + 8);
            }
            const bf16 *e0 = reinterpret_cast<const bf16 *>(&v0), *e1 = reinterpret_cast<const bf16 *>(&v1);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                xt[lh + j][lp] = e0[j];
                xt[lh + 8 + j][lp] = e1[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            const bf16x8 a = load8(&gt[l % 32][ks * 16 + (l / 32) * 8]);
            const bf16x8 b = load8(&xt[l % 32][ks * 16 + (l / 32) * 8]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int ci = ci0 + l % 32;
    if (ci >= cin) return;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int co = co0 + (e / 4) * 8 + (l / 32) * 4 + e % 4;
        atomicAdd(&gW[((size_t)tap * cout + co) * cin + ci], acc[e]);
    }
}

// Bias gradient: gb[c] += sum over pixels of G[p][c]. Thread t of a block
// takes channel group t % (c/8) (16-byte loads) for pixels t / (c/8) + k *
// lanes, grid-strided; the block then reduces its lanes in LDS and adds one
// value per channel. c is a multiple of 8 and at most 256.
__global__ __launch_bounds__(256) void bias_grad_kernel(const bf16 *__restrict__ G, int gs, int c, size_t n_px,
                                                        float *__restrict__ gb) {
    __shared__ float part[256 * 8];
    const int groups = c / 8, lanes = 256 / groups;
    const int g = threadIdx.x % groups, lane = threadIdx.x / groups;
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (lane < lanes)
        for (size_t p = (size_t)blockIdx.x * lanes + lane; p < n_px; p += (size_t)gridDim.x * lanes) {
            float v[8];
            unpack8(*reinterpret_cast<const uint4 *>(G + p * gs + g * 8), v);
#pragma unroll
            for (int i = 0; i < 8; i++) acc[i] += v[i];
        }
#pragma unroll
    for (int i = 0; i < 8; i++) part[threadIdx.x * 8 + i] = acc[i];
    __syncthreads();
    if ((int)threadIdx.x < c) {
        const int gg = threadIdx.x / 8, i = threadIdx.x % 8;
        float s = 0.0f;
        for (int l = 0; l < lanes; l++) s += part[(l * groups + gg) * 8 + i];
        atomicAdd(&gb[threadIdx.x], s);
    }
}

// 2x2 average pool of channels [0, c) of X (stride xs) into P (stride c), half
// resolution; grid.y = n * H/2 + y2.
__global__ __launch_bounds__(256) void pool_kernel(const bf16 *__restrict__ X, int xs, int c, bf16 *__restrict__ P,
                                                   int H, int W) {
    const int g = c / 8, w2 = W / 2, h2 = H / 2;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= w2 * g) return;
    const int x = t / g, cg = (t % g) * 8;
    const int n = blockIdx.y / h2, y = blockIdx.y % h2;
    const uint32_t p00 = ((uint32_t)(n * H + 2 * y) * W + 2 * x) * xs + cg;
    float a[8], b[8], d[8], e[8], o[8];
    unpack8(ld16(X + p00), a);
    unpack8(ld16(X + p00 + xs), b);
    unpack8(ld16(X + p00 + (uint32_t)W * xs), d);
    unpack8(ld16(X + p00 + (uint32_t)W * xs + xs), e);
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = 0.25f * (((a[i] + b[i]) + d[i]) + e[i]);
    st16(P + ((uint32_t)blockIdx.y * w2 + x) * c + cg, pack8(o));
}

// Nearest 2x upsample of B (stride c, half resolution) into channels [off, off+c)
// of Y (stride ys); grid.y = n * H + y.
__global__ __launch_bounds__(256) void upsample_kernel(const bf16 *__restrict__ B, int c, bf16 *__restrict__ Y, int ys,
                                                       int off, int H, int W) {
    const int g = c / 8;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= W * g) return;
    const int x = t / g, cg = (t % g) * 8;
    const int n = blockIdx.y / H, y = blockIdx.y % H;
    const uint4 v = ld16(B + ((uint32_t)(n * (H / 2) + y / 2) * (W / 2) + x / 2) * c + cg);
    st16(Y + ((uint32_t)blockIdx.y * W + x) * ys + off + cg, v);
}

// Backward of the pool into A2's gradient: gA[p] = gCat[p] + (gP[p/2] / 4) * (A[p] > 0).
// gCat already carries the mask; the pool branch gets it here. grid.y = n * H + y.
__global__ __launch_bounds__(256) void pool_back_kernel(const bf16 *__restrict__ gcat, int cs, const bf16 *__restrict__ gp,
                                                        const bf16 *__restrict__ A, int as, int c, bf16 *__restrict__ ga,
                                                        int H, int W) {
    const int g = c / 8;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= W * g) return;
    const int x = t / g, cg = (t % g) * 8;
    const int n = blockIdx.y / H, y = blockIdx.y % H;
    const uint32_t p = (uint32_t)blockIdx.y * W + x;
    float gc[8], pv[8], av[8], o[8];
    unpack8(ld16(gcat + p * cs + cg), gc);
    unpack8(ld16(gp + ((uint32_t)(n * (H / 2) + y / 2) * (W / 2) + x / 2) * c + cg), pv);
    unpack8(ld16(A + p * as + cg), av);
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = gc[i] + (0.25f * pv[i]) * (av[i] > 0.0f ? 1.0f : 0.0f);
    st16(ga + p * c + cg, pack8(o));
}

// Backward of the nearest upsample: gB[q] = sum of the 4 children's gY[p][off ..];
// grid.y = n * H/2 + y2.
__global__ __launch_bounds__(256) void upsample_back_kernel(const bf16 *__restrict__ gy, int ys, int off, int c,
                                                            bf16 *__restrict__ gb, int H, int W) {
    const int g = c / 8, w2 = W / 2, h2 = H / 2;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= w2 * g) return;
    const int x = t / g, cg = (t % g) * 8;
    const int n = blockIdx.y / h2, y = blockIdx.y % h2;
    const uint32_t p00 = ((uint32_t)(n * H + 2 * y) * W + 2 * x) * ys + off + cg;
    float a[8], b[8], d[8], e[8], o[8];
    unpack8(ld16(gy + p00), a);
    unpack8(ld16(gy + p00 + ys), b);
    unpack8(ld16(gy + p00 + (uint32_t)W * ys), d);
    unpack8(ld16(gy + p00 + (uint32_t)W * ys + ys), e);
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = ((a[i] + b[i]) + d[i]) + e[i];
    st16(gb + ((uint32_t)blockIdx.y * w2 + x) * c + cg, pack8(o));
}

// Output layer (1x1, 32 -> 3, FP32) plus the residual: pred = noisy + b + W·D.
// With a target: loss partial sums, and the backward pass: gD = (W^T dpred) * (D > 0)
// as bf16, gW += dpred D^T, gb += dpred (block reductions, then atomics).
constexpr int kOutC = 32;
template <bool TRAIN>
__global__ __launch_bounds__(256) void out_layer_kernel(const bf16 *__restrict__ D, const float *__restrict__ noisy,
                                                        const float *__restrict__ w, const float *__restrict__ b,
                                                        float *__restrict__ pred, const float *__restrict__ target,
                                                        size_t n_px, float grad_scale, bf16 *__restrict__ gd,
                                                        float *__restrict__ gw, float *__restrict__ gbias,
                                                        double *__restrict__ loss) {
    __shared__ float red[256];
    __shared__ float gacc[3 * kOutC + 3];
    for (int i = threadIdx.x; i < 3 * kOutC + 3; i += 256) gacc[i] = 0.0f;
    __syncthreads();
    float lsum = 0.0f;
    float lw[TRAIN ? 3 * kOutC + 3 : 1];
    if (TRAIN)
        for (int i = 0; i < 3 * kOutC + 3; i++) lw[i] = 0.0f;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < n_px; p += (size_t)gridDim.x * 256) {
        float d[kOutC];
#pragma unroll
        for (int i = 0; i < kOutC; i += 8) {
            const uint4 v = *reinterpret_cast<const uint4 *>(D + p * kOutC + i);
            const bf16 *e = reinterpret_cast<const bf16 *>(&v);
#pragma unroll
            for (int j = 0; j < 8; j++) d[i + j] = bf2f(e[j]);
        }
        float o[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float s = b[c];
#pragma unroll
            for (int i = 0; i < kOutC; i++) s += w[c * kOutC + i] * d[i];
            o[c] = noisy[p * 3 + c] + s;
            pred[p * 3 + c] = o[c];
        }
        if (TRAIN) {
            float g[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float diff = o[c] - target[p * 3 + c];
                lsum += diff * diff;
                g[c] = grad_scale * diff;
                lw[3 * kOutC + c] += g[c];
            }
#pragma unroll
            for (int i = 0; i < kOutC; i += 8) {
                uint32_t packed[4];
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int k = i + j;
                    float s = 0.0f;
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        s += w[c * kOutC + k] * g[c];
                        lw[c * kOutC + k] += g[c] * d[k];
                    }
                    const uint32_t h = f2bf(d[k] > 0.0f ? s : 0.0f);
                    if (j % 2 == 0) packed[j / 2] = h;
                    else packed[j / 2] |= h << 16;
                }
                *reinterpret_cast<uint4 *>(gd + p * kOutC + i) = make_uint4(packed[0], packed[1], packed[2], packed[3]);
            }
        }
    }
    if (!TRAIN) return;
    for (int i = 0; i < 3 * kOutC + 3; i++) atomicAdd(&gacc[i], lw[i]);
    red[threadIdx.x] = lsum;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, (double)red[0]);
    for (int i = threadIdx.x; i < 3 * kOutC; i += 256) atomicAdd(&gw[i], gacc[i]);
    if (threadIdx.x < 3) atomicAdd(&gbias[threadIdx.x], gacc[3 * kOutC + threadIdx.x]);
}

// Features from the renderer's buffers: accumulation (float4), G-buffer
// (normal, depth), albedo (float4), sample count. F: 16 bf16 per pixel; noisy:
// log1p of the accumulation (FP32 rgb).
__global__ __launch_bounds__(256) void features_kernel(const float4 *__restrict__ accum, const float4 *__restrict__ gbuf,
                                                       const float4 *__restrict__ albedo, float frames_feature,
                                                       size_t n_px, bf16 *__restrict__ F, float *__restrict__ noisy) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n_px) return;
    const float4 a = accum[p], g = gbuf[p], al = albedo[p];
    const float lr = log1pf(fmaxf(a.x, 0.0f)), lg = log1pf(fmaxf(a.y, 0.0f)), lb = log1pf(fmaxf(a.z, 0.0f));
    noisy[p * 3 + 0] = lr;
    noisy[p * 3 + 1] = lg;
    noisy[p * 3 + 2] = lb;
    const float depth = g.w < 1e29f ? log1pf(fminf(fmaxf(g.w, 0.0f), 1e4f)) * 0.25f : -1.0f;
    const float f[16] = {lr, lg, lb, al.x, al.y, al.z, g.x, g.y, g.z, depth, frames_feature, 0, 0, 0, 0, 0};
    uint32_t o[8];
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = (uint32_t)f2bf(f[2 * i]) | ((uint32_t)f2bf(f[2 * i + 1]) << 16);
    uint4 *dst = reinterpret_cast<uint4 *>(F + p * 16);
    dst[0] = make_uint4(o[0], o[1], o[2], o[3]);
    dst[1] = make_uint4(o[4], o[5], o[6], o[7]);
}

// log1p of a float4 image's rgb (training targets).
__global__ __launch_bounds__(256) void log_target_kernel(const float4 *__restrict__ img, size_t n_px,
                                                         float *__restrict__ out) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n_px) return;
    const float4 a = img[p];
    out[p * 3 + 0] = log1pf(fmaxf(a.x, 0.0f));
    out[p * 3 + 1] = log1pf(fmaxf(a.y, 0.0f));
    out[p * 3 + 2] = log1pf(fmaxf(a.z, 0.0f));
}

__global__ __launch_bounds__(256) void image_kernel(const float *__restrict__ pred, size_t n_px, float4 *__restrict__ out) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n_px) return;
    out[p] = make_float4(expm1f(pred[p * 3]), expm1f(pred[p * 3 + 1]), expm1f(pred[p * 3 + 2]), 1.0f);
}

// Crops of the training set into a batch: sample k of the batch is the
// (cw x ch) window at (ox[k], oy[k]) of image img[k] (all images iw x ih).
__global__ __launch_bounds__(256) void crop_kernel(const bf16 *__restrict__ F, const float *__restrict__ noisy,
                                                   const float *__restrict__ target, const int *__restrict__ sel,
                                                   int iw, int ih, int cw, int chh, bf16 *__restrict__ bF,
                                                   float *__restrict__ bnoisy, float *__restrict__ btarget) {
    const int k = blockIdx.y;
    const int img = sel[3 * k], ox = sel[3 * k + 1], oy = sel[3 * k + 2];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < cw * chh; i += gridDim.x * 256) {
        const int x = i % cw, y = i / cw;
        const size_t src = ((size_t)img * ih + oy + y) * iw + ox + x;
        const size_t dst = (size_t)k * cw * chh + i;
        const uint4 *s = reinterpret_cast<const uint4 *>(F + src * 16);
        uint4 *d = reinterpret_cast<uint4 *>(bF + dst * 16);
        d[0] = s[0];
        d[1] = s[1];
        for (int c = 0; c < 3; c++) {
            bnoisy[dst * 3 + c] = noisy[src * 3 + c];
            btarget[dst * 3 + c] = target[src * 3 + c];
        }
    }
}

// Adam on an FP32 parameter array, then the bf16 copies: forward layout
// [tap][co][ci] (same as the master) and, when wt is given, the backward-data
// layout wt[8 - tap][ci][co] (flipped taps, transposed channels).
__global__ __launch_bounds__(256) void adam_kernel(float *__restrict__ p, const float *__restrict__ g,
                                                   float *__restrict__ m, float *__restrict__ v, size_t n, float lr,
                                                   float b1, float b2, float c1, float c2, float eps,
                                                   bf16 *__restrict__ wf, bf16 *__restrict__ wt, int cout, int cin) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        p[i] = pi;
        if (wf) {
            const bf16 h = f2bf(pi);
            wf[i] = h;
            if (wt) {
                const int ci = (int)(i % cin), co = (int)((i / cin) % cout), tap = (int)(i / ((size_t)cin * cout));
                wt[((size_t)(8 - tap) * cin + ci) * cout + co] = h;
            }
        }
    }
}

__global__ __launch_bounds__(256) void weights_bf16_kernel(const float *__restrict__ p, size_t n, bf16 *__restrict__ wf,
                                                           bf16 *__restrict__ wt, int cout, int cin) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const bf16 h = f2bf(p[i]);
        wf[i] = h;
        if (wt) {
            const int ci = (int)(i % cin), co = (int)((i / cin) % cout), tap = (int)(i / ((size_t)cin * cout));
            wt[((size_t)(8 - tap) * cin + ci) * cout + co] = h;
        }
    }
}

__global__ __launch_bounds__(256) void zero_kernel(float *__restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0.0f;
}

}  // namespace rtu

// ─────────────────────────────── host side ───────────────────────────────

namespace {

using rt::ApiError;

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw ApiError(RT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T *dalloc(size_t count, const char *what) {
    void *p = nullptr;
    hip_check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), what);
    return static_cast<T *>(p);
}

unsigned blocks_for(size_t n, unsigned cap = 8192) { return (unsigned)std::min<size_t>((n + 255) / 256, cap); }

// compute units of the current device
int cu_count() {
    int dev = 0, cus = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    return std::max(cus, 1);
}

struct ConvLayer {
    int cin = 0, cout = 0;
    size_t nw() const { return (size_t)9 * cin * cout; }
};
// conv1..conv6 of the network described at the top of this file
const ConvLayer kLayers[6] = {{16, 32}, {32, 32}, {32, 64}, {64, 64}, {96, 32}, {32, 32}};
constexpr int kFeat = 16;

// Activations (and their gradients) of one batch shape.
struct Acts {
    int N = 0, H = 0, W = 0;
    size_t px() const { return (size_t)N * H * W; }
    rtu::bf16 *F = nullptr, *A1 = nullptr, *CAT = nullptr, *P = nullptr, *B1 = nullptr, *B2 = nullptr, *D1 = nullptr,
              *D2 = nullptr;
    float *noisy = nullptr, *target = nullptr, *pred = nullptr;
    // gradients
    rtu::bf16 *gD2 = nullptr, *gD1 = nullptr, *gCAT = nullptr, *gB2 = nullptr, *gB1 = nullptr, *gP = nullptr,
              *gA2 = nullptr, *gA1 = nullptr;
    bool grads = false;
    void release() {
        void *all[] = {F, A1, CAT, P, B1, B2, D1, D2, noisy, target, pred, gD2, gD1, gCAT, gB2, gB1, gP, gA2, gA1};
        for (void *p : all)
            if (p) (void)hipFree(p);
        *this = Acts{};
    }
    void ensure(int n, int h, int w, bool with_grads) {
        if (n == N && h == H && w == W && (grads || !with_grads)) return;
        release();
        N = n;
        H = h;
        W = w;
        grads = with_grads;
        const size_t P1 = px(), P2 = P1 / 4;
        F = dalloc<rtu::bf16>(P1 * 16, "unet F");
        A1 = dalloc<rtu::bf16>(P1 * 32, "unet A1");
        CAT = dalloc<rtu::bf16>(P1 * 96, "unet CAT");
        P = dalloc<rtu::bf16>(P2 * 32, "unet P");
        B1 = dalloc<rtu::bf16>(P2 * 64, "unet B1");
        B2 = dalloc<rtu::bf16>(P2 * 64, "unet B2");
        D1 = dalloc<rtu::bf16>(P1 * 32, "unet D1");
        D2 = dalloc<rtu::bf16>(P1 * 32, "unet D2");
        noisy = dalloc<float>(P1 * 3, "unet noisy");
        target = dalloc<float>(P1 * 3, "unet target");
        pred = dalloc<float>(P1 * 3, "unet pred");
        if (with_grads) {
            gD2 = dalloc<rtu::bf16>(P1 * 32, "unet gD2");
            gD1 = dalloc<rtu::bf16>(P1 * 32, "unet gD1");
            gCAT = dalloc<rtu::bf16>(P1 * 96, "unet gCAT");
            gB2 = dalloc<rtu::bf16>(P2 * 64, "unet gB2");
            gB1 = dalloc<rtu::bf16>(P2 * 64, "unet gB1");
            gP = dalloc<rtu::bf16>(P2 * 32, "unet gP");
            gA2 = dalloc<rtu::bf16>(P1 * 32, "unet gA2");
            gA1 = dalloc<rtu::bf16>(P1 * 32, "unet gA1");
        }
    }
};

}  // namespace

struct rt_unet {
    int device = 0;
    hipStream_t stream = nullptr;
    // FP32 master parameters: conv weights [9][cout][cin] and biases per layer,
    // then the output layer's w[3][32] and b[3]; one contiguous array.
    size_t n_params = 0;
    size_t w_off[6] = {}, b_off[6] = {}, ow_off = 0, ob_off = 0;
    float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
    rtu::bf16 *d_wf[6] = {}, *d_wt[6] = {};
    uint64_t adam_t = 0;
    Acts inf, train;
    // training set
    uint32_t ds_w = 0, ds_h = 0, ds_cap = 0;
    rtu::bf16 *ds_F = nullptr;
    float *ds_noisy = nullptr, *ds_target = nullptr;
    std::vector<uint8_t> ds_has_in, ds_has_tgt;
    int *d_sel = nullptr;
    size_t sel_cap = 0;
    double *d_loss = nullptr;
    size_t loss_cap = 0;

    ~rt_unet() {
        if (stream) (void)hipStreamSynchronize(stream);
        inf.release();
        train.release();
        void *all[] = {d_p, d_g, d_m, d_v, ds_F, ds_noisy, ds_target, d_sel, d_loss};
        for (void *p : all)
            if (p) (void)hipFree(p);
        for (int k = 0; k < 6; k++) {
            if (d_wf[k]) (void)hipFree(d_wf[k]);
            if (d_wt[k]) (void)hipFree(d_wt[k]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    void refresh_bf16() {
        for (int k = 0; k < 6; k++) {
            const ConvLayer &L = kLayers[k];
            hipLaunchKernelGGL(rtu::weights_bf16_kernel, dim3(blocks_for(L.nw())), dim3(256), 0, stream,
                               d_p + w_off[k], L.nw(), d_wf[k], d_wt[k], L.cout, L.cin);
        }
        hip_check(hipGetLastError(), "weights to bf16");
    }

    // src: rtu::SRC_POOL (X at 2H x 2W, pooled while staged) or rtu::SRC_CONCAT
    // (channels from c1 on read from X2, half resolution) for the inference pass.
    void conv(const rtu::bf16 *X, int xs, int k, bool bwd, rtu::bf16 *Y, int ys, int yoff, bool relu,
              const rtu::bf16 *mask, int ms, int N, int H, int W, int src = rtu::SRC_PLAIN,
              const rtu::bf16 *X2 = nullptr, int x2s = 0, int c1 = 0) {
        const ConvLayer &L = kLayers[k];
        const int cin = bwd ? L.cout : L.cin, cout = bwd ? L.cin : L.cout;
        const int tiles = ((W + 31) / 32) * ((H + rtu::kConvRows - 1) / rtu::kConvRows) * N;
        // the input tile, then the weights [9][cout][cin + 8] (bf16)
        const size_t lds = (size_t)(rtu::kConvRows + 2) * 34 * (cin + 8) * 2 + (size_t)9 * cout * (cin + 8) * 2;
        const rtu::bf16 *w = bwd ? d_wt[k] : d_wf[k];
        const float *b = bwd ? nullptr : d_p + b_off[k];
        // the channel pairs of the six layers, forward (cin, cout) and backward (cout, cin)
#define RT_CONV_SRC(CI, NCO, SRC)                                                                                \
    do {                                                                                                         \
        /* once per instance: raise the dynamic LDS limit (above 64 KB) and size the persistent grid */          \
        static int resident = 0;                                                                                 \
        if (resident == 0) {                                                                                     \
            if (lds > 65536)                                                                                     \
                hip_check(hipFuncSetAttribute((const void *)rtu::conv3x3_kernel<CI, NCO, SRC>,                   \
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),             \
                          "conv LDS size");                                                                      \
            int per_cu = 0;                                                                                      \
            hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rtu::conv3x3_kernel<CI, NCO, SRC>,   \
                                                                   256, lds),                                    \
                      "conv occupancy");                                                                         \
            resident = std::max(per_cu, 1) * cu_count();                                                         \
        }                                                                                                        \
        const dim3 grid((unsigned)std::max(1, std::min(tiles, resident)));                                       \
        hipLaunchKernelGGL((rtu::conv3x3_kernel<CI, NCO, SRC>), grid, dim3(256), lds, stream, X, xs, w, b, Y,   \
                           ys, yoff, relu ? 1 : 0, mask, ms, H, W, X2, x2s, c1, N);                              \
    } while (0)
#define RT_CONV(CI, NCO) RT_CONV_SRC(CI, NCO, rtu::SRC_PLAIN)
        if (src == rtu::SRC_POOL) {
            if (cin == 32 && cout == 64) RT_CONV_SRC(32, 2, rtu::SRC_POOL);
            else throw ApiError(RT_ERR_INVALID, "unsupported pooled convolution");
        } else if (src == rtu::SRC_CONCAT) {
            if (cin == 96 && cout == 32) RT_CONV_SRC(96, 1, rtu::SRC_CONCAT);
            else throw ApiError(RT_ERR_INVALID, "unsupported concatenated convolution");
        } else if (cin == 16 && cout == 32) RT_CONV(16, 1);
        else if (cin == 32 && cout == 32) RT_CONV(32, 1);
        else if (cin == 32 && cout == 64) RT_CONV(32, 2);
        else if (cin == 64 && cout == 64) RT_CONV(64, 2);
        else if (cin == 96 && cout == 32) RT_CONV(96, 1);
        else if (cin == 32 && cout == 96) RT_CONV(32, 3);
        else if (cin == 64 && cout == 32) RT_CONV(64, 1);
        else throw ApiError(RT_ERR_INVALID, "unsupported channel count");
#undef RT_CONV
#undef RT_CONV_SRC
    }
    void wgrad(const rtu::bf16 *G, int k, const rtu::bf16 *X, int xs, int N, int H, int W) {
        const ConvLayer &L = kLayers[k];
        const int segs = N * H * ((W + 31) / 32);
        const dim3 grid((unsigned)std::min(segs, 384), 9, (L.cout / 32) * ((L.cin + 31) / 32));
        hipLaunchKernelGGL(rtu::conv3x3_wgrad_kernel, grid, dim3(64), 0, stream, G, L.cout, L.cout, X, xs, L.cin,
                           d_g + w_off[k], N, H, W);
        hipLaunchKernelGGL(rtu::bias_grad_kernel, dim3(512), dim3(256), 0, stream, G, L.cout, L.cout,
                           (size_t)N * H * W, d_g + b_off[k]);
    }

    // Forward pass over a.F / a.noisy; with `target` the loss goes to loss_slot
    // and the output layer's backward pass fills a.gD2 and the output-layer grads.
    void forward(Acts &a, bool with_target, double *loss_slot) {
        const int N = a.N, H = a.H, W = a.W, H2 = H / 2, W2 = W / 2;
        conv(a.F, 16, 0, false, a.A1, 32, 0, true, nullptr, 0, N, H, W);
        conv(a.A1, 32, 1, false, a.CAT, 96, 0, true, nullptr, 0, N, H, W);
        if (with_target) {  // training: P and CAT[32:96] are kept for the backward pass
            hipLaunchKernelGGL(rtu::pool_kernel, dim3((W2 * 4 + 255) / 256, N * H2), dim3(256), 0, stream, a.CAT,
                               96, 32, a.P, H, W);
            conv(a.P, 32, 2, false, a.B1, 64, 0, true, nullptr, 0, N, H2, W2);
        } else {  // inference: pooled while staged
            conv(a.CAT, 96, 2, false, a.B1, 64, 0, true, nullptr, 0, N, H2, W2, rtu::SRC_POOL);
        }
        conv(a.B1, 64, 3, false, a.B2, 64, 0, true, nullptr, 0, N, H2, W2);
        if (with_target) {
            hipLaunchKernelGGL(rtu::upsample_kernel, dim3((W * 8 + 255) / 256, N * H), dim3(256), 0, stream, a.B2,
                               64, a.CAT, 96, 32, H, W);
            conv(a.CAT, 96, 4, false, a.D1, 32, 0, true, nullptr, 0, N, H, W);
        } else {  // inference: B2 upsampled and concatenated while staged
            conv(a.CAT, 96, 4, false, a.D1, 32, 0, true, nullptr, 0, N, H, W, rtu::SRC_CONCAT, a.B2, 64, 32);
        }
        conv(a.D1, 32, 5, false, a.D2, 32, 0, true, nullptr, 0, N, H, W);
        const float scale = 2.0f / (3.0f * (float)a.px());
        if (with_target)
            hipLaunchKernelGGL(rtu::out_layer_kernel<true>, dim3(blocks_for(a.px(), 1024)), dim3(256), 0, stream, a.D2,
                               a.noisy, d_p + ow_off, d_p + ob_off, a.pred, a.target, a.px(), scale, a.gD2,
                               d_g + ow_off, d_g + ob_off, loss_slot);
        else
            hipLaunchKernelGGL(rtu::out_layer_kernel<false>, dim3(blocks_for(a.px())), dim3(256), 0, stream, a.D2,
                               a.noisy, d_p + ow_off, d_p + ob_off, a.pred, (const float *)nullptr, a.px(), scale,
                               (rtu::bf16 *)nullptr, (float *)nullptr, (float *)nullptr, (double *)nullptr);
        hip_check(hipGetLastError(), "unet forward");
    }

    void backward(Acts &a) {
        const int N = a.N, H = a.H, W = a.W, H2 = H / 2, W2 = W / 2;
        wgrad(a.gD2, 5, a.D1, 32, N, H, W);
        conv(a.gD2, 32, 5, true, a.gD1, 32, 0, false, a.D1, 32, N, H, W);
        wgrad(a.gD1, 4, a.CAT, 96, N, H, W);
        conv(a.gD1, 32, 4, true, a.gCAT, 96, 0, false, a.CAT, 96, N, H, W);
        hipLaunchKernelGGL(rtu::upsample_back_kernel, dim3((W2 * 8 + 255) / 256, N * H2), dim3(256), 0, stream,
                           a.gCAT, 96, 32, 64, a.gB2, H, W);
        wgrad(a.gB2, 3, a.B1, 64, N, H2, W2);
        conv(a.gB2, 64, 3, true, a.gB1, 64, 0, false, a.B1, 64, N, H2, W2);
        wgrad(a.gB1, 2, a.P, 32, N, H2, W2);
        conv(a.gB1, 64, 2, true, a.gP, 32, 0, false, nullptr, 0, N, H2, W2);
        hipLaunchKernelGGL(rtu::pool_back_kernel, dim3((W * 4 + 255) / 256, N * H), dim3(256), 0, stream, a.gCAT, 96,
                           a.gP, a.CAT, 96, 32, a.gA2, H, W);
        wgrad(a.gA2, 1, a.A1, 32, N, H, W);
        conv(a.gA2, 32, 1, true, a.gA1, 32, 0, false, a.A1, 32, N, H, W);
        wgrad(a.gA1, 0, a.F, 16, N, H, W);
        hip_check(hipGetLastError(), "unet backward");
    }

    void adam(float lr) {
        adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, (float)adam_t), c2 = 1.0f - std::pow(b2, (float)adam_t);
        for (int k = 0; k < 6; k++) {
            const ConvLayer &L = kLayers[k];
            hipLaunchKernelGGL(rtu::adam_kernel, dim3(blocks_for(L.nw())), dim3(256), 0, stream, d_p + w_off[k],
                               d_g + w_off[k], d_m + w_off[k], d_v + w_off[k], L.nw(), lr, b1, b2, c1, c2, 1e-8f,
                               d_wf[k], d_wt[k], L.cout, L.cin);
            hipLaunchKernelGGL(rtu::adam_kernel, dim3(1), dim3(256), 0, stream, d_p + b_off[k], d_g + b_off[k],
                               d_m + b_off[k], d_v + b_off[k], (size_t)L.cout, lr, b1, b2, c1, c2, 1e-8f,
                               (rtu::bf16 *)nullptr, (rtu::bf16 *)nullptr, 0, 0);
        }
        const size_t no = (size_t)3 * rtu::kOutC + 3;
        hipLaunchKernelGGL(rtu::adam_kernel, dim3(1), dim3(256), 0, stream, d_p + ow_off, d_g + ow_off, d_m + ow_off,
                           d_v + ow_off, no, lr, b1, b2, c1, c2, 1e-8f, (rtu::bf16 *)nullptr, (rtu::bf16 *)nullptr, 0,
                           0);
        hip_check(hipGetLastError(), "adam");
    }

    void features(Acts &a, const rtu::FrameViews &v) {
        const size_t n = (size_t)v.width * v.height;
        const float ff = std::log2((float)std::max(v.frames, 1u) + 1.0f) / 8.0f;
        hipLaunchKernelGGL(rtu::features_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                           reinterpret_cast<const float4 *>(v.accum), reinterpret_cast<const float4 *>(v.gbuf),
                           reinterpret_cast<const float4 *>(v.albedo), ff, n, a.F, a.noisy);
    }
};

namespace {

rt_unet *create_unet(int device, uint64_t seed) {
    hip_check(hipSetDevice(device), "hipSetDevice");
    auto u = std::make_unique<rt_unet>();
    u->device = device;
    hip_check(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking), "hipStreamCreate");
    size_t off = 0;
    for (int k = 0; k < 6; k++) {
        u->w_off[k] = off;
        off += kLayers[k].nw();
        u->b_off[k] = off;
        off += kLayers[k].cout;
    }
    u->ow_off = off;
    off += (size_t)3 * rtu::kOutC;
    u->ob_off = off;
    off += 3;
    u->n_params = off;
    std::vector<float> p(off, 0.0f);
    std::mt19937_64 rng(seed);
    for (int k = 0; k < 6; k++) {
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (9.0f * kLayers[k].cin)));
        for (size_t i = 0; i < kLayers[k].nw(); i++) p[u->w_off[k] + i] = nd(rng);
    }
    // the output layer starts at zero: the untrained network returns its input
    u->d_p = dalloc<float>(off, "unet params");
    u->d_g = dalloc<float>(off, "unet grads");
    u->d_m = dalloc<float>(off, "unet adam m");
    u->d_v = dalloc<float>(off, "unet adam v");
    hip_check(hipMemcpy(u->d_p, p.data(), off * 4, hipMemcpyHostToDevice), "upload params");
    for (float *b : {u->d_g, u->d_m, u->d_v}) hip_check(hipMemset(b, 0, off * 4), "memset");
    for (int k = 0; k < 6; k++) {
        u->d_wf[k] = dalloc<rtu::bf16>(kLayers[k].nw(), "unet bf16 weights");
        u->d_wt[k] = dalloc<rtu::bf16>(kLayers[k].nw(), "unet bf16 weights (backward)");
    }
    hip_check(hipDeviceSynchronize(), "sync");
    u->refresh_bf16();
    hip_check(hipStreamSynchronize(u->stream), "sync");
    return u.release();
}

void check_views(const rtu::FrameViews &v) {
    if (!v.accum || !v.gbuf || !v.albedo) throw ApiError(RT_ERR_INVALID, "frame views incomplete");
    if (v.width % 2 || v.height % 2) throw ApiError(RT_ERR_INVALID, "the denoiser needs even width and height");
}

}  // namespace

namespace rtu {

void denoise(rt_unet *u, const FrameViews &v, void *out, hipStream_t stream) {
    check_views(v);
    // the network runs on the caller's stream
    hipStream_t saved = u->stream;
    u->stream = stream;
    try {
        u->inf.ensure(1, (int)v.height, (int)v.width, false);
        u->features(u->inf, v);
        u->forward(u->inf, false, nullptr);
        const size_t n = (size_t)v.width * v.height;
        hipLaunchKernelGGL(image_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, u->inf.pred, n,
                           reinterpret_cast<float4 *>(out));
        hip_check(hipGetLastError(), "unet image");
    } catch (...) {
        u->stream = saved;
        throw;
    }
    u->stream = saved;
}

}  // namespace rtu

extern "C" {

int rt_unet_create(int device, uint64_t seed, rt_unet **out) {
    return rt::api_call([&] {
        if (!out) throw ApiError(RT_ERR_INVALID, "null argument");
        *out = create_unet(device, seed);
    });
}

int rt_unet_destroy(rt_unet *u) {
    return rt::api_call([&] { delete u; });
}

int rt_unet_param_count(const rt_unet *u, size_t *n) {
    return rt::api_call([&] {
        if (!u || !n) throw ApiError(RT_ERR_INVALID, "null argument");
        *n = u->n_params;
    });
}

int rt_unet_get_params(rt_unet *u, float *out, size_t n) {
    return rt::api_call([&] {
        if (!u || !out) throw ApiError(RT_ERR_INVALID, "null argument");
        if (n < u->n_params) throw ApiError(RT_ERR_INVALID, "buffer smaller than the parameter count");
        hip_check(hipStreamSynchronize(u->stream), "sync");
        hip_check(hipMemcpy(out, u->d_p, u->n_params * 4, hipMemcpyDeviceToHost), "read params");
    });
}

int rt_unet_set_params(rt_unet *u, const float *in, size_t n) {
    return rt::api_call([&] {
        if (!u || !in) throw ApiError(RT_ERR_INVALID, "null argument");
        if (n != u->n_params) throw ApiError(RT_ERR_INVALID, "parameter count mismatch");
        hip_check(hipStreamSynchronize(u->stream), "sync");
        hip_check(hipMemcpy(u->d_p, in, n * 4, hipMemcpyHostToDevice), "write params");
        u->refresh_bf16();
        hip_check(hipStreamSynchronize(u->stream), "sync");
    });
}

int rt_unet_dataset_reset(rt_unet *u, uint32_t width, uint32_t height, uint32_t capacity) {
    return rt::api_call([&] {
        if (!u) throw ApiError(RT_ERR_INVALID, "null argument");
        if (width % 2 || height % 2 || width < 32 || height < 32 || capacity == 0)
            throw ApiError(RT_ERR_INVALID, "dataset images must be even-sized, >= 32, capacity > 0");
        hip_check(hipStreamSynchronize(u->stream), "sync");
        for (void *p : {(void *)u->ds_F, (void *)u->ds_noisy, (void *)u->ds_target})
            if (p) (void)hipFree(p);
        const size_t px = (size_t)width * height * capacity;
        u->ds_F = dalloc<rtu::bf16>(px * 16, "dataset features");
        u->ds_noisy = dalloc<float>(px * 3, "dataset inputs");
        u->ds_target = dalloc<float>(px * 3, "dataset targets");
        u->ds_w = width;
        u->ds_h = height;
        u->ds_cap = capacity;
        u->ds_has_in.assign(capacity, 0);
        u->ds_has_tgt.assign(capacity, 0);
    });
}

int rt_unet_dataset_set(rt_unet *u, uint32_t slot, int which, rt_renderer *r) {
    return rt::api_call([&] {
        if (!u || !r) throw ApiError(RT_ERR_INVALID, "null argument");
        if (slot >= u->ds_cap) throw ApiError(RT_ERR_INVALID, "slot out of range");
        hip_check(hipStreamSynchronize(u->stream), "sync");
        const rtu::FrameViews v = rtu::frame_views(r);
        if (v.width != u->ds_w || v.height != u->ds_h) throw ApiError(RT_ERR_INVALID, "image size differs from the dataset");
        const size_t n = (size_t)v.width * v.height, base = (size_t)slot * n;
        if (which == 0) {
            check_views(v);
            const float ff = std::log2((float)std::max(v.frames, 1u) + 1.0f) / 8.0f;
            hipLaunchKernelGGL(rtu::features_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, u->stream,
                               reinterpret_cast<const float4 *>(v.accum), reinterpret_cast<const float4 *>(v.gbuf),
                               reinterpret_cast<const float4 *>(v.albedo), ff, n, u->ds_F + base * 16,
                               u->ds_noisy + base * 3);
            u->ds_has_in[slot] = 1;
        } else {
            hipLaunchKernelGGL(rtu::log_target_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, u->stream,
                               reinterpret_cast<const float4 *>(v.accum), n, u->ds_target + base * 3);
            u->ds_has_tgt[slot] = 1;
        }
        hip_check(hipGetLastError(), "dataset set");
        hip_check(hipStreamSynchronize(u->stream), "sync");
    });
}

int rt_unet_train(rt_unet *u, uint32_t steps, uint32_t batch, uint32_t crop, float lr, uint64_t seed,
                  float *losses) {
    return rt::api_call([&] {
        if (!u) throw ApiError(RT_ERR_INVALID, "null argument");
        std::vector<uint32_t> ready;
        for (uint32_t k = 0; k < u->ds_cap; k++)
            if (u->ds_has_in[k] && u->ds_has_tgt[k]) ready.push_back(k);
        if (ready.empty()) throw ApiError(RT_ERR_INVALID, "no dataset slot holds both an input and a target");
        if (crop % 32 || crop < 32 || crop > u->ds_w || crop > u->ds_h || batch == 0 || batch > 256)
            throw ApiError(RT_ERR_INVALID, "crop must be a multiple of 32 within the images; 1 <= batch <= 256");
        u->train.ensure((int)batch, (int)crop, (int)crop, true);
        const size_t sel_n = (size_t)3 * batch * std::max(steps, 1u);
        if (u->sel_cap < sel_n) {
            if (u->d_sel) (void)hipFree(u->d_sel);
            u->d_sel = dalloc<int>(sel_n, "crop selection");
            u->sel_cap = sel_n;
        }
        if (u->loss_cap < steps) {
            if (u->d_loss) (void)hipFree(u->d_loss);
            u->d_loss = dalloc<double>(steps, "losses");
            u->loss_cap = steps;
        }
        hip_check(hipMemsetAsync(u->d_loss, 0, 8ull * std::max(steps, 1u), u->stream), "clear losses");
        // every step's crops drawn up front and uploaded once
        std::mt19937_64 rng(seed);
        std::vector<int> sel((size_t)3 * batch * std::max(steps, 1u));
        for (size_t k = 0; k < (size_t)batch * steps; k++) {
            sel[3 * k] = (int)ready[rng() % ready.size()];
            sel[3 * k + 1] = (int)(rng() % ((u->ds_w - crop) / 2 + 1)) * 2;
            sel[3 * k + 2] = (int)(rng() % ((u->ds_h - crop) / 2 + 1)) * 2;
        }
        hip_check(hipStreamSynchronize(u->stream), "sync");
        hip_check(hipMemcpy(u->d_sel, sel.data(), sel.size() * 4, hipMemcpyHostToDevice), "upload crops");
        for (uint32_t s = 0; s < steps; s++) {
            hipLaunchKernelGGL(rtu::crop_kernel, dim3(std::max(1u, crop * crop / 256), batch), dim3(256), 0,
                               u->stream, u->ds_F, u->ds_noisy, u->ds_target, u->d_sel + (size_t)3 * batch * s,
                               (int)u->ds_w, (int)u->ds_h, (int)crop, (int)crop, u->train.F, u->train.noisy,
                               u->train.target);
            hipLaunchKernelGGL(rtu::zero_kernel, dim3(blocks_for(u->n_params)), dim3(256), 0, u->stream, u->d_g,
                               u->n_params);
            u->forward(u->train, true, u->d_loss + s);
            u->backward(u->train);
            u->adam(lr);
        }
        hip_check(hipStreamSynchronize(u->stream), "train sync");
        if (losses && steps) {
            std::vector<double> l(steps);
            hip_check(hipMemcpy(l.data(), u->d_loss, 8ull * steps, hipMemcpyDeviceToHost), "read losses");
            const double norm = 1.0 / (3.0 * (double)u->train.px());
            for (uint32_t s = 0; s < steps; s++) losses[s] = (float)(l[s] * norm);
        }
    });
}

// Test hooks: one forward (and backward) pass on caller-supplied host tensors.
int rt_unet_eval(rt_unet *u, uint32_t n, uint32_t h, uint32_t w, const uint16_t *features, const float *noisy,
                 const float *target, float *pred, float *grads, double *loss) {
    return rt::api_call([&] {
        if (!u || !features || !noisy || !pred) throw ApiError(RT_ERR_INVALID, "null argument");
        if (h % 2 || w % 2 || n == 0) throw ApiError(RT_ERR_INVALID, "even sizes required");
        Acts &a = u->train;
        a.ensure((int)n, (int)h, (int)w, true);
        const size_t px = (size_t)n * h * w;
        hip_check(hipStreamSynchronize(u->stream), "sync");
        hip_check(hipMemcpy(a.F, features, px * 16 * 2, hipMemcpyHostToDevice), "upload features");
        hip_check(hipMemcpy(a.noisy, noisy, px * 12, hipMemcpyHostToDevice), "upload inputs");
        if (target) hip_check(hipMemcpy(a.target, target, px * 12, hipMemcpyHostToDevice), "upload targets");
        if (u->loss_cap < 1) {
            u->d_loss = dalloc<double>(1, "loss");
            u->loss_cap = 1;
        }
        hip_check(hipMemsetAsync(u->d_loss, 0, 8, u->stream), "clear loss");
        hipLaunchKernelGGL(rtu::zero_kernel, dim3(blocks_for(u->n_params)), dim3(256), 0, u->stream, u->d_g,
                           u->n_params);
        u->forward(a, target != nullptr, u->d_loss);
        if (target) u->backward(a);
        hip_check(hipStreamSynchronize(u->stream), "sync");
        hip_check(hipMemcpy(pred, a.pred, px * 12, hipMemcpyDeviceToHost), "read pred");
        if (target && grads) hip_check(hipMemcpy(grads, u->d_g, u->n_params * 4, hipMemcpyDeviceToHost), "read grads");
        if (target && loss) {
            double l = 0;
            hip_check(hipMemcpy(&l, u->d_loss, 8, hipMemcpyDeviceToHost), "read loss");
            *loss = l / (3.0 * (double)px);
        }
    });
}

}  // extern "C"
