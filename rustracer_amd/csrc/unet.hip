// unet.hip — the learned denoiser: a convolutional U-Net on the displayed
// (accumulated) image, with hand-written BF16 MFMA convolution kernels for the
// forward pass and both backward passes, FP32 accumulation, and Adam, for gfx950.
//
// This part has no counterpart in the reference renderer; its inputs are the
// renderer's own buffers (accumulation, accumulated luminance moments,
// G-buffer) plus a first-hit albedo pass.
//
// Network (NHWC bf16 activations; image sizes padded to multiples of 16):
//   F: 12 features (log1p rgb, albedo, normal, depth, 2 variances) + 4 zeros
//   encoder level l = 0..4 (channels 32, 64, 128, 128, 128):
//     T_l = relu(conv3x3(X_l)), S_l = relu(conv3x3(T_l)); X_0 = F,
//     X_l = avgpool2(S_{l-1}); S_4 = B is the bottleneck
//   decoder level l = 3..0 (channels 128, 128, 64, 32):
//     CAT_l = [S_l, upsample2(V_{l+1})] (V_4 = B; 256, 256, 192, 96 channels)
//     U_l = relu(conv3x3(CAT_l)), V_l = relu(conv3x3(U_l))
//   pred = log1p(noisy) + conv1x1(V_0, 32->3) + b (FP32), image = expm1(pred)
// Loss: mean over pixels and rgb of |pred - log1p(target)| (L1 on log radiance).
// Every gradient is a fixed-order sum (per-workgroup partials, then a
// reduction in index order), so a training run is reproducible bit for bit.
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
#include <array>
#include <map>
#include <cmath>
#include <cstdio>
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
// offset 0; Wt: [9][cout_total][CIN]; Y: [N][H][W][ys], cout_total channels at
// offset yoff.
//
// Work items are (tile, output-channel block): a tile is 32 pixels of
// kConvRows = 4 * RPW consecutive rows, a block CB = 32 * NCO output channels.
// A workgroup of four wavefronts computes one item at a time, each wavefront
// RPW rows (2 or 4) for all CB channels; the weights are the MFMA's A operand
// (row = output channel) and the pixels its B operand (column = pixel), so
// every weight fragment feeds RPW MFMAs and each lane ends with 4 consecutive
// output channels of one pixel per accumulator group (8-byte epilogue
// stores). Per 16-channel group a wavefront reads its RPW + 2 input rows at
// the three column shifts once (3 * (RPW + 2) fragments serve the 9 * RPW
// products).
//
// The reduction over 9 taps x CIN runs in CIN / KC slices of KC input
// channels, always in the same order (16-channel groups in increasing order,
// the 9 taps inside each), so every variant of a layer gives the same bits. Per slice, the input rows y0-1 .. y0+kConvRows, pixels x0-1 ..
// x0+32 of those channels are staged in LDS (pixel stride KC + 8 to spread
// banks) and serve all 9 taps, and so are the slice's weights [9][CB][KC + 8].
// With WONCE (one slice and one block: the whole layer's weights fit) the
// weights are staged once per workgroup and only the input tiles move. The
// loads of the next (item, slice) step are issued into registers before the
// current step's MFMAs and written to LDS after them, so their latency
// overlaps the compute. The workgroups are persistent (as many as are
// resident at once) and take items blockIdx.x, blockIdx.x + gridDim.x, ...;
// the output-channel block varies fastest, so concurrent workgroups share
// input tiles in L2.
//
// The inference pass folds the pool and the upsample into the staging of the
// layers that read them (the training pass keeps the pooled and concatenated
// tensors for its backward): SRC_POOL reads X at twice the output resolution
// and stages the 2x2 average of pool_kernel; SRC_CONCAT stages channels [0, c1)
// from X and [c1, CIN) from X2 (half resolution, stride x2s) at (y/2, x/2), as
// upsample_kernel writes them into the concatenation. The staged values are
// bit-identical to those kernels' outputs.
enum ConvSrc { SRC_PLAIN = 0, SRC_POOL = 1, SRC_CONCAT = 2 };

// rpw rows per wavefront, 4 * rpw rows per tile
constexpr size_t conv_lds_bytes(int kc, int nco, int rpw) {
    return ((size_t)(4 * rpw + 2) * 34 * (kc + 8) + (size_t)9 * 32 * nco * (kc + 8)) * 2;
}

template <int CIN, int KC, int NCO, int SRC, bool WONCE, int RPW>
__global__ __launch_bounds__(256) void conv3x3_kernel(const bf16 *__restrict__ X, int xs,
                                                      const bf16 *__restrict__ Wt, int cout_total,
                                                      const float *__restrict__ bias, bf16 *__restrict__ Y, int ys,
                                                      int yoff, int relu, const bf16 *__restrict__ mask, int ms,
                                                      int H, int W, const bf16 *__restrict__ X2, int x2s, int c1,
                                                      int N) {
    static_assert(CIN % KC == 0 && KC % 16 == 0, "input channels in slices of a multiple of 16");
    static_assert(!WONCE || CIN == KC, "weights staged once: one slice");
    constexpr int R = RPW;
    constexpr int kConvRows = 4 * RPW;
    constexpr int NS = CIN / KC;
    constexpr int PS = KC + 8;  // LDS pixel and weight-row stride (elements)
    constexpr int GROUPS = KC / 8;
    constexpr int CB = 32 * NCO;
    constexpr int TITEMS = (kConvRows + 2) * 34 * GROUPS;
    constexpr int TPER = (TITEMS + 255) / 256;
    constexpr int WITEMS = 9 * CB * GROUPS;
    constexpr int WPER = WONCE ? 1 : (WITEMS + 255) / 256;
    constexpr int NL = SRC == SRC_POOL ? 4 : 1;
    // dynamic LDS: the input slice [kConvRows + 2][34][PS], then the weight
    // slice [9][CB][PS] (rows tap * CB + output channel)
    extern __shared__ __attribute__((aligned(16))) bf16 tile[];
    bf16 *wl = tile + (kConvRows + 2) * 34 * PS;
    const int l = threadIdx.x % 64, wv = threadIdx.x / 64;
    const int tiles_x = (W + 31) / 32, tiles_y = (H + kConvRows - 1) / kConvRows;
    const int ncb = cout_total / CB;
    const int items = tiles_x * tiles_y * N * ncb;
    // weight item i: row (tap * CB + j), 8-channel group g of the slice
    auto w_src = [&](int i, int cb, int s) {
        const int row = i / GROUPS, g = i % GROUPS;
        const int tap = row / CB, j = row % CB;
        return Wt + ((size_t)tap * cout_total + cb * CB + j) * CIN + s * KC + g * 8;
    };
    if (WONCE) {
        constexpr int WB = 8;
        constexpr int WALL = (WITEMS + 255) / 256;
#pragma unroll
        for (int k0 = 0; k0 < WALL; k0 += WB) {
            uint4 v[WB];
#pragma unroll
            for (int k = 0; k < WB; k++) {
                const int i = threadIdx.x + (k0 + k) * 256;
                if (k0 + k < WALL && i < WITEMS) v[k] = ld16(w_src(i, 0, 0));
            }
#pragma unroll
            for (int k = 0; k < WB; k++) {
                const int i = threadIdx.x + (k0 + k) * 256;
                if (k0 + k < WALL && i < WITEMS)
                    *reinterpret_cast<uint4 *>(wl + (i / GROUPS) * PS + (i % GROUPS) * 8) = v[k];
            }
        }
    }
    auto item_of = [&](int q) { return (int)blockIdx.x + (q / NS) * (int)gridDim.x; };
    // Per-thread staging items, worked out once: item k is (row, pixel, group)
    // packed as row << 16 | px << 8 | g (row 0x7FFF past the last item), its
    // element offset from the tile's origin in X (it_o1) and, when the
    // concatenation's second part comes from X2, in X2 (it_o2, less c1). A
    // step then needs a few adds per item: the tile's origin, the bounds test
    // and the address, with no branches (an item outside the image loads the
    // tensor's first element and keeps zero instead).
    int it_rp[TPER], it_o1[TPER], it_o2[SRC == SRC_CONCAT ? TPER : 1];
#pragma unroll
    for (int k = 0; k < TPER; k++) {
        const int i = threadIdx.x + k * 256;
        const int g = i % GROUPS, qq = i / GROUPS;
        const int px = qq % 34, row = qq / 34;
        it_rp[k] = i < TITEMS ? (row << 16) | (px << 8) | g : 0x7FFF0000;
        it_o1[k] = SRC == SRC_POOL ? ((2 * row) * (2 * W) + 2 * px) * xs + g * 8 : (row * W + px) * xs + g * 8;
        if (SRC == SRC_CONCAT) it_o2[k] = (((row - 1) >> 1) * (W / 2) + ((px - 1) >> 1)) * x2s + g * 8 - c1;
    }
    int it_w[WONCE ? 1 : WPER];
    if (!WONCE) {
#pragma unroll
        for (int k = 0; k < WPER; k++) {
            const int i = threadIdx.x + k * 256;
            const int row = i / GROUPS, g = i % GROUPS;
            it_w[k] = ((row / CB) * cout_total + row % CB) * CIN + g * 8;
        }
    }
    uint4 rt[TPER][NL];
    uint4 rw[WPER];
    uint32_t okm = 0;  // bit k: item k of the fetched step lies inside the image
    auto fetch = [&](int q) {
        const int item = item_of(q), s = q % NS;
        const int cb = item % ncb, t = item / ncb;
        const int bx = t % tiles_x, rest = t / tiles_x;
        const int x0 = bx * 32, y0 = (rest % tiles_y) * kConvRows, n = rest / tiles_y;
        // element offsets of the tile's origin (row y0 - 1, column x0 - 1; may be
        // negative, but never for an item inside the image)
        const long long b1 = SRC == SRC_POOL
                                 ? ((long long)(n * 2 * H + 2 * (y0 - 1)) * (2 * W) + 2 * (x0 - 1)) * xs + s * KC
                                 : ((long long)(n * H + y0 - 1) * W + (x0 - 1)) * xs + s * KC;
        const long long b2 = SRC == SRC_CONCAT ? ((long long)(n * (H / 2) + y0 / 2) * (W / 2) + x0 / 2) * x2s + s * KC : 0;
        okm = 0;
#pragma unroll
        for (int k = 0; k < TPER; k++) {
            const int rp = it_rp[k];
            const int yy = y0 - 1 + (rp >> 16), xx = x0 - 1 + ((rp >> 8) & 0xFF);
            const bool ok = (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W;
            okm |= ok ? 1u << k : 0u;
            const bf16 *p;
            if (SRC == SRC_CONCAT) {
                const bool second = s * KC + (rp & 0xFF) * 8 >= c1;
                p = second ? X2 + (ok ? b2 + it_o2[SRC == SRC_CONCAT ? k : 0] : 0) : X + (ok ? b1 + it_o1[k] : 0);
            } else {
                p = X + (ok ? b1 + it_o1[k] : 0);
            }
            if (SRC == SRC_POOL) {
                const size_t down = (size_t)(2 * W) * xs;
                rt[k][0] = ld16(p);
                rt[k][NL > 1 ? 1 : 0] = ld16(p + xs);
                rt[k][NL > 2 ? 2 : 0] = ld16(p + down);
                rt[k][NL > 3 ? 3 : 0] = ld16(p + down + xs);
            } else {
                rt[k][0] = ld16(p);
            }
        }
        if (!WONCE) {
            const bf16 *wb = Wt + (size_t)cb * CB * CIN + s * KC;
#pragma unroll
            for (int k = 0; k < WPER; k++) {
                const int i = threadIdx.x + k * 256;
                if (i < WITEMS) rw[k] = ld16(wb + it_w[k]);
            }
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int k = 0; k < TPER; k++) {
            const int i = threadIdx.x + k * 256;
            if (i < TITEMS) {
                const int rp = it_rp[k];
                uint4 v = rt[k][0];
                // an item outside the image stages zero (for the pool too: the
                // pool of a padded position is 0.25 * (0 + 0 + 0 + 0) = 0)
                if (SRC == SRC_POOL && (okm >> k & 1u)) {
                    float a[8], b[8], d[8], e[8], o[8];
                    unpack8(rt[k][0], a);
                    unpack8(rt[k][NL > 1 ? 1 : 0], b);
                    unpack8(rt[k][NL > 2 ? 2 : 0], d);
                    unpack8(rt[k][NL > 3 ? 3 : 0], e);
#pragma unroll
                    for (int c8 = 0; c8 < 8; c8++) o[c8] = 0.25f * (((a[c8] + b[c8]) + d[c8]) + e[c8]);
                    v = pack8(o);
                }
                if (!(okm >> k & 1u)) v = make_uint4(0u, 0u, 0u, 0u);
                *reinterpret_cast<uint4 *>(tile + ((rp >> 16) * 34 + ((rp >> 8) & 0xFF)) * PS + (rp & 0xFF) * 8) = v;
            }
        }
        if (!WONCE) {
#pragma unroll
            for (int k = 0; k < WPER; k++) {
                const int i = threadIdx.x + k * 256;
                if (i < WITEMS) *reinterpret_cast<uint4 *>(wl + (i / GROUPS) * PS + (i % GROUPS) * 8) = rw[k];
            }
        }
    };
    floatx16 acc[R][NCO];
    if (item_of(0) < items) fetch(0);
    for (int q = 0; item_of(q) < items; q++) {
        __syncthreads();  // every wavefront is done with the previous step (and the weights are in)
        commit();
        __syncthreads();
        if (item_of(q + 1) < items) fetch(q + 1);
        const int item = item_of(q), s = q % NS;
        const int cb = item % ncb, t = item / ncb;
        const int bx = t % tiles_x, rest = t / tiles_x;
        const int x0 = bx * 32, y0 = (rest % tiles_y) * kConvRows, n = rest / tiles_y;
        const int c = l % 32, kg = (l / 32) * 8;
        const bf16 *trow = tile + (wv * R * 34 + c) * PS + kg;
        if (s == 0) {
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int j = 0; j < NCO; j++) acc[r][j] = floatx16{};
        }
        // canonical order: 16-channel groups in increasing order, the 9 taps
        // inside each; every slicing and blocking of a layer therefore sums in
        // the same order and gives the same bits
#pragma unroll
        for (int c0 = 0; c0 < KC; c0 += 16) {
            // the wavefront's R + 2 input rows at the three column shifts, read
            // once for all nine taps (tap (dy, dx) of row r is row r + dy, shift dx)
            bf16x8 b[R + 2][3];
#pragma unroll
            for (int rr = 0; rr < R + 2; rr++)
#pragma unroll
                for (int dx = 0; dx < 3; dx++) b[rr][dx] = load8(trow + (rr * 34 + dx) * PS + c0);
#pragma unroll
            for (int tap = 0; tap < 9; tap++) {
                const bf16 *wp = wl + (tap * CB + c) * PS + kg + c0;
                bf16x8 a[NCO];
#pragma unroll
                for (int j = 0; j < NCO; j++) a[j] = load8(wp + j * 32 * PS);
#pragma unroll
                for (int r = 0; r < R; r++)
#pragma unroll
                    for (int j = 0; j < NCO; j++)
                        acc[r][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[j], b[r + tap / 3][tap % 3],
                                                                            acc[r][j], 0, 0, 0);
            }
        }
        if (s != NS - 1) continue;
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
                        const int co = cb * CB + j * 32 + g * 8 + (l / 32) * 4;
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

// Weight gradient of a 3x3 convolution: gW[tap][co][ci] = sum over pixels of
// G[p][co] * X[p + tap][ci], as an MFMA GEMM with M = 32 output channels, N =
// 32 input channels and K = pixels. A workgroup of three wavefronts takes one
// (32 co, 32 ci) block (blockIdx.x, which varies fastest, so the workgroups of
// one tile range run together and share its loads in L2) and all nine taps,
// over the tiles blockIdx.y, blockIdx.y + gridDim.y, ... of 32 pixels x
// kWgRows rows. Per
// tile, G and the (kWgRows + 2) x 34 input window are staged transposed in LDS
// ([channel][pixel]); the window is staged once per column shift dx, so every
// MFMA fragment is an aligned 16-byte read. A tile's global loads are all
// issued before any of them is written to LDS, and the next tile's loads are
// issued before this tile's MFMAs. Wavefront w accumulates the taps
// of row dy = w (3w .. 3w + 2), each G fragment feeding its three MFMAs. The
// workgroup's sum goes to its own partial array part[blockIdx.y][tap][co][ci]
// (no atomics): reduce_partials_kernel adds the partials in a fixed order, so
// the gradient does not depend on scheduling. The workgroups of input-channel
// block 0 also write the bias gradient's partials bpart[blockIdx.y][wavefront][co]:
// G summed over their pixels by one more MFMA per K step, against ones.
constexpr int kWgRows = 4;                      // rows per tile
constexpr int kWgK = 32 * kWgRows;              // pixels per tile
constexpr int kWgGS = kWgK + 8;                 // LDS stride of one G channel (elements)
constexpr int kWgXS = (kWgRows + 2) * 32 + 8;   // LDS stride of one input channel in one dx copy

// Four bf16 of one channel (16-bit word j of the four 8-channel vectors a..d)
// as one 8-byte LDS word.
__device__ __forceinline__ uint2 pick4(const uint4 &a, const uint4 &b, const uint4 &c, const uint4 &d, int j) {
    const uint32_t wa = (&a.x)[j >> 1], wb = (&b.x)[j >> 1], wc = (&c.x)[j >> 1], wd = (&d.x)[j >> 1];
    if (j & 1) return make_uint2((wa >> 16) | (wb & 0xFFFF0000u), (wc >> 16) | (wd & 0xFFFF0000u));
    return make_uint2((wa & 0xFFFFu) | (wb << 16), (wc & 0xFFFFu) | (wd << 16));
}

__global__ __launch_bounds__(192) void conv3x3_wgrad_kernel(const bf16 *__restrict__ G, int gs, int cout,
                                                            const bf16 *__restrict__ X, int xs, int cin,
                                                            float *__restrict__ part, float *__restrict__ bpart,
                                                            int N, int H, int W) {
    __shared__ __attribute__((aligned(16))) bf16 gt[32 * kWgGS];      // [co][row * 32 + x]
    __shared__ __attribute__((aligned(16))) bf16 xt[3][32 * kWgXS];   // [dx][ci][row * 32 + x], row 0 = y0 - 1
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nci = (cin + 31) / 32;
    const int co0 = (blockIdx.x / nci) * 32, ci0 = (blockIdx.x % nci) * 32;
    const int tiles_x = (W + 31) / 32, tiles_y = (H + kWgRows - 1) / kWgRows;
    const int total = N * tiles_y * tiles_x;
    // Staging items (one per thread): 8 channels x 4 consecutive pixels of one
    // row. G: rows 0 .. kWgRows - 1 (threads below 32 * kWgRows); the window:
    // rows 0 .. kWgRows + 1, where a thread loads window columns 4m .. 4m + 5
    // and writes columns 4m + dx .. 4m + dx + 3 to copy dx. A tile's loads are
    // all issued before any LDS write, and the next tile's before the MFMAs.
    static_assert((kWgRows + 2) * 32 == 192, "one window item per thread");
    const int sg = threadIdx.x & 3, sm = (threadIdx.x >> 2) & 7, sr = threadIdx.x >> 5;
    const bool g_item = sr < kWgRows;
    uint4 rg[4], rx[6];
    auto fetch = [&](int t) {
        const int tx = t % tiles_x, rest = t / tiles_x;
        const int x0 = tx * 32, y0 = (rest % tiles_y) * kWgRows, n = rest / tiles_y;
        {
            const int yy = y0 + sr;
#pragma unroll
            for (int p = 0; p < 4; p++) {
                const int xx = x0 + 4 * sm + p;
                rg[p] = make_uint4(0u, 0u, 0u, 0u);
                if (g_item && yy < H && xx < W)
                    rg[p] = ld16(G + ((size_t)(n * H + yy) * W + xx) * gs + co0 + sg * 8);
            }
        }
        {
            const int yy = y0 - 1 + sr, ch = ci0 + sg * 8;
#pragma unroll
            for (int p = 0; p < 6; p++) {
                const int xx = x0 - 1 + 4 * sm + p;
                rx[p] = make_uint4(0u, 0u, 0u, 0u);
                if (yy >= 0 && yy < H && xx >= 0 && xx < W && ch < cin)
                    rx[p] = ld16(X + ((size_t)(n * H + yy) * W + xx) * xs + ch);
            }
        }
    };
    auto commit = [&]() {
        if (g_item) {
#pragma unroll
            for (int j = 0; j < 8; j++)
                *reinterpret_cast<uint2 *>(&gt[(sg * 8 + j) * kWgGS + sr * 32 + 4 * sm]) =
                    pick4(rg[0], rg[1], rg[2], rg[3], j);
        }
#pragma unroll
        for (int dx = 0; dx < 3; dx++)
#pragma unroll
            for (int j = 0; j < 8; j++)
                *reinterpret_cast<uint2 *>(&xt[dx][(sg * 8 + j) * kWgXS + sr * 32 + 4 * sm]) =
                    pick4(rx[dx], rx[dx + 1], rx[dx + 2], rx[dx + 3], j);
    };
    floatx16 acc[3] = {};
    // bias gradient: the workgroups of input-channel block 0 also sum G over
    // their pixels, one more MFMA against ones per K step, taken by wavefront
    // ks % 3 (each wavefront keeps its own partial)
    const bool bias_wg = blockIdx.x % nci == 0;
    const bf16x8 ones = __builtin_bit_cast(bf16x8, make_uint4(0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u));
    floatx16 accb = {};
    if ((int)blockIdx.y < total) fetch(blockIdx.y);
    for (int t = blockIdx.y; t < total; t += gridDim.y) {
        __syncthreads();  // every wavefront is done with the previous tile
        commit();
        __syncthreads();
        if (t + (int)gridDim.y < total) fetch(t + gridDim.y);
        // K steps of 16 pixels in pixel order; a lane's 8 pixels lie in one row
#pragma unroll 2
        for (int ks = 0; ks < kWgK / 16; ks++) {
            const int k = ks * 16 + (l >> 5) * 8;
            const bf16x8 a = load8(&gt[(l & 31) * kWgGS + k]);
            const bf16 *xb = &xt[0][(l & 31) * kWgXS + ((k >> 5) + wv) * 32 + (k & 31)];
#pragma unroll
            for (int dx = 0; dx < 3; dx++)
                acc[dx] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, load8(xb + dx * 32 * kWgXS), acc[dx], 0, 0, 0);
            if (bias_wg && ks % 3 == wv) accb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, ones, accb, 0, 0, 0);
        }
    }
    // every column of accb holds the sum over the pixels for output channels
    // (e / 4) * 8 + (l / 32) * 4 + e % 4; column 0 (lanes 0 and 32) stores it
    if (bias_wg && (l & 31) == 0) {
        float *pb = bpart + ((size_t)blockIdx.y * 3 + wv) * cout + co0;
#pragma unroll
        for (int e = 0; e < 16; e++) pb[(e / 4) * 8 + (l / 32) * 4 + e % 4] = accb[e];
    }
    const int ci = ci0 + (l & 31);
    if (ci >= cin) return;
    float *pg = part + (size_t)blockIdx.y * 9 * cout * cin;
#pragma unroll
    for (int dx = 0; dx < 3; dx++) {
        const int tap = wv * 3 + dx;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int co = co0 + (e / 4) * 8 + (l / 32) * 4 + e % 4;
            pg[((size_t)tap * cout + co) * cin + ci] = acc[dx][e];
        }
    }
}

// out[i] = sum over g = 0 .. n_part-1 (in that order) of part[g * stride + i],
// for i < n.
__global__ __launch_bounds__(256) void reduce_partials_kernel(const float *__restrict__ part, int n_part, size_t stride,
                                                              size_t n, float *__restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        // eight loads in flight, added in index order
        float s = 0.0f;
        int g = 0;
        for (; g + 8 <= n_part; g += 8) {
            float v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = part[(size_t)(g + k) * stride + i];
#pragma unroll
            for (int k = 0; k < 8; k++) s += v[k];
        }
        for (; g < n_part; g++) s += part[(size_t)g * stride + i];
        out[i] = s;
    }
}

// First level of a two-level fixed-order sum: out2[c * n + i] = sum over g in
// [c * chunk, min((c + 1) * chunk, n_part)), in that order, of part[g * stride
// + i]; grid.y = c. reduce_partials_kernel then adds the chunks in order.
__global__ __launch_bounds__(256) void reduce_chunks_kernel(const float *__restrict__ part, int n_part, size_t stride,
                                                            size_t n, int chunk, float *__restrict__ out2) {
    const int g0 = (int)blockIdx.y * chunk, g1 = min(g0 + chunk, n_part);
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        float s = 0.0f;
        int g = g0;
        for (; g + 8 <= g1; g += 8) {
            float v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = part[(size_t)(g + k) * stride + i];
#pragma unroll
            for (int k = 0; k < 8; k++) s += v[k];
        }
        for (; g < g1; g++) s += part[(size_t)g * stride + i];
        out2[(size_t)blockIdx.y * n + i] = s;
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

// Output layer (1x1, 32 -> 3, FP32) plus the residual: pred = noisy + b + W·D.
// With a target, the L1 loss and the backward pass: dpred = grad_scale *
// sign(pred - target), gD = (W^T dpred) * (D > 0) as bf16, and per block the
// partial sums of gW = dpred D^T, gb = dpred and |pred - target| in
// part[blockIdx.x][kOutParts] (fixed-order wavefront and block reductions, no
// atomics; reduce_partials_kernel adds the blocks).
constexpr int kOutC = 32;
constexpr int kOutParts = 3 * kOutC + 3 + 1;  // gW, gb, loss
template <bool TRAIN>
__global__ __launch_bounds__(256) void out_layer_kernel(const bf16 *__restrict__ D, const float *__restrict__ noisy,
                                                        const float *__restrict__ w, const float *__restrict__ b,
                                                        float *__restrict__ pred, const float *__restrict__ target,
                                                        size_t n_px, float grad_scale, bf16 *__restrict__ gd,
                                                        float *__restrict__ part) {
    __shared__ float red[4][kOutParts];
    float lw[TRAIN ? kOutParts : 1];
    if (TRAIN)
        for (int i = 0; i < kOutParts; i++) lw[i] = 0.0f;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < n_px; p += (size_t)gridDim.x * 256) {
        float d[kOutC];
#pragma unroll
        for (int i = 0; i < kOutC; i += 8) {
            const uint4 v = *reinterpret_cast<const uint4 *>(D + p * kOutC + i);
            unpack8(v, d + i);
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
                lw[kOutParts - 1] += fabsf(diff);
                g[c] = diff > 0.0f ? grad_scale : (diff < 0.0f ? -grad_scale : 0.0f);
                lw[3 * kOutC + c] += g[c];
            }
#pragma unroll
            for (int i = 0; i < kOutC; i += 8) {
                float h[8];
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int k = i + j;
                    float s = 0.0f;
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        s += w[c * kOutC + k] * g[c];
                        lw[c * kOutC + k] += g[c] * d[k];
                    }
                    h[j] = d[k] > 0.0f ? s : 0.0f;
                }
                *reinterpret_cast<uint4 *>(gd + p * kOutC + i) = pack8(h);
            }
        }
    }
    if (!TRAIN) return;
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    for (int i = 0; i < kOutParts; i++) {
        float v = lw[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[wv][i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kOutParts; i += 256)
        part[(size_t)blockIdx.x * kOutParts + i] = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
}

// Features of the network from the renderer's buffers, on a padded image of
// pw x ph (multiples of 16) whose pixels past the w x h frame repeat its last
// row and column. F: 16 bf16 per pixel: log1p rgb of the accumulation,
// first-hit albedo, G-buffer normal and depth, log1p of the luminance variance
// of the samples (E[l^2] - E[l]^2 of the accumulated moments) and log1p of that
// variance times the accumulation's attenuation alpha/(2 - alpha) (the
// residual variance of the mean; it falls as frames add up), then 4 zeros.
// noisy: log1p of the accumulation (FP32 rgb).
__global__ __launch_bounds__(256) void features_kernel(const float4 *__restrict__ accum, const float4 *__restrict__ gbuf,
                                                       const float4 *__restrict__ albedo, const float2 *__restrict__ mom,
                                                       float ema, uint32_t w, uint32_t h, uint32_t pw, uint32_t ph,
                                                       bf16 *__restrict__ F, float *__restrict__ noisy) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)pw * ph) return;
    const uint32_t x = (uint32_t)(p % pw), y = (uint32_t)(p / pw);
    const size_t src = (size_t)min(y, h - 1u) * w + min(x, w - 1u);
    const float4 a = accum[src], g = gbuf[src], al = albedo[src];
    const float2 m = mom[src];
    const float lr = log1pf(fmaxf(a.x, 0.0f)), lg = log1pf(fmaxf(a.y, 0.0f)), lb = log1pf(fmaxf(a.z, 0.0f));
    noisy[p * 3 + 0] = lr;
    noisy[p * 3 + 1] = lg;
    noisy[p * 3 + 2] = lb;
    const float depth = g.w < 1e29f ? log1pf(fminf(fmaxf(g.w, 0.0f), 1e4f)) * 0.25f : -1.0f;
    const float var = fmaxf(m.y - m.x * m.x, 0.0f);
    const float f[16] = {lr, lg, lb, al.x, al.y, al.z, g.x, g.y, g.z, depth, log1pf(var), log1pf(var * ema),
                         0.0f, 0.0f, 0.0f, 0.0f};
    uint4 *dst = reinterpret_cast<uint4 *>(F + p * 16);
    dst[0] = pack8(f);
    dst[1] = pack8(f + 8);
}

// The displayed image: expm1 of the prediction over the w x h frame of the
// padded pw-wide prediction.
__global__ __launch_bounds__(256) void image_kernel(const float *__restrict__ pred, uint32_t w, uint32_t h,
                                                    uint32_t pw, float4 *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)w * h) return;
    const size_t p = (i / w) * pw + i % w;
    out[i] = make_float4(expm1f(pred[p * 3]), expm1f(pred[p * 3 + 1]), expm1f(pred[p * 3 + 2]), 1.0f);
}

// The training loss of one step: the sum of the out-layer blocks' |diff|
// partials (in block order, in double) over the element count.
__global__ void loss_kernel(const float *__restrict__ part, int n_part, double inv_count, double *__restrict__ loss) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    int g = 0;
    for (; g + 8 <= n_part; g += 8) {  // eight loads in flight, added in block order
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = part[(size_t)(g + k) * kOutParts + kOutParts - 1];
#pragma unroll
        for (int k = 0; k < 8; k++) s += (double)v[k];
    }
    for (; g < n_part; g++) s += (double)part[(size_t)g * kOutParts + kOutParts - 1];
    *loss = s * inv_count;
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

// Fixed-order sum of n_part partial arrays (part[g * stride + i], i < n) into
// out: one pass up to kRedChunk * 2 partials, else chunks of kRedChunk summed
// into scratch (reduce_scratch_floats(n_part, n) floats), then the chunks.
constexpr int kRedChunk = 32;
size_t reduce_scratch_floats(int n_part, size_t n) {
    return n_part <= 2 * kRedChunk ? 0 : (size_t)((n_part + kRedChunk - 1) / kRedChunk) * n;
}
void reduce_fixed(const float *part, int n_part, size_t stride, size_t n, float *out, float *scratch, hipStream_t s) {
    if (n_part <= 2 * kRedChunk) {
        hipLaunchKernelGGL(rtu::reduce_partials_kernel, dim3(blocks_for(n, 2048)), dim3(256), 0, s, part, n_part,
                           stride, n, out);
        return;
    }
    const int chunks = (n_part + kRedChunk - 1) / kRedChunk;
    hipLaunchKernelGGL(rtu::reduce_chunks_kernel, dim3(blocks_for(n, 256), chunks), dim3(256), 0, s, part, n_part,
                       stride, n, kRedChunk, scratch);
    hipLaunchKernelGGL(rtu::reduce_partials_kernel, dim3(blocks_for(n, 2048)), dim3(256), 0, s, scratch, chunks, n, n,
                       out);
}

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
// The network described at the top of this file: levels 0..4 of the encoder
// (layers 2l, 2l+1), then decoder levels 3..0 (layers 10 + 2(3 - l), +1).
constexpr int kLevels = 5;
constexpr int kNumLayers = 18;
constexpr int kFeat = 16;
constexpr int kEnc[kLevels] = {32, 64, 128, 128, 128};
constexpr int kDec[kLevels - 1] = {32, 64, 128, 128};
// channels of the concatenation at decoder level l: skip, then the upsampled
// output of the level below (the bottleneck for level 3)
constexpr int cat_up(int l) { return l == kLevels - 2 ? kEnc[kLevels - 1] : kDec[l + 1]; }
constexpr int cat_ch(int l) { return kEnc[l] + cat_up(l); }
constexpr int dec_layer(int l) { return 2 * kLevels + 2 * (kLevels - 2 - l); }
const ConvLayer kLayers[kNumLayers] = {{kFeat, 32}, {32, 32}, {32, 64}, {64, 64}, {64, 128}, {128, 128},
                                       {128, 128}, {128, 128}, {128, 128}, {128, 128}, {256, 128}, {128, 128},
                                       {256, 128}, {128, 128}, {192, 64}, {64, 64}, {96, 32}, {32, 32}};
// image sizes are padded to a multiple of this (four 2x2 pools)
constexpr uint32_t kPadTo = 1u << (kLevels - 1);

// Activations (and their gradients) of one batch shape; H and W multiples of kPadTo.
struct Acts {
    int N = 0, H = 0, W = 0;
    int h(int l) const { return H >> l; }
    int w(int l) const { return W >> l; }
    size_t px(int l = 0) const { return (size_t)N * h(l) * w(l); }
    rtu::bf16 *F = nullptr, *B = nullptr;
    rtu::bf16 *T[kLevels] = {}, *P[kLevels] = {}, *CAT[kLevels - 1] = {}, *U[kLevels - 1] = {},
              *V[kLevels - 1] = {};
    float *noisy = nullptr, *target = nullptr, *pred = nullptr;
    // gradients
    rtu::bf16 *gB = nullptr;
    rtu::bf16 *gT[kLevels] = {}, *gP[kLevels] = {}, *gS[kLevels - 1] = {}, *gCAT[kLevels - 1] = {},
              *gU[kLevels - 1] = {}, *gV[kLevels - 1] = {};
    float *part = nullptr;  // partial sums of the weight/bias/out-layer gradients
    size_t part_floats = 0;
    bool grads = false;
    std::vector<void *> owned;

    void release() {
        for (void *p : owned) (void)hipFree(p);
        *this = Acts{};
    }
    rtu::bf16 *act(size_t n, const char *what) {
        rtu::bf16 *p = dalloc<rtu::bf16>(n, what);
        owned.push_back(p);
        return p;
    }
    void ensure(int n, int hh, int ww, bool with_grads, size_t part_need) {
        if (n == N && hh == H && ww == W && (grads || !with_grads) && part_floats >= part_need) return;
        release();
        N = n;
        H = hh;
        W = ww;
        grads = with_grads;
        F = act(px(0) * kFeat, "unet F");
        for (int l = 0; l < kLevels; l++) {
            T[l] = act(px(l) * kEnc[l], "unet T");
            if (l > 0) P[l] = act(px(l) * kEnc[l - 1], "unet P");
            if (l < kLevels - 1) {
                CAT[l] = act(px(l) * cat_ch(l), "unet CAT");
                U[l] = act(px(l) * kDec[l], "unet U");
                V[l] = act(px(l) * kDec[l], "unet V");
            }
        }
        B = act(px(kLevels - 1) * kEnc[kLevels - 1], "unet B");
        for (float **f : {&noisy, &target, &pred}) {
            *f = dalloc<float>(px(0) * 3, "unet rgb");
            owned.push_back(*f);
        }
        if (with_grads) {
            gB = act(px(kLevels - 1) * kEnc[kLevels - 1], "unet gB");
            for (int l = 0; l < kLevels; l++) {
                gT[l] = act(px(l) * kEnc[l], "unet gT");
                if (l > 0) gP[l] = act(px(l) * kEnc[l - 1], "unet gP");
                if (l < kLevels - 1) {
                    gS[l] = act(px(l) * kEnc[l], "unet gS");
                    gCAT[l] = act(px(l) * cat_ch(l), "unet gCAT");
                    gU[l] = act(px(l) * kDec[l], "unet gU");
                    gV[l] = act(px(l) * kDec[l], "unet gV");
                }
            }
            part = dalloc<float>(part_need, "unet gradient partials");
            owned.push_back(part);
            part_floats = part_need;
        }
    }
};

// Partial arrays of a layer's weight gradient: enough workgroups to fill the
// GPU (about 1024 of three wavefronts, four per CU), at most one per tile.
int wgrad_parts(const ConvLayer &L, int N, int H, int W) {
    const int nblk = (L.cout / 32) * ((L.cin + 31) / 32);
    const int tiles = N * ((H + rtu::kWgRows - 1) / rtu::kWgRows) * ((W + 31) / 32);
    return std::max(1, std::min({1024 / nblk, tiles, 1024}));
}
constexpr int kOutBlocks = 1024;

// Gradient partial floats a training pass of this shape needs.
size_t part_floats_for(int N, int H, int W) {
    size_t need = (size_t)kOutBlocks * rtu::kOutParts + reduce_scratch_floats(kOutBlocks, 3 * rtu::kOutC + 3);
    for (int k = 0; k < kNumLayers; k++) {
        const int l = k < 2 * kLevels ? k / 2 : (kLevels - 2) - (k - 2 * kLevels) / 2;
        const int parts = wgrad_parts(kLayers[k], N, H >> l, W >> l);
        const ConvLayer &L = kLayers[k];
        need = std::max(need, (size_t)parts * (L.nw() + 3 * L.cout) +
                                  std::max(reduce_scratch_floats(parts, L.nw()), reduce_scratch_floats(3 * parts, L.cout)));
    }
    return need;
}

// One convolution launch's arguments (see conv3x3_kernel).
struct ConvArgs {
    const rtu::bf16 *X = nullptr;
    int xs = 0;
    const rtu::bf16 *w = nullptr;
    int cout = 0;
    const float *b = nullptr;
    rtu::bf16 *Y = nullptr;
    int ys = 0, yoff = 0, relu = 0;
    const rtu::bf16 *mask = nullptr;
    int ms = 0, N = 0, H = 0, W = 0;
    const rtu::bf16 *X2 = nullptr;
    int x2s = 0, c1 = 0;
};
typedef void (*ConvLaunch)(const ConvArgs &, hipStream_t);
// cin, cout, src, N, H, W, mask, bias
typedef std::array<int, 8> ConvKey;

template <int CIN, int KC, int NCO, int SRC, bool WONCE, int RPW>
void conv_launch(const ConvArgs &a, hipStream_t s) {
    constexpr size_t lds = rtu::conv_lds_bytes(KC, NCO, RPW);
    static_assert(lds <= 160 * 1024, "LDS of one workgroup");
    // once per instance (a function-local static, initialised once even
    // across threads): raise the dynamic LDS limit (above 64 KB) and size
    // the persistent grid
    static const int resident = [] {
        const void *fn = (const void *)rtu::conv3x3_kernel<CIN, KC, NCO, SRC, WONCE, RPW>;
        if (lds > 65536)
            hip_check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds), "conv LDS size");
        int per_cu = 0;
        hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rtu::conv3x3_kernel<CIN, KC, NCO, SRC, WONCE, RPW>,
                                                               256, lds),
                  "conv occupancy");
        return std::max(per_cu, 1) * cu_count();
    }();
    constexpr int rows = 4 * RPW;
    const int items = ((a.W + 31) / 32) * ((a.H + rows - 1) / rows) * a.N * (a.cout / (32 * NCO));
    const dim3 grid((unsigned)std::max(1, std::min(items, resident)));
    hipLaunchKernelGGL((rtu::conv3x3_kernel<CIN, KC, NCO, SRC, WONCE, RPW>), grid, dim3(256), lds, s, a.X, a.xs, a.w, a.cout,
                       a.b, a.Y, a.ys, a.yoff, a.relu, a.mask, a.ms, a.H, a.W, a.X2, a.x2s, a.c1, a.N);
}

// The compiled variants: input channels, slice width KC, output-channel block
// 32 * NCO, staging source, weights staged once (one slice, one block). The
// first listed for a shape is the choice without tuning.
struct ConvVariant {
    int cin, kc, nco, src;
    bool wonce;
    int rpw;
    ConvLaunch fn;
};
#define RT_CV(CI, KC, NCO, SRC, ONCE) {CI, KC, NCO, rtu::SRC, ONCE, 2, conv_launch<CI, KC, NCO, rtu::SRC, ONCE, 2>}
#define RT_CV4(CI, KC, NCO, SRC, ONCE) {CI, KC, NCO, rtu::SRC, ONCE, 4, conv_launch<CI, KC, NCO, rtu::SRC, ONCE, 4>}
#define RT_CV1(CI, KC, NCO, SRC, ONCE) {CI, KC, NCO, rtu::SRC, ONCE, 1, conv_launch<CI, KC, NCO, rtu::SRC, ONCE, 1>}
const ConvVariant kConvVariants[] = {
    RT_CV(16, 16, 1, SRC_PLAIN, true),
    RT_CV(32, 32, 1, SRC_PLAIN, true),    RT_CV(32, 16, 1, SRC_PLAIN, false),
    RT_CV(32, 32, 2, SRC_PLAIN, true),    RT_CV(32, 16, 2, SRC_PLAIN, false),
    RT_CV(32, 32, 3, SRC_PLAIN, true),    RT_CV(32, 16, 3, SRC_PLAIN, false),
    RT_CV(32, 32, 2, SRC_POOL, true),     RT_CV(32, 16, 2, SRC_POOL, false),   RT_CV(32, 16, 1, SRC_POOL, false),
    RT_CV(64, 64, 1, SRC_PLAIN, true),    RT_CV(64, 32, 1, SRC_PLAIN, false),  RT_CV(64, 16, 1, SRC_PLAIN, false),
    RT_CV(64, 64, 2, SRC_PLAIN, true),    RT_CV(64, 32, 2, SRC_PLAIN, false),  RT_CV(64, 16, 2, SRC_PLAIN, false),
    RT_CV(64, 32, 4, SRC_PLAIN, false),   RT_CV(64, 16, 4, SRC_PLAIN, false),
    RT_CV(96, 96, 1, SRC_PLAIN, true),    RT_CV(96, 32, 1, SRC_PLAIN, false),  RT_CV(96, 16, 1, SRC_PLAIN, false),
    RT_CV(96, 96, 1, SRC_CONCAT, true),   RT_CV(96, 32, 1, SRC_CONCAT, false), RT_CV(96, 16, 1, SRC_CONCAT, false),
    RT_CV(128, 32, 4, SRC_PLAIN, false),  RT_CV(128, 16, 4, SRC_PLAIN, false), RT_CV(128, 32, 2, SRC_PLAIN, false),
    RT_CV(128, 16, 2, SRC_PLAIN, false),  RT_CV(128, 32, 1, SRC_PLAIN, false),
    RT_CV(192, 32, 2, SRC_PLAIN, false),  RT_CV(192, 16, 2, SRC_PLAIN, false), RT_CV(192, 32, 1, SRC_PLAIN, false),
    RT_CV(192, 32, 2, SRC_CONCAT, false), RT_CV(192, 16, 2, SRC_CONCAT, false), RT_CV(192, 32, 1, SRC_CONCAT, false),
    RT_CV(256, 32, 4, SRC_PLAIN, false),  RT_CV(256, 16, 4, SRC_PLAIN, false), RT_CV(256, 32, 2, SRC_PLAIN, false),
    RT_CV(256, 32, 4, SRC_CONCAT, false), RT_CV(256, 16, 4, SRC_CONCAT, false), RT_CV(256, 32, 2, SRC_CONCAT, false),
    // sixteen-row tiles (four rows per wavefront)
    RT_CV4(16, 16, 1, SRC_PLAIN, true),   RT_CV4(32, 32, 1, SRC_PLAIN, true),   RT_CV4(32, 16, 2, SRC_PLAIN, false),
    RT_CV4(32, 16, 1, SRC_POOL, false),   RT_CV4(64, 16, 2, SRC_PLAIN, false),  RT_CV4(64, 32, 1, SRC_PLAIN, false),
    RT_CV4(96, 16, 1, SRC_CONCAT, false), RT_CV4(96, 32, 1, SRC_CONCAT, false), RT_CV4(96, 16, 1, SRC_PLAIN, false),
    RT_CV4(128, 16, 2, SRC_PLAIN, false), RT_CV4(128, 32, 1, SRC_PLAIN, false), RT_CV4(192, 16, 2, SRC_CONCAT, false),
    RT_CV4(192, 16, 2, SRC_PLAIN, false), RT_CV4(256, 16, 2, SRC_CONCAT, false), RT_CV4(256, 16, 2, SRC_PLAIN, false),
    // four-row tiles (one row per wavefront): twice the work items of eight-row
    // tiles, for the levels too small to fill the chip
    RT_CV1(128, 16, 2, SRC_PLAIN, false), RT_CV1(128, 32, 1, SRC_PLAIN, false), RT_CV1(128, 16, 4, SRC_PLAIN, false),
    RT_CV1(256, 16, 2, SRC_CONCAT, false), RT_CV1(256, 16, 2, SRC_PLAIN, false),
};
#undef RT_CV
#undef RT_CV4
#undef RT_CV1

}  // namespace

struct rt_unet {
    int device = 0;
    hipStream_t stream = nullptr;
    // FP32 master parameters: conv weights [9][cout][cin] and biases per layer,
    // then the output layer's w[3][32] and b[3]; one contiguous array.
    size_t n_params = 0;
    size_t w_off[kNumLayers] = {}, b_off[kNumLayers] = {}, ow_off = 0, ob_off = 0;
    float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
    rtu::bf16 *d_wf[kNumLayers] = {}, *d_wt[kNumLayers] = {};
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
    // convolution variant chosen per shape (autotuned at first use)
    std::map<ConvKey, int> conv_choice;
    int variant_mode = 1;  // 0 first listed, 1 autotuned, k >= 2 the (k-2)-th that fits
    // inference: per (level, N, H, W), whether the level's pool (level 1) or
    // upsample + concatenation (decoder levels) is folded into the next
    // convolution's staging (true) or run as its own kernel before a plain
    // convolution; both forms give the same bits. Timed at first use in
    // autotuned mode, folded in every other mode.
    std::map<std::array<int, 4>, bool> fold_choice;

    ~rt_unet() {
        if (stream) (void)hipStreamSynchronize(stream);
        inf.release();
        train.release();
        void *all[] = {d_p, d_g, d_m, d_v, ds_F, ds_noisy, ds_target, d_sel, d_loss};
        for (void *p : all)
            if (p) (void)hipFree(p);
        for (int k = 0; k < kNumLayers; k++) {
            if (d_wf[k]) (void)hipFree(d_wf[k]);
            if (d_wt[k]) (void)hipFree(d_wt[k]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    void refresh_bf16() {
        for (int k = 0; k < kNumLayers; k++) {
            const ConvLayer &L = kLayers[k];
            hipLaunchKernelGGL(rtu::weights_bf16_kernel, dim3(blocks_for(L.nw())), dim3(256), 0, stream,
                               d_p + w_off[k], L.nw(), d_wf[k], d_wt[k], L.cout, L.cin);
        }
        hip_check(hipGetLastError(), "weights to bf16");
    }

    // Layer k forward (bias, ReLU) or backward-data (bwd: transposed weights,
    // no bias, mask = the ReLU derivative of the layer's input, or none).
    // src: rtu::SRC_POOL (X at 2H x 2W, pooled while staged) or rtu::SRC_CONCAT
    // (channels from c1 on read from X2, half resolution) for the inference pass.
    void conv(int k, bool bwd, const rtu::bf16 *X, int xs, rtu::bf16 *Y, int ys, int yoff, const rtu::bf16 *mask,
              int ms, int N, int H, int W, int src = rtu::SRC_PLAIN, const rtu::bf16 *X2 = nullptr, int x2s = 0,
              int c1 = 0) {
        const ConvLayer &L = kLayers[k];
        ConvArgs a;
        a.X = X;
        a.xs = xs;
        a.w = bwd ? d_wt[k] : d_wf[k];
        a.cout = bwd ? L.cin : L.cout;
        a.b = bwd ? nullptr : d_p + b_off[k];
        a.Y = Y;
        a.ys = ys;
        a.yoff = yoff;
        a.relu = bwd ? 0 : 1;
        a.mask = mask;
        a.ms = ms;
        a.N = N;
        a.H = H;
        a.W = W;
        a.X2 = X2;
        a.x2s = x2s;
        a.c1 = c1;
        run_conv(bwd ? L.cout : L.cin, src, a);
    }

    // The variants compiled for (cin, src) that can produce a.cout channels;
    // with more than one, the fastest on this shape is timed once and kept.
    // All variants sum in the same order, so the choice never changes a bit.
    void run_conv(int cin, int src, const ConvArgs &a) {
        std::vector<int> cands;
        for (int i = 0; i < (int)(sizeof(kConvVariants) / sizeof(kConvVariants[0])); i++) {
            const ConvVariant &v = kConvVariants[i];
            if (v.cin == cin && v.src == src && a.cout % (32 * v.nco) == 0 && (!v.wonce || a.cout == 32 * v.nco))
                cands.push_back(i);
        }
        if (cands.empty()) throw ApiError(RT_ERR_INVALID, "no convolution variant for this channel count");
        int pick = cands[0];
        if (variant_mode >= 2) {
            pick = cands[std::min((size_t)(variant_mode - 2), cands.size() - 1)];
        } else if (cands.size() > 1 && variant_mode == 1) {
            const ConvKey key{cin, a.cout, src, a.N, a.H, a.W, a.mask ? 1 : 0, a.b ? 1 : 0};
            auto it = conv_choice.find(key);
            if (it != conv_choice.end()) {
                pick = it->second;
            } else {
                pick = tune_conv(cands, a);
                conv_choice[key] = pick;
            }
        }
        kConvVariants[pick].fn(a, stream);
    }

    // timed launches per candidate (two gave picks that changed from run to run
    // on the small layers)
    static constexpr int kTuneLaunches = 6;
    int tune_conv(const std::vector<int> &cands, const ConvArgs &a) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hip_check(hipEventCreate(&e0), "tuning event");
        hip_check(hipEventCreate(&e1), "tuning event");
        int best = cands[0];
        float best_ms = 1e30f;
        try {
            for (int c : cands) {
                kConvVariants[c].fn(a, stream);  // first launch: one-time set-up, caches
                hip_check(hipEventRecord(e0, stream), "tuning event");
                for (int rep = 0; rep < kTuneLaunches; rep++) kConvVariants[c].fn(a, stream);
                hip_check(hipEventRecord(e1, stream), "tuning event");
                hip_check(hipEventSynchronize(e1), "tuning sync");
                float ms = 0.0f;
                hip_check(hipEventElapsedTime(&ms, e0, e1), "tuning time");
                if (ms < best_ms) {
                    best_ms = ms;
                    best = c;
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return best;
    }

    // Weight and bias gradients of layer k from its output gradient G and its input X.
    void wgrad(int k, const rtu::bf16 *G, const rtu::bf16 *X, int xs, Acts &a, int N, int H, int W) {
        const ConvLayer &L = kLayers[k];
        const int parts = wgrad_parts(L, N, H, W);
        // weight partials [parts][nw], then bias partials [parts][3 wavefronts][cout],
        // then the reductions' scratch
        float *bpart = a.part + (size_t)parts * L.nw();
        float *scratch = bpart + (size_t)parts * 3 * L.cout;
        if ((size_t)parts * (L.nw() + 3 * L.cout) +
                std::max(reduce_scratch_floats(parts, L.nw()), reduce_scratch_floats(3 * parts, L.cout)) >
            a.part_floats)
            throw ApiError(RT_ERR_INVALID, "gradient partials larger than their buffer");
        const dim3 grid((L.cout / 32) * ((L.cin + 31) / 32), (unsigned)parts);
        hipLaunchKernelGGL(rtu::conv3x3_wgrad_kernel, grid, dim3(192), 0, stream, G, L.cout, L.cout, X, xs, L.cin,
                           a.part, bpart, N, H, W);
        reduce_fixed(a.part, parts, L.nw(), L.nw(), d_g + w_off[k], scratch, stream);
        reduce_fixed(bpart, 3 * parts, (size_t)L.cout, (size_t)L.cout, d_g + b_off[k], scratch, stream);
    }

    void pool(const rtu::bf16 *X, int xs, int c, rtu::bf16 *P, int N, int H, int W) {  // H, W: input size
        hipLaunchKernelGGL(rtu::pool_kernel, dim3((unsigned)((W / 2 * (c / 8) + 255) / 256), (unsigned)(N * H / 2)),
                           dim3(256), 0, stream, X, xs, c, P, H, W);
    }
    void upsample(const rtu::bf16 *Bsrc, int c, rtu::bf16 *Y, int ys, int off, int N, int H, int W) {  // H, W: output
        hipLaunchKernelGGL(rtu::upsample_kernel, dim3((unsigned)((W * (c / 8) + 255) / 256), (unsigned)(N * H)),
                           dim3(256), 0, stream, Bsrc, c, Y, ys, off, H, W);
    }

    // The inference convolution that reads a resampled input: which 1 is layer
    // 2 over the pooled level-0 output, which -1 - l the first convolution of
    // decoder level l over [skip, upsampled]. fold: stage the pool or the
    // upsample inside the convolution; otherwise run pool_kernel /
    // upsample_kernel into P[1] / CAT[l] first and convolve that.
    void level_input(Acts &a, int which, bool fold) {
        const int N = a.N;
        if (which == 1) {
            if (fold) {
                conv(2, false, a.CAT[0], cat_ch(0), a.T[1], kEnc[1], 0, nullptr, 0, N, a.h(1), a.w(1),
                     rtu::SRC_POOL);
            } else {
                pool(a.CAT[0], cat_ch(0), kEnc[0], a.P[1], N, a.h(0), a.w(0));
                conv(2, false, a.P[1], kEnc[0], a.T[1], kEnc[1], 0, nullptr, 0, N, a.h(1), a.w(1));
            }
            return;
        }
        const int l = -1 - which, k = dec_layer(l);
        const rtu::bf16 *below = l == kLevels - 2 ? a.B : a.V[l + 1];
        const int bc = cat_up(l);
        if (fold) {
            conv(k, false, a.CAT[l], cat_ch(l), a.U[l], kDec[l], 0, nullptr, 0, N, a.h(l), a.w(l), rtu::SRC_CONCAT,
                 below, bc, kEnc[l]);
        } else {
            upsample(below, bc, a.CAT[l], cat_ch(l), kEnc[l], N, a.h(l), a.w(l));
            conv(k, false, a.CAT[l], cat_ch(l), a.U[l], kDec[l], 0, nullptr, 0, N, a.h(l), a.w(l));
        }
    }

    // Folded or not for this level and shape: timed once in autotuned mode
    // (each form's first run tunes its convolution, then kTuneLaunches runs
    // of each are timed); the timing runs only rewrite buffers that the pass
    // writes again right after.
    bool fold_here(Acts &a, int which) {
        if (variant_mode != 1) return true;
        const std::array<int, 4> key{which, a.N, a.H, a.W};
        auto it = fold_choice.find(key);
        if (it != fold_choice.end()) return it->second;
        level_input(a, which, true);
        level_input(a, which, false);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hip_check(hipEventCreate(&e0), "tuning event");
        hip_check(hipEventCreate(&e1), "tuning event");
        float ms[2] = {0.0f, 0.0f};
        try {
            for (int f = 0; f < 2; f++) {
                hip_check(hipEventRecord(e0, stream), "tuning event");
                for (int rep = 0; rep < kTuneLaunches; rep++) level_input(a, which, f == 0);
                hip_check(hipEventRecord(e1, stream), "tuning event");
                hip_check(hipEventSynchronize(e1), "tuning sync");
                hip_check(hipEventElapsedTime(&ms[f], e0, e1), "tuning time");
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const bool fold = ms[0] <= ms[1];
        fold_choice[key] = fold;
        return fold;
    }

    // Forward pass over a.F / a.noisy. Training (with_target) keeps the pooled
    // and concatenated tensors for the backward pass, writes the loss partials
    // and runs the output layer's backward pass into a.gV[0] and the
    // output-layer gradients; inference folds the first pool and every
    // upsample + concatenation into the convolutions' staging.
    void forward(Acts &a, bool with_target, double *loss_slot) {
        const int N = a.N;
        conv(0, false, a.F, kFeat, a.T[0], kEnc[0], 0, nullptr, 0, N, a.h(0), a.w(0));
        conv(1, false, a.T[0], kEnc[0], a.CAT[0], cat_ch(0), 0, nullptr, 0, N, a.h(0), a.w(0));
        for (int l = 1; l < kLevels; l++) {
            const int cs = kEnc[l - 1];
            if (!with_target && l == 1) {
                level_input(a, 1, fold_here(a, 1));
            } else {
                pool(a.CAT[l - 1], cat_ch(l - 1), cs, a.P[l], N, a.h(l - 1), a.w(l - 1));
                conv(2 * l, false, a.P[l], cs, a.T[l], kEnc[l], 0, nullptr, 0, N, a.h(l), a.w(l));
            }
            rtu::bf16 *out = l < kLevels - 1 ? a.CAT[l] : a.B;
            const int os = l < kLevels - 1 ? cat_ch(l) : kEnc[l];
            conv(2 * l + 1, false, a.T[l], kEnc[l], out, os, 0, nullptr, 0, N, a.h(l), a.w(l));
        }
        for (int l = kLevels - 2; l >= 0; l--) {
            const int k = dec_layer(l);
            const rtu::bf16 *below = l == kLevels - 2 ? a.B : a.V[l + 1];
            const int bc = cat_up(l);
            if (with_target) {
                upsample(below, bc, a.CAT[l], cat_ch(l), kEnc[l], N, a.h(l), a.w(l));
                conv(k, false, a.CAT[l], cat_ch(l), a.U[l], kDec[l], 0, nullptr, 0, N, a.h(l), a.w(l));
            } else {
                level_input(a, -1 - l, fold_here(a, -1 - l));
            }
            conv(k + 1, false, a.U[l], kDec[l], a.V[l], kDec[l], 0, nullptr, 0, N, a.h(l), a.w(l));
        }
        const float scale = 1.0f / (3.0f * (float)a.px());
        if (with_target) {
            hipLaunchKernelGGL(rtu::out_layer_kernel<true>, dim3(kOutBlocks), dim3(256), 0, stream, a.V[0], a.noisy,
                               d_p + ow_off, d_p + ob_off, a.pred, a.target, a.px(), scale, a.gV[0], a.part);
            // output-layer gradients (w then b, contiguous) and the loss from the block partials
            reduce_fixed(a.part, kOutBlocks, (size_t)rtu::kOutParts, (size_t)(3 * rtu::kOutC + 3), d_g + ow_off,
                         a.part + (size_t)kOutBlocks * rtu::kOutParts, stream);
            hipLaunchKernelGGL(rtu::loss_kernel, dim3(1), dim3(64), 0, stream, a.part, kOutBlocks,
                               1.0 / (3.0 * (double)a.px()), loss_slot);
        } else {
            hipLaunchKernelGGL(rtu::out_layer_kernel<false>, dim3(blocks_for(a.px())), dim3(256), 0, stream, a.V[0],
                               a.noisy, d_p + ow_off, d_p + ob_off, a.pred, (const float *)nullptr, a.px(), scale,
                               (rtu::bf16 *)nullptr, (float *)nullptr);
        }
        hip_check(hipGetLastError(), "unet forward");
    }

    void backward(Acts &a) {
        const int N = a.N;
        // decoder, output side first
        for (int l = 0; l < kLevels - 1; l++) {
            const int k = dec_layer(l);
            if (l > 0)  // V_l's gradient: the 2x2 sums of the upsampled part of CAT_{l-1}'s gradient
                hipLaunchKernelGGL(rtu::upsample_back_kernel,
                                   dim3((unsigned)((a.w(l) * (kDec[l] / 8) + 255) / 256), (unsigned)(N * a.h(l))),
                                   dim3(256), 0, stream, a.gCAT[l - 1], cat_ch(l - 1), kEnc[l - 1], kDec[l], a.gV[l],
                                   a.h(l - 1), a.w(l - 1));
            wgrad(k + 1, a.gV[l], a.U[l], kDec[l], a, N, a.h(l), a.w(l));
            conv(k + 1, true, a.gV[l], kDec[l], a.gU[l], kDec[l], 0, a.U[l], kDec[l], N, a.h(l), a.w(l));
            wgrad(k, a.gU[l], a.CAT[l], cat_ch(l), a, N, a.h(l), a.w(l));
            conv(k, true, a.gU[l], kDec[l], a.gCAT[l], cat_ch(l), 0, a.CAT[l], cat_ch(l), N, a.h(l), a.w(l));
        }
        // the bottleneck's gradient, then the encoder from the bottom up
        const int lb = kLevels - 1;
        hipLaunchKernelGGL(rtu::upsample_back_kernel,
                           dim3((unsigned)((a.w(lb) * (kEnc[lb] / 8) + 255) / 256), (unsigned)(N * a.h(lb))),
                           dim3(256), 0, stream, a.gCAT[lb - 1], cat_ch(lb - 1), kEnc[lb - 1], kEnc[lb], a.gB,
                           a.h(lb - 1), a.w(lb - 1));
        for (int l = kLevels - 1; l >= 0; l--) {
            const rtu::bf16 *gx = l == kLevels - 1 ? a.gB : a.gS[l];
            wgrad(2 * l + 1, gx, a.T[l], kEnc[l], a, N, a.h(l), a.w(l));
            conv(2 * l + 1, true, gx, kEnc[l], a.gT[l], kEnc[l], 0, a.T[l], kEnc[l], N, a.h(l), a.w(l));
            const rtu::bf16 *xin = l == 0 ? a.F : a.P[l];
            const int xc = l == 0 ? kFeat : kEnc[l - 1];
            wgrad(2 * l, a.gT[l], xin, xc, a, N, a.h(l), a.w(l));
            if (l == 0) break;
            conv(2 * l, true, a.gT[l], kEnc[l], a.gP[l], xc, 0, nullptr, 0, N, a.h(l), a.w(l));
            // S_{l-1}'s gradient: its part of CAT_{l-1}'s gradient plus the pool's share, masked by S
            hipLaunchKernelGGL(rtu::pool_back_kernel,
                               dim3((unsigned)((a.w(l - 1) * (xc / 8) + 255) / 256), (unsigned)(N * a.h(l - 1))),
                               dim3(256), 0, stream, a.gCAT[l - 1], cat_ch(l - 1), a.gP[l], a.CAT[l - 1],
                               cat_ch(l - 1), xc, a.gS[l - 1], a.h(l - 1), a.w(l - 1));
        }
        hip_check(hipGetLastError(), "unet backward");
    }

    void adam(float lr) {
        adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, (float)adam_t), c2 = 1.0f - std::pow(b2, (float)adam_t);
        for (int k = 0; k < kNumLayers; k++) {
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

    // Features of a frame into a (padded) activation set whose level-0 size
    // is the frame's rounded up to kPadTo.
    void features(const rtu::FrameViews &v, rtu::bf16 *F, float *noisy, uint32_t pw, uint32_t ph) {
        const size_t n = (size_t)pw * ph;
        hipLaunchKernelGGL(rtu::features_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                           reinterpret_cast<const float4 *>(v.accum), reinterpret_cast<const float4 *>(v.gbuf),
                           reinterpret_cast<const float4 *>(v.albedo), reinterpret_cast<const float2 *>(v.moments),
                           v.ema, v.width, v.height, pw, ph, F, noisy);
    }
};

namespace {

rt_unet *create_unet(int device, uint64_t seed) {
    hip_check(hipSetDevice(device), "hipSetDevice");
    auto u = std::make_unique<rt_unet>();
    u->device = device;
    hip_check(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking), "hipStreamCreate");
    size_t off = 0;
    for (int k = 0; k < kNumLayers; k++) {
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
    for (int k = 0; k < kNumLayers; k++) {
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
    for (int k = 0; k < kNumLayers; k++) {
        u->d_wf[k] = dalloc<rtu::bf16>(kLayers[k].nw(), "unet bf16 weights");
        u->d_wt[k] = dalloc<rtu::bf16>(kLayers[k].nw(), "unet bf16 weights (backward)");
    }
    hip_check(hipDeviceSynchronize(), "sync");
    u->refresh_bf16();
    hip_check(hipStreamSynchronize(u->stream), "sync");
    return u.release();
}

void check_views(const rtu::FrameViews &v) {
    if (!v.accum || !v.gbuf || !v.albedo || !v.moments) throw ApiError(RT_ERR_INVALID, "frame views incomplete");
    if (v.width == 0 || v.height == 0) throw ApiError(RT_ERR_INVALID, "empty frame");
}

uint32_t padded(uint32_t n) { return (n + kPadTo - 1) / kPadTo * kPadTo; }

}  // namespace

namespace rtu {

void denoise(rt_unet *u, const FrameViews &v, void *out, hipStream_t stream) {
    check_views(v);
    // the network runs on the caller's stream
    hipStream_t saved = u->stream;
    u->stream = stream;
    try {
        const uint32_t pw = padded(v.width), ph = padded(v.height);
        u->inf.ensure(1, (int)ph, (int)pw, false, 0);
        u->features(v, u->inf.F, u->inf.noisy, pw, ph);
        u->forward(u->inf, false, nullptr);
        const size_t n = (size_t)v.width * v.height;
        hipLaunchKernelGGL(image_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, u->inf.pred,
                           v.width, v.height, pw, reinterpret_cast<float4 *>(out));
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
        // new parameters restart the optimizer: Adam's moments and step count
        hip_check(hipMemset(u->d_m, 0, n * 4), "clear Adam moments");
        hip_check(hipMemset(u->d_v, 0, n * 4), "clear Adam moments");
        u->adam_t = 0;
        u->refresh_bf16();
        hip_check(hipStreamSynchronize(u->stream), "sync");
    });
}

int rt_unet_dataset_reset(rt_unet *u, uint32_t width, uint32_t height, uint32_t capacity) {
    return rt::api_call([&] {
        if (!u) throw ApiError(RT_ERR_INVALID, "null argument");
        if (width % 32 || height % 32 || width < 32 || height < 32 || capacity == 0)
            throw ApiError(RT_ERR_INVALID, "dataset images must be multiples of 32 in size, capacity > 0");
        hip_check(hipStreamSynchronize(u->stream), "sync");
        for (void *p : {(void *)u->ds_F, (void *)u->ds_noisy, (void *)u->ds_target})
            if (p) (void)hipFree(p);
        u->ds_F = nullptr;
        u->ds_noisy = u->ds_target = nullptr;
        const size_t px = (size_t)width * height * capacity;
        u->ds_F = dalloc<rtu::bf16>(px * kFeat, "dataset features");
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
        if (v.width != u->ds_w || v.height != u->ds_h)
            throw ApiError(RT_ERR_INVALID, "image size differs from the dataset");
        const size_t n = (size_t)v.width * v.height, base = (size_t)slot * n;
        if (which == 0) {
            check_views(v);
            u->features(v, u->ds_F + base * kFeat, u->ds_noisy + base * 3, v.width, v.height);
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
        u->train.ensure((int)batch, (int)crop, (int)crop, true, part_floats_for((int)batch, (int)crop, (int)crop));
        const size_t sel_n = (size_t)3 * batch * std::max(steps, 1u);
        if (u->sel_cap < sel_n) {
            if (u->d_sel) (void)hipFree(u->d_sel);
            u->d_sel = nullptr;
            u->d_sel = dalloc<int>(sel_n, "crop selection");
            u->sel_cap = sel_n;
        }
        if (u->loss_cap < steps) {
            if (u->d_loss) (void)hipFree(u->d_loss);
            u->d_loss = nullptr;
            u->d_loss = dalloc<double>(steps, "losses");
            u->loss_cap = steps;
        }
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
            // every gradient is overwritten (reduced from its partials), so none is cleared
            u->forward(u->train, true, u->d_loss + s);
            u->backward(u->train);
            u->adam(lr);
        }
        hip_check(hipStreamSynchronize(u->stream), "train sync");
        if (losses && steps) {
            std::vector<double> l(steps);
            hip_check(hipMemcpy(l.data(), u->d_loss, 8ull * steps, hipMemcpyDeviceToHost), "read losses");
            for (uint32_t s = 0; s < steps; s++) losses[s] = (float)l[s];
        }
    });
}

int rt_unet_set_variant_mode(rt_unet *u, int mode) {
    return rt::api_call([&] {
        if (!u || mode < 0) throw ApiError(RT_ERR_INVALID, "null argument or negative mode");
        hip_check(hipStreamSynchronize(u->stream), "sync");
        u->variant_mode = mode;
    });
}

// The convolution variant chosen for each shape so far, one line per shape:
// "cin cout src N H W mask bias -> kc nco wonce rows_per_wavefront".
int rt_unet_conv_choices(const rt_unet *u, char *buf, size_t cap, size_t *needed) {
    return rt::api_call([&] {
        if (!u) throw ApiError(RT_ERR_INVALID, "null argument");
        std::string out;
        for (const auto &kv : u->conv_choice) {
            const ConvKey &k = kv.first;
            const ConvVariant &v = kConvVariants[kv.second];
            char line[160];
            std::snprintf(line, sizeof(line), "%d %d %d %d %d %d %d %d -> %d %d %d %d\n", k[0], k[1], k[2], k[3],
                          k[4], k[5], k[6], k[7], v.kc, v.nco, v.wonce ? 1 : 0, v.rpw);
            out += line;
        }
        for (const auto &kv : u->fold_choice) {
            char line[96];
            std::snprintf(line, sizeof(line), "fold %d %d %d %d -> %d\n", kv.first[0], kv.first[1], kv.first[2],
                          kv.first[3], kv.second ? 1 : 0);
            out += line;
        }
        if (needed) *needed = out.size() + 1;
        if (buf && cap) {
            const size_t n = std::min(cap - 1, out.size());
            std::memcpy(buf, out.data(), n);
            buf[n] = '\0';
        }
    });
}

// Test hook: one forward (and backward) pass on caller-supplied host tensors
// (h and w multiples of 16, features padded to 16 channels).
int rt_unet_eval(rt_unet *u, uint32_t n, uint32_t h, uint32_t w, const uint16_t *features, const float *noisy,
                 const float *target, float *pred, float *grads, double *loss) {
    return rt::api_call([&] {
        if (!u || !features || !noisy || !pred) throw ApiError(RT_ERR_INVALID, "null argument");
        if (h % kPadTo || w % kPadTo || n == 0 || h == 0 || w == 0)
            throw ApiError(RT_ERR_INVALID, "image sizes must be multiples of 16");
        Acts &a = u->train;
        a.ensure((int)n, (int)h, (int)w, true, part_floats_for((int)n, (int)h, (int)w));
        const size_t px = (size_t)n * h * w;
        hip_check(hipStreamSynchronize(u->stream), "sync");
        hip_check(hipMemcpy(a.F, features, px * kFeat * 2, hipMemcpyHostToDevice), "upload features");
        hip_check(hipMemcpy(a.noisy, noisy, px * 12, hipMemcpyHostToDevice), "upload inputs");
        if (target) hip_check(hipMemcpy(a.target, target, px * 12, hipMemcpyHostToDevice), "upload targets");
        if (u->loss_cap < 1) {
            u->d_loss = dalloc<double>(1, "loss");
            u->loss_cap = 1;
        }
        u->forward(a, target != nullptr, u->d_loss);
        if (target) u->backward(a);
        hip_check(hipStreamSynchronize(u->stream), "sync");
        hip_check(hipMemcpy(pred, a.pred, px * 12, hipMemcpyDeviceToHost), "read pred");
        if (target && grads) hip_check(hipMemcpy(grads, u->d_g, u->n_params * 4, hipMemcpyDeviceToHost), "read grads");
        if (target && loss) hip_check(hipMemcpy(loss, u->d_loss, 8, hipMemcpyDeviceToHost), "read loss");
    });
}

}  // extern "C"
