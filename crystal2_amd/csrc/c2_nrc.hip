// Radiance-cache MLP on gfx950: MFMA (v_mfma_f32_16x16x32_f16) inference, a fused training
// forward + backward, split-K weight gradients, and the gradient reduction fused with FP32
// Adam. See c2_nrc.h for the network, the query format and the fragment layouts.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "c2_host.h"
#include "c2_nrc.h"

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef float float4_t __attribute__((ext_vector_type(4)));

namespace c2 {

#define NRC_CHECK(x)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) fail(C2_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));  \
    } while (0)

constexpr int kNrcBlock = 256;        // 4 wavefronts
constexpr int kInferTiles = 4;        // inference: 4 x 16 queries per wavefront
constexpr int kPartial = kNrcW * kNrcW + kNrcW;  // dW + db of one layer
constexpr int kRowH = 72;             // LDS row stride in halves: 64 + 8 (16-byte reads of 16 rows
                                      // fall on distinct banks)

// v_mfma_f32_16x16x32_f16, lane l = 16 g + c (c < 16):
//   A (16 x 32): A[c][8 g + j], j < 8      B (32 x 16): B[8 g + j][c]
//   D (16 x 16): D[4 g + i][c], i < 4
__device__ __forceinline__ float4_t mfma32(half8_t a, half8_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// K order of a 32-feature k-tile: slot j of lane group g holds feature
// kfeat(t, g, j) = 32 t + 16 (j / 4) + 4 g + j % 4. With it, the D fragments of the output
// tiles 2t and 2t + 1 of one layer, concatenated per lane, are the next layer's B
// fragment of k-tile t: activations chain through the layers in registers. The weight
// matrices are staged into LDS with their K columns in the same order.
__device__ __forceinline__ int kfeat(int t, int g, int j) { return 32 * t + 16 * (j >> 2) + 4 * g + (j & 3); }

// Training slice js of ns over the *countp rows collected on the device: equal slices of
// ceil(count / ns) rows (at most nfixed), the last one shorter. Without a device count
// (countp == nullptr) the batch is rows [0, nfixed). Returns the row count, sets *off.
__device__ __forceinline__ int slice_rows(const int32_t* countp, int32_t js, int32_t ns, int32_t nfixed, int* off) {
    if (!countp) {
        *off = 0;
        return nfixed;
    }
    const int count = *countp;
    const int per = min((count + ns - 1) / ns, nfixed);
    *off = js * per;
    return min(max(count - *off, 0), per);
}

// B fragments (both k-tiles) of the 64-dim input encoding of one packed query (c2_nrc.h:
// q0 = position and roughness, qg = q1 for lane group 0, q2 for lane group 2) for lane group
// g. Features (the order of the weight rows of layer 0):
//   16 c + 2 b + {0, 1}: sin / cos (2^b pi u_c), u_c = the position normalised to [0, 1], b < 8
//   48 / 50: azimuth of the outgoing direction / normal, atan2(z, x) / 2 pi + 0.5,
//   49 / 51: their cos theta * 0.5 + 0.5 (q1, computed by whoever writes the query)
//   52 + k: one-blob roughness, exp(-((rough - (k + 0.5) / 4) * 4)^2), k < 4
//   56..58: albedo, 59: 1 (q2), 60..63: 0
// Lane group g holds, per coordinate, the features 16 c + 4 g + 0..3 (bands 2g and 2g + 1; the
// second from the first by the double-angle identities) and 48 + 4 g + 0..3; every lane
// computes the same instructions (no branch on g).
// The hardware sine and cosine take revolutions (v_sin_f32(x) = sin(2 pi x)): band 2g's angle
// 2^(2g) pi u is u * 2^(2g - 1) revolutions, an exact scaling. The one-blob is
// exp2(-16 log2(e) (rough - centre)^2).
__device__ __forceinline__ void encode_frag(const NrcDev& N, float4 q0, float4 qg, int g, half8_t x[2]) {
    const float rev = (float)(1 << (2 * g)) * 0.5f;
    const float pos[3] = {q0.x, q0.y, q0.z};
    float pe[3][4];
    for (int c = 0; c < 3; c++) {
        const float u = (pos[c] - N.bmin[c]) * N.binv[c];
        const float a = u * rev;
        const float s0 = __builtin_amdgcn_sinf(a), c0 = __builtin_amdgcn_cosf(a);
        pe[c][0] = s0;
        pe[c][1] = c0;
        pe[c][2] = 2.f * s0 * c0;
        pe[c][3] = c0 * c0 - s0 * s0;
    }
    const float gq[4] = {qg.x, qg.y, qg.z, qg.w};
    // the lane group's four features as a bitwise select (as a conditional, the compiler
    // branches around the exponentials and serialises the tiles' loads behind them)
    const uint32_t mb = g == 1 ? ~0u : 0u, mq = (g == 0 || g == 2) ? ~0u : 0u;
    float o[4];
    for (int k = 0; k < 4; k++) {
        const float t = q0.w - ((float)k + 0.5f) * 0.25f;
        const float blob = __builtin_amdgcn_exp2f(t * t * (-16.f * 1.44269504f));
        o[k] = __uint_as_float((__float_as_uint(blob) & mb) | (__float_as_uint(gq[k]) & mq));
    }
    for (int i = 0; i < 4; i++) {
        x[0][i] = (_Float16)pe[0][i];
        x[0][4 + i] = (_Float16)pe[1][i];
        x[1][i] = (_Float16)pe[2][i];
        x[1][4 + i] = (_Float16)o[i];
    }
}

// The two float4 of packed query q that lane group g encodes (q0, and q2 for group 2, q1
// otherwise).
__device__ __forceinline__ void load_query(const float4* __restrict__ feats, int64_t q, int g, float4& q0,
                                           float4& qg) {
    q0 = feats[q * kNrcQ4];
    qg = feats[q * kNrcQ4 + (g == 2 ? 2 : 1)];
}

// Stages ROWS rows of a [ROWS][64] half matrix into LDS ([ROWS][kRowH]) with the columns in
// k-tile order, dst[row][32 t + 8 g + j] = src[row][kfeat(t, g, j)], by a block of NTHR threads.
template <int ROWS, int NTHR>
__device__ __forceinline__ void stage_perm(const _Float16* __restrict__ src, _Float16* dst) {
    // every load of the thread is issued before the first LDS store (one round trip, not one
    // per item)
    constexpr int kItems = ROWS * 8, kPer = (kItems + NTHR - 1) / NTHR;
    uint4 v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        const int u = (int)threadIdx.x + k * NTHR;
        if (kItems % NTHR == 0 || u < kItems) {
            const int row = u >> 3, t = (u >> 2) & 1, g = u & 3;
            const uint2 lo = *reinterpret_cast<const uint2*>(src + row * kNrcW + 32 * t + 4 * g);
            const uint2 hi = *reinterpret_cast<const uint2*>(src + row * kNrcW + 32 * t + 16 + 4 * g);
            v[k] = make_uint4(lo.x, lo.y, hi.x, hi.y);
        }
    }
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        const int u = (int)threadIdx.x + k * NTHR;
        if (kItems % NTHR == 0 || u < kItems) {
            const int row = u >> 3, t = (u >> 2) & 1, g = u & 3;
            *reinterpret_cast<uint4*>(dst + row * kRowH + 32 * t + 8 * g) = v[k];
        }
    }
}

// LDS image of the forward weights: hidden layers 0..3 (64 output rows each, W^T) and the
// output layer's first 16 rows; biases in fp32.
struct FwdLds {
    _Float16 w[(4 * kNrcW + 16) * kRowH];
    float b[kNrcLayers * kNrcW];
};
template <int NTHR>
__device__ __forceinline__ void stage_forward(const NrcDev& N, FwdLds& L) {
    // layers 0..3 (consecutive [64][64] blocks) and the output layer's first 16 rows
    stage_perm<4 * kNrcW + 16, NTHR>(N.wT, L.w);
    for (int i = threadIdx.x; i < kNrcLayers * kNrcW; i += NTHR) L.b[i] = N.b[i];
}

// ReLU of four fp32 accumulator rows as fp16: converted in pairs (round to nearest) and
// clamped at 0 in packed fp16. max(cvt(x), 0) = cvt(max(x, 0)) up to the sign of a zero,
// since the conversion is monotone and maps 0 to 0.
__device__ __forceinline__ void relu_f16(float4_t a, half8_t& x, int half) {
    const half2_t z = {(_Float16)0.f, (_Float16)0.f};
    const half2_t lo = __builtin_elementwise_max(half2_t{(_Float16)a[0], (_Float16)a[1]}, z);
    const half2_t hi = __builtin_elementwise_max(half2_t{(_Float16)a[2], (_Float16)a[3]}, z);
    x[4 * half + 0] = lo[0];
    x[4 * half + 1] = lo[1];
    x[4 * half + 2] = hi[0];
    x[4 * half + 3] = hi[1];
}

// One hidden layer for NT query tiles, in place: x = relu(W^T x + b) as the next layer's B
// fragments (the bias enters as the first MFMA's accumulator). Each weight fragment read
// from LDS feeds NT MFMAs.
template <int NT>
__device__ __forceinline__ void hidden_layer(const _Float16* W, const float* bias, half8_t x[2][NT]) {
    const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
    float4_t acc[4][NT];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const float4_t b4 = *reinterpret_cast<const float4_t*>(bias + 16 * m + 4 * g);
        const half8_t a = *reinterpret_cast<const half8_t*>(W + (16 * m + c) * kRowH + 8 * g);
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[m][nt] = mfma32(a, x[0][nt], b4);
    }
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const half8_t a = *reinterpret_cast<const half8_t*>(W + (16 * m + c) * kRowH + 32 + 8 * g);
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[m][nt] = mfma32(a, x[1][nt], acc[m][nt]);
    }
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) relu_f16(acc[m][nt], x[m >> 1][nt], m & 1);
}

// Output layer (rows 0..15 of W^T, 3 of them live): lane group 0 holds y[0..2] of its query.
template <int NT>
__device__ __forceinline__ void output_layer(const FwdLds& L, const half8_t in[2][NT], float4_t y[NT]) {
    const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
    const _Float16* W = L.w + 4 * kNrcW * kRowH;
    const float4_t b4 = *reinterpret_cast<const float4_t*>(L.b + 4 * kNrcW + 4 * g);
#pragma unroll
    for (int nt = 0; nt < NT; nt++) y[nt] = b4;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const half8_t a = *reinterpret_cast<const half8_t*>(W + c * kRowH + 32 * t + 8 * g);
#pragma unroll
        for (int nt = 0; nt < NT; nt++) y[nt] = mfma32(a, in[t][nt], y[nt]);
    }
}

// Inference over dense queries 0..n-1. recs == nullptr: kNrcQ4 float4 per query in feats,
// results in out[q] (float4). Otherwise the renderer's records in recs (kNrcRec4 float4: the
// query, then the path's throughput): the throughput's float4 becomes throughput x max(0,
// prediction), which k_accumulate adds to the path's radiance, so every access of the kernel
// is dense. (Adding it to the path's radiance here took a scattered read-modify-write per
// query; as no-return float atomics it measured 1.75x slower still.) A wavefront takes 4 tiles of
// 16 queries per iteration;
// four blocks of 40 KB LDS fit a CU, and the MFMA accumulators live in VGPRs (the file is
// built with -amdgpu-mfma-vgpr-form), so four waves per SIMD.
__global__ void __launch_bounds__(kNrcBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_nrc_infer(NrcDev N, const int32_t* __restrict__ countp, int32_t nfixed, const float4* feats, float4* recs,
            float4* __restrict__ out, unsigned long long* tot) {
    constexpr int NT = kInferTiles;
    __shared__ __align__(16) FwdLds S;
    stage_forward<kNrcBlock>(N, S);
    __syncthreads();
    const int n = countp ? min(*countp, nfixed) : nfixed;
    if (tot && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(tot, (unsigned long long)n);
    const int stride = recs ? kNrcRec4 : kNrcQ4;
    const float4* src = recs ? recs : feats;  // a record's query floats (slots 0..2) are only read
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 4, c = l & 15;
    for (int base = (blockIdx.x * 4 + wv) * 16 * NT; base < n; base += gridDim.x * 4 * 16 * NT) {  // wave-uniform
        half8_t x[2][NT];
        float4 q0[NT], qg[NT];  // every tile's loads are issued before the first encoding
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const int q = min(base + 16 * nt + c, n - 1);  // past the end: a copy of the last query
            q0[nt] = src[(int64_t)q * stride];
            qg[nt] = src[(int64_t)q * stride + (g == 2 ? 2 : 1)];
        }
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            half8_t e[2];
            encode_frag(N, q0[nt], qg[nt], g, e);
            x[0][nt] = e[0];
            x[1][nt] = e[1];
        }
#pragma unroll 1  // unrolled, the layers' LDS reads are hoisted and the registers run out
        for (int layer = 0; layer < 4; layer++)
            hidden_layer<NT>(S.w + layer * kNrcW * kRowH, S.b + layer * kNrcW, x);
        float4_t y[NT];
        output_layer<NT>(S, x, y);
        if (g == 0) {
#pragma unroll
            for (int nt = 0; nt < NT; nt++) {
                const int q = base + 16 * nt + c;
                if (q >= n) continue;
                const float r = fmaxf(y[nt][0], 0.f), gg = fmaxf(y[nt][1], 0.f), bb = fmaxf(y[nt][2], 0.f);
                if (recs) {
                    float4* tb = recs + (int64_t)q * kNrcRec4 + kNrcQ4;  // the throughput
                    const float4 b4 = *tb;
                    *tb = make_float4(b4.x * r, b4.y * gg, b4.z * bb, 0.f);
                } else {
                    out[q] = make_float4(r, gg, bb, 0.f);
                }
            }
        }
    }
}

// One training step's gradient in one kernel. A block of 16 wavefronts (four per SIMD)
// takes one tile of 256 rows (16 per wavefront); the grid covers the slice:
//   forward   every layer's input H0 (encoding) .. H4 stays in registers;
//   loss      dY = 2 (y - t) / (y^2 + 0.01) / 3 (clamped to +-1000, fp16) = D5;
//   backward  D_k = (W_k D_{k+1}) * (H_k > 0), k = 4..1;
//   gradient  for each layer L, H_L and D_{L+1} of the 256 rows go through LDS (feature-major,
//             [feature][row]); wavefront w computes the tile dW_L[16 (w % 4) + ..][16 (w / 4) + ..]
//             = H_L^T D_{L+1} over the tile's rows (K = 32 rows per MFMA) and, if w % 4 == 0, the
//             bias gradient of its 16 outputs, the column sums of D_{L+1}, as the MFMA of the
//             same D fragments with a ones matrix; both go to the block's partial at once.
// The block writes its sums as partial[layer][block]; k_nrc_reduce_adam adds the blocks.
// Rows past the slice enter as zeros. Backward weights W_k ([in][out]) are staged like the
// forward ones (K = out in k-tile order).
constexpr int kTrainWaves = 16;
constexpr int kTrainThreads = 64 * kTrainWaves;
constexpr int kTrainRows = 16 * kTrainWaves;  // rows per block
constexpr int kRowT = kTrainRows + 8;         // LDS row stride (halves) of the transposed tiles
struct TrainLds {
    FwdLds f;
    _Float16 wb[4 * kNrcW * kRowH];  // W_1 .. W_4, [in][out]
    _Float16 hs[kNrcW * kRowT];      // H_L of the block's rows, [feature][row]
    _Float16 ds[kNrcW * kRowT];      // D_{L+1}, [feature][row]
};

__global__ void __launch_bounds__(kTrainThreads) k_nrc_train(NrcDev N, const float4* __restrict__ feats,
                                                             const float4* __restrict__ targets,
                                                             const int32_t* __restrict__ countp, int32_t js,
                                                             int32_t ns, int32_t nfixed, float* __restrict__ partial,
                                                             int32_t maxBlocks, float* __restrict__ lossPart) {
    __shared__ __align__(16) TrainLds S;
    __shared__ float waveLoss[kTrainWaves];
    if ((int)blockIdx.x >= maxBlocks) return;  // the launch never has more blocks than partials
    int off;
    const int n = slice_rows(countp, js, ns, nfixed, &off);
    const int base = blockIdx.x * kTrainRows;
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 4, c = l & 15;
    const int it = wv & 3, mo = wv >> 2;  // the dW tile: input rows 16 it .., output columns 16 mo ..
    float* const part = partial + (int64_t)blockIdx.x * kPartial;  // + L * maxBlocks * kPartial
    if (base >= n) {  // block-uniform: no rows, zero partials
        if (threadIdx.x == 0) lossPart[blockIdx.x] = 0.f;
        for (int L = 0; L < kNrcLayers; L++) {
            if (L == 4 && mo > 0) continue;
            float* out = part + (int64_t)L * maxBlocks * kPartial;
            for (int i = 0; i < 4; i++) out[(16 * it + 4 * g + i) * kNrcW + 16 * mo + c] = 0.f;
            if (it == 0 && c == 0)
                for (int i = 0; i < 4; i++) out[kNrcW * kNrcW + 16 * mo + 4 * g + i] = 0.f;
        }
        return;
    }
    stage_forward<kTrainThreads>(N, S.f);
    stage_perm<4 * kNrcW, kTrainThreads>(N.w + kNrcW * kNrcW, S.wb);  // layers 1..4 of [layer][in][out]
    __syncthreads();
    feats += (int64_t)off * kNrcQ4;
    targets += off;
    const half8_t ones = {(_Float16)1.f, (_Float16)1.f, (_Float16)1.f, (_Float16)1.f,
                          (_Float16)1.f, (_Float16)1.f, (_Float16)1.f, (_Float16)1.f};
    const int q = base + 16 * wv + c;
    const bool valid = q < n;
    half8_t h[kNrcLayers][2][1];
    {
        float4 q0, qg;
        load_query(feats, min(q, n - 1), g, q0, qg);  // past the slice: masked below
        half8_t e[2];
        encode_frag(N, q0, qg, g, e);
        h[0][0][0] = e[0];
        h[0][1][0] = e[1];
    }
    for (int layer = 0; layer < 4; layer++) {
        h[layer + 1][0][0] = h[layer][0][0];
        h[layer + 1][1][0] = h[layer][1][0];
        hidden_layer<1>(S.f.w + layer * kNrcW * kRowH, S.f.b + layer * kNrcW, h[layer + 1]);
    }
    float4_t y[1];
    output_layer<1>(S.f, h[4], y);
    // D5 as the B fragment of the output layer's backward (k-tile 0; rows 0..2 are lane
    // group 0's slots 0..2, every other slot is 0)
    half8_t d[2][1];
    d[0][0] = half8_t{0, 0, 0, 0, 0, 0, 0, 0};
    d[1][0] = half8_t{0, 0, 0, 0, 0, 0, 0, 0};
    float lsum = 0.f;
    if (valid && g == 0) {
        const float4 tg = targets[q];
        const float tc[3] = {tg.x, tg.y, tg.z};
        for (int k = 0; k < 3; k++) {
            const float e = y[0][k] - tc[k];
            const float den = y[0][k] * y[0][k] + 0.01f;
            lsum += e * e / den;
            d[0][0][k] = (_Float16)fminf(fmaxf(2.f * e / den / 3.f, -1000.f), 1000.f);
        }
    }
    // the block's loss sum (wavefront sums through LDS, added in wavefront order after the
    // first barrier below); k_nrc_reduce_adam adds the blocks' sums: no atomics on one address
    for (int sh = 32; sh > 0; sh >>= 1) lsum += __shfl_xor(lsum, sh);
    if (l == 0) waveLoss[wv] = lsum;
    for (int L = 4; L >= 0; L--) {
        // H_L and D_{L+1} of the block's rows into LDS, [feature][row] (rows past the slice: 0)
        const int row = 16 * wv + c;
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int f = kfeat(t, g, j);
                S.hs[f * kRowT + row] = valid ? h[L][t][0][j] : (_Float16)0.f;
                S.ds[f * kRowT + row] = d[t][0][j];
            }
        __syncthreads();
        if (L == 4 && threadIdx.x == 0) {
            float bl = 0.f;
            for (int w = 0; w < kTrainWaves; w++) bl += waveLoss[w];
            lossPart[blockIdx.x] = bl;
        }
        if (L < 4 || mo == 0) {  // wavefront-uniform; the output layer has one column tile
            float4_t gw = {0.f, 0.f, 0.f, 0.f}, gb = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < kTrainRows / 32; ks++) {
                const half8_t A = *reinterpret_cast<const half8_t*>(&S.hs[(16 * it + c) * kRowT + 32 * ks + 8 * g]);
                // D_{L+1}[16 mo + c][rows]: the weight gradient's B fragment and, as an A
                // fragment, the bias gradient's rows
                const half8_t B = *reinterpret_cast<const half8_t*>(&S.ds[(16 * mo + c) * kRowT + 32 * ks + 8 * g]);
                gw = mfma32(A, B, gw);
                if (it == 0) gb = mfma32(B, ones, gb);  // wavefront-uniform
            }
            float* out = part + (int64_t)L * maxBlocks * kPartial;
            for (int i = 0; i < 4; i++) out[(16 * it + 4 * g + i) * kNrcW + 16 * mo + c] = gw[i];
            if (it == 0 && c == 0)
                for (int i = 0; i < 4; i++) out[kNrcW * kNrcW + 16 * mo + 4 * g + i] = gb[i];
        }
        if (L == 0) break;
        __syncthreads();  // the tiles are rewritten for the next layer
        // backward through layer L: D_L = (W_L D_{L+1}) * (H_L > 0)
        const _Float16* W = S.wb + (L - 1) * kNrcW * kRowH;
        float4_t acc[4];
        for (int m = 0; m < 4; m++) acc[m] = float4_t{0.f, 0.f, 0.f, 0.f};
        const int kt = L == 4 ? 1 : 2;  // D5 lives in k-tile 0 only
        for (int t = 0; t < kt; t++)
            for (int m = 0; m < 4; m++) {
                const half8_t a = *reinterpret_cast<const half8_t*>(W + (16 * m + c) * kRowH + 32 * t + 8 * g);
                acc[m] = mfma32(a, d[t][0], acc[m]);
            }
        for (int m = 0; m < 4; m++)
            for (int i = 0; i < 4; i++) {
                const int slot = 4 * (m & 1) + i;
                const bool on = (float)h[L][m >> 1][0][slot] > 0.f;
                d[m >> 1][0][slot] = (_Float16)(on ? acc[m][i] : 0.f);
            }
    }
}

// The gradient (mean over the slice's rows of the per-block partials) and, with adam, the
// Adam step on the FP32 master parameters with the fp16 copies refreshed in both layouts.
// A block takes 64 consecutive parameters (one layer's); its wavefront w adds the partials of
// blocks w, w + kReduceParts, ... in order (each load a contiguous 256 B row segment), and
// wavefront 0 adds the kReduceParts sums in order. The output layer's parameters beyond its
// first 16 columns have no partials: their gradient is 0. An empty slice takes no step.
constexpr int kReduceParts = 16;
__global__ void __launch_bounds__(64 * kReduceParts)
k_nrc_reduce_adam(const int32_t* __restrict__ countp, int32_t js, int32_t ns, int32_t nfixed,
                  const float* __restrict__ partial, int32_t maxBlocks, int32_t used_blocks, float* grad,
                  unsigned long long* tot, int32_t adam, float* master, float* m1, float* m2, _Float16* w,
                  _Float16* wT, float lr, float beta1, float beta2, float eps, float c1, float c2,
                  const float* __restrict__ lossPart, float* lossAcc) {
    __shared__ float red[kReduceParts][64];
    int off;
    const int n = slice_rows(countp, js, ns, nfixed, &off);
    if (tot && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(tot, (unsigned long long)n);
    const int lane = threadIdx.x & 63, part = threadIdx.x >> 6;
    if (lossAcc && blockIdx.x == 0 && part == 1) {  // the loss: the blocks' sums, in a fixed order
        float ls = 0.f;
        for (int k = lane; k < min(maxBlocks, used_blocks); k += 64) ls += lossPart[k];
        for (int sh = 32; sh > 0; sh >>= 1) ls += __shfl_xor(ls, sh);
        if (lane == 0) atomicAdd(lossAcc, ls);
    }
    const int i = blockIdx.x * 64 + lane;
    const bool inRange = i < kNrcParams;
    const int ii = inRange ? i : 0;
    const bool isW = ii < kNrcLayers * kNrcW * kNrcW;
    const int layer = isW ? ii / (kNrcW * kNrcW) : (ii - kNrcLayers * kNrcW * kNrcW) / kNrcW;
    const int e = isW ? ii % (kNrcW * kNrcW) : kNrcW * kNrcW + (ii - kNrcLayers * kNrcW * kNrcW) % kNrcW;
    const int out = isW ? ii % kNrcW : (ii - kNrcLayers * kNrcW * kNrcW) % kNrcW;
    const bool stored = inRange && (layer < 4 || out < 16);
    const int used = min(maxBlocks, used_blocks);
    float s = 0.f;
    if (stored) {
        const float* src = partial + (int64_t)layer * maxBlocks * kPartial + e;
#pragma unroll 8
        for (int k = part; k < used; k += kReduceParts) s += src[(int64_t)k * kPartial];
    }
    red[part][lane] = s;
    __syncthreads();
    if (part != 0 || !inRange) return;
    float sacc = 0.f;
#pragma unroll
    for (int k = 0; k < kReduceParts; k++) sacc += red[k][lane];
    const float gr = n > 0 ? sacc / (float)n : 0.f;
    grad[i] = gr;
    if (!adam || n == 0) return;
    const bool active = layer < 4 || out < kNrcOut;  // output layer: 3 of 64 padded columns
    float p = master[i];
    if (active) {
        const float a = beta1 * m1[i] + (1.f - beta1) * gr;
        const float b = beta2 * m2[i] + (1.f - beta2) * gr * gr;
        m1[i] = a;
        m2[i] = b;
        p -= lr * (a / c1) / (sqrtf(b / c2) + eps);
        master[i] = p;
    }
    if (isW) {
        const int in = (i % (kNrcW * kNrcW)) / kNrcW;
        w[i] = (_Float16)p;
        wT[layer * kNrcW * kNrcW + out * kNrcW + in] = (_Float16)p;
    }
}

// ------------------------------------------------------------------------------------ host

struct NrcArena {
    std::vector<void*> ptrs;
    template <class T>
    T* alloc(size_t n) {
        void* p = nullptr;
        const size_t b = std::max<size_t>(n * sizeof(T), 16);
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            (void)hipGetLastError();  // clear it: a later launch check must not report this failure
            fail(C2_ERR_NOMEM, "hipMalloc(" + std::to_string(b) + " bytes): " + hipGetErrorString(e));
        }
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    ~NrcArena() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

}  // namespace c2

using namespace c2;

struct c2_nrc {
    NrcArena A;
    c2_nrc_opts opts{};
    NrcDev dev{};
    float* master = nullptr;
    float* m1 = nullptr;
    float* m2 = nullptr;
    float* grad = nullptr;
    _Float16* w = nullptr;
    _Float16* wT = nullptr;
    int32_t cap = 0;          // training rows capacity
    float* partial = nullptr;
    int32_t maxBlocks = 0;  // training blocks for `cap` rows (one row tile and one partial each)
    float* loss = nullptr;       // accumulated loss sum (rows x 3 terms)
    float* loss_part = nullptr;  // [maxBlocks] the training blocks' loss sums of the last step
    unsigned long long* totals = nullptr;  // inference queries, training rows
    int64_t step = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double infer_ms = 0, train_ms = 0, infer_flop = 0, train_flop = 0;
    int64_t infer_queries = 0, train_rows = 0;
};

namespace c2 {

static void nrc_upload_params(c2_nrc* c, const std::vector<float>& params) {
    NRC_CHECK(hipMemcpy(c->master, params.data(), params.size() * sizeof(float), hipMemcpyHostToDevice));
    std::vector<_Float16> w(kNrcLayers * kNrcW * kNrcW), wT(w.size());
    for (int L = 0; L < kNrcLayers; L++)
        for (int i = 0; i < kNrcW; i++)
            for (int o = 0; o < kNrcW; o++) {
                float v = params[(size_t)L * kNrcW * kNrcW + i * kNrcW + o];
                w[(size_t)L * kNrcW * kNrcW + i * kNrcW + o] = (_Float16)v;
                wT[(size_t)L * kNrcW * kNrcW + o * kNrcW + i] = (_Float16)v;
            }
    NRC_CHECK(hipMemcpy(c->w, w.data(), w.size() * sizeof(_Float16), hipMemcpyHostToDevice));
    NRC_CHECK(hipMemcpy(c->wT, wT.data(), wT.size() * sizeof(_Float16), hipMemcpyHostToDevice));
}

// Blocks for `rows` rows at `per_block` rows per block iteration, at most `cap` blocks (each
// block stages the weights once and loops over its share).
static int nrc_grid(int64_t rows, int per_block, int cap) {
    int64_t blocks = (rows + per_block - 1) / per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(blocks, cap));
}
constexpr int kInferBlocks = 1024;  // 4 blocks of 40 KB LDS per CU

// Launches used by the renderer (c2_device.hip) on its stream.
void nrc_launch_infer(c2_nrc* c, const int32_t* countp, int32_t maxn, float4* recs, hipStream_t s) {
    hipLaunchKernelGGL(k_nrc_infer, dim3(nrc_grid(maxn, 4 * 16 * kInferTiles, kInferBlocks)), dim3(kNrcBlock), 0, s,
                       c->dev, countp, maxn, (const float4*)nullptr, recs, (float4*)nullptr, c->totals);
}

// One training step on slice js of ns of the *countp rows collected on the device (at most
// maxn rows per slice); a host-sized batch passes countp = nullptr, js = 0, ns = 1.
void nrc_launch_train(c2_nrc* c, const float4* feats, const float4* targets, const int32_t* countp, int32_t js,
                      int32_t ns, int32_t maxn, hipStream_t s, bool adam) {
    if (maxn > c->cap) fail(C2_ERR_INVALID, "radiance cache: training batch exceeds capacity");
    const int blocks = nrc_grid(maxn, kTrainRows, c->maxBlocks);  // maxn <= cap: one row tile per block
    hipLaunchKernelGGL(k_nrc_train, dim3(blocks), dim3(kTrainThreads), 0, s, c->dev, feats, targets, countp, js, ns, maxn,
                       c->partial, c->maxBlocks, c->loss_part);
    float c1 = 1.f, c2v = 1.f;
    if (adam) {
        c->step++;
        c1 = 1.f - (float)std::pow((double)c->opts.beta1, (double)c->step);
        c2v = 1.f - (float)std::pow((double)c->opts.beta2, (double)c->step);
    }
    const c2_nrc_opts& o = c->opts;
    hipLaunchKernelGGL(k_nrc_reduce_adam, dim3((kNrcParams + 63) / 64), dim3(64 * kReduceParts), 0, s, countp, js,
                       ns, maxn, c->partial, c->maxBlocks, blocks, c->grad, c->totals + 1, adam ? 1 : 0, c->master,
                       c->m1, c->m2, c->w, c->wT, o.learning_rate, o.beta1, o.beta2, o.epsilon, c1, c2v, c->loss_part,
                       c->loss);
    NRC_CHECK(hipGetLastError());
}

NrcDev nrc_dev(const c2_nrc* c) { return c->dev; }

// Queries from the host (kNrcFeat raw floats: position 0..2, outgoing direction 3..5, facing
// normal 6..8, roughness 9, albedo 10..12) in the packed layout the kernels read (c2_nrc.h),
// with the same float operations as the renderer's query writer.
static std::vector<float4> pack_queries(const float* feats, int64_t n) {
    std::vector<float4> q((size_t)n * kNrcQ4);
    for (int64_t i = 0; i < n; i++) {
        const float* r = feats + (size_t)i * kNrcFeat;
        float4* o = q.data() + (size_t)i * kNrcQ4;
        o[0] = make_float4(r[0], r[1], r[2], r[9]);
        o[1] = make_float4(std::atan2(r[5], r[3]) * 0.15915494f + 0.5f, r[4] * 0.5f + 0.5f,
                           std::atan2(r[8], r[6]) * 0.15915494f + 0.5f, r[7] * 0.5f + 0.5f);
        o[2] = make_float4(r[10], r[11], r[12], 1.f);
    }
    return q;
}

}  // namespace c2

extern "C" {

int c2_nrc_create(const c2_nrc_opts* opts, c2_nrc** out) {
    return guard([&] {
        if (!opts || !out) fail(C2_ERR_INVALID, "null argument");
        if (opts->batch_capacity <= 0) fail(C2_ERR_INVALID, "batch_capacity must be positive");
        std::unique_ptr<c2_nrc> c(new c2_nrc());
        c->opts = *opts;
        c->master = c->A.alloc<float>(kNrcParams);
        c->m1 = c->A.alloc<float>(kNrcParams);
        c->m2 = c->A.alloc<float>(kNrcParams);
        c->grad = c->A.alloc<float>(kNrcParams);
        c->w = c->A.alloc<_Float16>(kNrcLayers * kNrcW * kNrcW);
        c->wT = c->A.alloc<_Float16>(kNrcLayers * kNrcW * kNrcW);
        c->cap = opts->batch_capacity;
        c->maxBlocks = (int32_t)((c->cap + kTrainRows - 1) / kTrainRows);
        c->partial = c->A.alloc<float>((size_t)kNrcLayers * c->maxBlocks * kPartial);
        c->loss_part = c->A.alloc<float>(c->maxBlocks);
        c->loss = c->A.alloc<float>(1);
        c->totals = c->A.alloc<unsigned long long>(2);
        NRC_CHECK(hipMemset(c->totals, 0, 2 * sizeof(unsigned long long)));
        NRC_CHECK(hipMemset(c->m1, 0, kNrcParams * sizeof(float)));
        NRC_CHECK(hipMemset(c->m2, 0, kNrcParams * sizeof(float)));
        NRC_CHECK(hipMemset(c->loss, 0, sizeof(float)));
        // He-uniform initialisation (deterministic), output layer: 3 active columns
        std::mt19937 rng((uint32_t)opts->seed);
        std::vector<float> params(kNrcParams, 0.f);
        for (int L = 0; L < kNrcLayers; L++) {
            float lim = std::sqrt(6.f / kNrcW);
            if (L == kNrcLayers - 1) lim = std::sqrt(6.f / kNrcW) * 0.1f;
            std::uniform_real_distribution<float> U(-lim, lim);
            for (int i = 0; i < kNrcW; i++)
                for (int o = 0; o < kNrcW; o++)
                    if (L < kNrcLayers - 1 || o < kNrcOut) params[(size_t)L * kNrcW * kNrcW + i * kNrcW + o] = U(rng);
        }
        c->dev.w = c->w;
        c->dev.wT = c->wT;
        c->dev.b = c->master + kNrcLayers * kNrcW * kNrcW;
        for (int k = 0; k < 3; k++) {
            c->dev.bmin[k] = opts->bbox_min[k];
            float e = opts->bbox_max[k] - opts->bbox_min[k];
            c->dev.binv[k] = e > 0 ? 1.f / e : 1.f;
        }
        nrc_upload_params(c.get(), params);
        *out = c.release();
    });
}

void c2_nrc_destroy(c2_nrc* c) { delete c; }

int c2_nrc_get_params(c2_nrc* c, float* params) {
    return guard([&] {
        if (!c || !params) fail(C2_ERR_INVALID, "null argument");
        NRC_CHECK(hipDeviceSynchronize());
        NRC_CHECK(hipMemcpy(params, c->master, kNrcParams * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int c2_nrc_set_params(c2_nrc* c, const float* params) {
    return guard([&] {
        if (!c || !params) fail(C2_ERR_INVALID, "null argument");
        std::vector<float> p(params, params + kNrcParams);
        nrc_upload_params(c, p);
    });
}

int c2_nrc_infer(c2_nrc* c, const float* feats, int64_t n, float* out) {
    return guard([&] {
        if (!c || n < 0 || (n > 0 && (!feats || !out))) fail(C2_ERR_INVALID, "bad arguments");
        if (n == 0) return;
        NrcArena T;
        float4* df = T.alloc<float4>((size_t)n * kNrcQ4);
        float4* dout = T.alloc<float4>(n);
        const std::vector<float4> q = pack_queries(feats, n);
        NRC_CHECK(hipMemcpy(df, q.data(), q.size() * sizeof(float4), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_nrc_infer, dim3(nrc_grid(n, 4 * 16 * kInferTiles, kInferBlocks)), dim3(kNrcBlock), 0, 0,
                           c->dev, (const int32_t*)nullptr, (int32_t)n, df, (float4*)nullptr, dout,
                           (unsigned long long*)nullptr);
        NRC_CHECK(hipGetLastError());
        std::vector<float4> o(n);
        NRC_CHECK(hipMemcpy(o.data(), dout, n * sizeof(float4), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; i++) {
            out[i * 3] = o[i].x;
            out[i * 3 + 1] = o[i].y;
            out[i * 3 + 2] = o[i].z;
        }
    });
}

int c2_nrc_train_step(c2_nrc* c, const float* feats, const float* targets, int64_t n, int32_t apply, float* grad_out,
                      float* loss_out) {
    return guard([&] {
        if (!c || n <= 0 || !feats || !targets) fail(C2_ERR_INVALID, "bad arguments");
        if (n > c->cap) fail(C2_ERR_INVALID, "batch exceeds capacity");
        NrcArena T;
        float4* df = T.alloc<float4>((size_t)n * kNrcQ4);
        float4* dt = T.alloc<float4>(n);
        const std::vector<float4> q = pack_queries(feats, n);
        NRC_CHECK(hipMemcpy(df, q.data(), q.size() * sizeof(float4), hipMemcpyHostToDevice));
        std::vector<float4> t4(n);
        for (int64_t i = 0; i < n; i++) t4[i] = make_float4(targets[i * 3], targets[i * 3 + 1], targets[i * 3 + 2], 0);
        NRC_CHECK(hipMemcpy(dt, t4.data(), n * sizeof(float4), hipMemcpyHostToDevice));
        NRC_CHECK(hipMemset(c->loss, 0, sizeof(float)));
        nrc_launch_train(c, df, dt, nullptr, 0, 1, (int32_t)n, nullptr, apply != 0);
        NRC_CHECK(hipDeviceSynchronize());
        if (grad_out) NRC_CHECK(hipMemcpy(grad_out, c->grad, kNrcParams * sizeof(float), hipMemcpyDeviceToHost));
        if (loss_out) {
            float l = 0;
            NRC_CHECK(hipMemcpy(&l, c->loss, sizeof(float), hipMemcpyDeviceToHost));
            *loss_out = l / (float)(3 * n);
        }
    });
}

int c2_nrc_get_stats(const c2_nrc* c, c2_nrc_stats* out) {
    return guard([&] {
        if (!c || !out) fail(C2_ERR_INVALID, "null argument");
        unsigned long long t[2] = {0, 0};
        NRC_CHECK(hipMemcpy(t, c->totals, sizeof(t), hipMemcpyDeviceToHost));
        out->steps = c->step;
        out->flop_forward_per_query = kNrcFlopFwd;
        out->queries = (int64_t)t[0];
        out->train_rows = (int64_t)t[1];
    });
}

}  // extern "C"
