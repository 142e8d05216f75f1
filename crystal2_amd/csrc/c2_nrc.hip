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

// B fragments (both k-tiles) of the 64-dim input encoding of raw query r[16] for lane group
// g. Features (the order of the weight rows of layer 0):
//   16 c + 2 b + {0, 1}: sin / cos (2^b pi u_c), u_c = the position normalised to [0, 1], b < 8
//   48 / 50: azimuth of the outgoing direction / normal, atan2(z, x) / 2 pi + 0.5 (r[13], r[14],
//            filled by whoever writes the query), 49 / 51: their cos theta * 0.5 + 0.5
//   52 + k: one-blob roughness, exp(-((r[9] - (k + 0.5) / 4) * 4)^2), k < 4
//   56..58: albedo, 59: 1, 60..63: 0
// Lane group g holds, per coordinate, the features 16 c + 4 g + 0..3 (bands 2g and 2g + 1; the
// second from the first by the double-angle identities) and 48 + 4 g + 0..3; every lane
// computes the same instructions (no branch on g).
__device__ __forceinline__ void encode_frag(const NrcDev& N, const float* r, int g, half8_t x[2]) {
    const float scale = (float)(1 << (2 * g)) * 3.14159265f;
    float pe[3][4];
    for (int c = 0; c < 3; c++) {
        const float u = (r[c] - N.bmin[c]) * N.binv[c];
        const float a = u * scale;
        const float s0 = __sinf(a), c0 = __cosf(a);
        pe[c][0] = s0;
        pe[c][1] = c0;
        pe[c][2] = 2.f * s0 * c0;
        pe[c][3] = c0 * c0 - s0 * s0;
    }
    const float g0[4] = {r[13], r[4] * 0.5f + 0.5f, r[14], r[7] * 0.5f + 0.5f};
    const float g2[4] = {r[10], r[11], r[12], 1.f};
    float o[4];
    for (int k = 0; k < 4; k++) {
        const float d = (r[9] - ((float)k + 0.5f) * 0.25f) * 4.f;
        const float blob = __expf(-d * d);
        o[k] = g == 0 ? g0[k] : g == 1 ? blob : g == 2 ? g2[k] : 0.f;
    }
    for (int i = 0; i < 4; i++) {
        x[0][i] = (_Float16)pe[0][i];
        x[0][4 + i] = (_Float16)pe[1][i];
        x[1][i] = (_Float16)pe[2][i];
        x[1][4 + i] = (_Float16)o[i];
    }
}

__device__ __forceinline__ void load_raw(const float4* __restrict__ feats, int64_t p, bool valid, float raw[kNrcFeat]) {
    for (int k = 0; k < 4; k++) {
        float4 v = valid ? feats[p * 4 + k] : make_float4(0.f, 0.f, 0.f, 0.f);
        raw[4 * k] = v.x;
        raw[4 * k + 1] = v.y;
        raw[4 * k + 2] = v.z;
        raw[4 * k + 3] = v.w;
    }
}

// Stages `rows` rows of a [rows][64] half matrix into LDS ([rows][kRowH]) with the columns in
// k-tile order: dst[row][32 t + 8 g + j] = src[row][kfeat(t, g, j)].
__device__ __forceinline__ void stage_perm(const _Float16* __restrict__ src, _Float16* dst, int rows) {
    for (int u = threadIdx.x; u < rows * 8; u += blockDim.x) {
        const int row = u >> 3, t = (u >> 2) & 1, g = u & 3;
        const uint2 lo = *reinterpret_cast<const uint2*>(src + row * kNrcW + 32 * t + 4 * g);
        const uint2 hi = *reinterpret_cast<const uint2*>(src + row * kNrcW + 32 * t + 16 + 4 * g);
        *reinterpret_cast<uint4*>(dst + row * kRowH + 32 * t + 8 * g) = make_uint4(lo.x, lo.y, hi.x, hi.y);
    }
}

// LDS image of the forward weights: hidden layers 0..3 (64 output rows each, W^T) and the
// output layer's first 16 rows; biases in fp32.
struct FwdLds {
    _Float16 w[(4 * kNrcW + 16) * kRowH];
    float b[kNrcLayers * kNrcW];
};
__device__ __forceinline__ void stage_forward(const NrcDev& N, FwdLds& L) {
    stage_perm(N.wT, L.w, 4 * kNrcW);  // layers 0..3 are consecutive [64][64] blocks
    stage_perm(N.wT + 4 * kNrcW * kNrcW, L.w + 4 * kNrcW * kRowH, 16);
    for (int i = threadIdx.x; i < kNrcLayers * kNrcW; i += blockDim.x) L.b[i] = N.b[i];
}

// One hidden layer for NT query tiles, in place: x = relu(W^T x + b) as the next layer's B
// fragments (the bias enters as the first MFMA's accumulator).
template <int NT>
__device__ __forceinline__ void hidden_layer(const _Float16* W, const float* bias, half8_t x[2][NT]) {
    const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
    float4_t acc[4][NT];
    for (int m = 0; m < 4; m++) {
        const float4_t b4 = *reinterpret_cast<const float4_t*>(bias + 16 * m + 4 * g);
        const half8_t a = *reinterpret_cast<const half8_t*>(W + (16 * m + c) * kRowH + 8 * g);
        for (int nt = 0; nt < NT; nt++) acc[m][nt] = mfma32(a, x[0][nt], b4);
    }
    for (int m = 0; m < 4; m++) {
        const half8_t a = *reinterpret_cast<const half8_t*>(W + (16 * m + c) * kRowH + 32 + 8 * g);
        for (int nt = 0; nt < NT; nt++) acc[m][nt] = mfma32(a, x[1][nt], acc[m][nt]);
    }
    for (int m = 0; m < 4; m++)
        for (int nt = 0; nt < NT; nt++)
            for (int i = 0; i < 4; i++) x[m >> 1][nt][4 * (m & 1) + i] = (_Float16)fmaxf(acc[m][nt][i], 0.f);
}

// Output layer (rows 0..15 of W^T, 3 of them live): lane group 0 holds y[0..2] of its query.
template <int NT>
__device__ __forceinline__ void output_layer(const FwdLds& L, const half8_t in[2][NT], float4_t y[NT]) {
    const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
    const _Float16* W = L.w + 4 * kNrcW * kRowH;
    const float4_t b4 = *reinterpret_cast<const float4_t*>(L.b + 4 * kNrcW + 4 * g);
    for (int nt = 0; nt < NT; nt++) y[nt] = b4;
    for (int t = 0; t < 2; t++) {
        const half8_t a = *reinterpret_cast<const half8_t*>(W + c * kRowH + 32 * t + 8 * g);
        for (int nt = 0; nt < NT; nt++) y[nt] = mfma32(a, in[t][nt], y[nt]);
    }
}

// Inference. list == nullptr: queries 0..n-1 with results in out[q] (float4).
// Otherwise query i is path list[i]: L[p] += beta[p] * max(0, prediction).
__global__ void __launch_bounds__(kNrcBlock) k_nrc_infer(NrcDev N, const int32_t* __restrict__ list,
                                                         const int32_t* __restrict__ countp, int32_t nfixed,
                                                         const float4* __restrict__ feats, float4* __restrict__ L,
                                                         const float4* __restrict__ beta, float4* __restrict__ out,
                                                         unsigned long long* tot) {
    constexpr int NT = kInferTiles;
    __shared__ __align__(16) FwdLds S;
    stage_forward(N, S);
    __syncthreads();
    const int n = countp ? min(*countp, nfixed) : nfixed;
    if (tot && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(tot, (unsigned long long)n);
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 4, c = l & 15;
    for (int base = (blockIdx.x * 4 + wv) * 16 * NT; base < n; base += gridDim.x * 4 * 16 * NT) {
        half8_t x[2][NT];
        int pq[NT];
        for (int nt = 0; nt < NT; nt++) {
            const int q = base + 16 * nt + c;
            const bool valid = q < n;
            pq[nt] = valid ? (list ? list[q] : q) : -1;
            float raw[kNrcFeat];
            load_raw(feats, valid ? pq[nt] : 0, valid, raw);
            half8_t e[2];
            encode_frag(N, raw, g, e);
            x[0][nt] = e[0];
            x[1][nt] = e[1];
        }
        for (int layer = 0; layer < 4; layer++) hidden_layer<NT>(S.w + layer * kNrcW * kRowH, S.b + layer * kNrcW, x);
        float4_t y[NT];
        output_layer<NT>(S, x, y);
        if (g == 0)
            for (int nt = 0; nt < NT; nt++) {
                const int p = pq[nt];
                if (p < 0) continue;
                const float r = fmaxf(y[nt][0], 0.f), gg = fmaxf(y[nt][1], 0.f), bb = fmaxf(y[nt][2], 0.f);
                if (list) {
                    float4 b4 = beta[p], L4 = L[p];
                    L[p] = make_float4(L4.x + b4.x * r, L4.y + b4.y * gg, L4.z + b4.z * bb, L4.w);
                } else {
                    out[p] = make_float4(r, gg, bb, 0.f);
                }
            }
    }
}

// One training step's gradient in one kernel. A block of 4 wavefronts takes 64 rows at a
// time (16 per wavefront) and loops over its share of the slice:
//   forward   every layer's input H0 (encoding) .. H4 stays in registers;
//   loss      dY = 2 (y - t) / (y^2 + 0.01) / 3 (clamped to +-1000, fp16) = D5;
//   backward  D_k = (W_k D_{k+1}) * (H_k > 0), k = 4..1;
//   gradient  for each layer L, H_L and D_{L+1} of the 64 rows go through LDS (feature-major,
//             [feature][row]) and wavefront w accumulates dW_L[16 w .. 16 w + 15][all outputs]
//             = H_L^T D_{L+1} (K = 32 rows per MFMA) and the bias gradient (column sums of
//             D_{L+1}) for outputs 16 w .. 16 w + 15, in registers across the block's rows.
// The block writes its sums as partial[layer][block]; k_nrc_reduce_adam adds the blocks.
// Rows past the slice enter as zeros. Backward weights W_k ([in][out]) are staged like the
// forward ones (K = out in k-tile order).
constexpr int kTrainRows = 64;   // rows per block iteration (4 wavefronts x 16)
constexpr int kRowT = 72;        // LDS row stride (halves) of the transposed tiles: 64 rows + 8
struct TrainLds {
    FwdLds f;
    _Float16 wb[4 * kNrcW * kRowH];  // W_1 .. W_4, [in][out]
    _Float16 hs[kNrcW * kRowT];      // H_L of the block's 64 rows, [feature][row]
    _Float16 ds[kNrcW * kRowT];      // D_{L+1}, [feature][row]
};

__global__ void __launch_bounds__(kNrcBlock) k_nrc_train(NrcDev N, const float4* __restrict__ feats,
                                                         const float4* __restrict__ targets,
                                                         const int32_t* __restrict__ countp, int32_t js, int32_t ns,
                                                         int32_t nfixed, float* __restrict__ partial,
                                                         int32_t maxBlocks, float* lossAcc) {
    __shared__ __align__(16) TrainLds S;
    stage_forward(N, S.f);
    stage_perm(N.w + kNrcW * kNrcW, S.wb, 4 * kNrcW);  // layers 1..4 of [layer][in][out]
    __syncthreads();
    int off;
    const int n = slice_rows(countp, js, ns, nfixed, &off);
    feats += (int64_t)off * 4;  // 4 float4 per row
    targets += off;
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 4, c = l & 15;
    float lsum = 0.f;
    float4_t gw[kNrcLayers][4];  // dW_L[16 wv + 4 g + i][16 mo + c]
    float gb[kNrcLayers];        // db_L[16 wv + c] (lane group 0)
    for (int L = 0; L < kNrcLayers; L++) {
        gb[L] = 0.f;
        for (int mo = 0; mo < 4; mo++) gw[L][mo] = float4_t{0.f, 0.f, 0.f, 0.f};
    }
    for (int base = blockIdx.x * kTrainRows; base < n; base += gridDim.x * kTrainRows) {  // block-uniform
        const int q = base + 16 * wv + c;
        const bool valid = q < n;
        half8_t h[kNrcLayers][2][1];
        {
            float raw[kNrcFeat];
            load_raw(feats, valid ? q : 0, valid, raw);
            half8_t e[2];
            encode_frag(N, raw, g, e);
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
        for (int L = 4; L >= 0; L--) {
            // H_L and D_{L+1} of the 64 rows into LDS, [feature][row] (rows past the slice: 0)
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
            const int mo_tiles = L == 4 ? 1 : 4;
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                const half8_t A = *reinterpret_cast<const half8_t*>(&S.hs[(16 * wv + c) * kRowT + 32 * ks + 8 * g]);
#pragma unroll
                for (int mo = 0; mo < 4; mo++) {
                    if (mo >= mo_tiles) continue;  // block-uniform
                    const half8_t B = *reinterpret_cast<const half8_t*>(&S.ds[(16 * mo + c) * kRowT + 32 * ks + 8 * g]);
                    gw[L][mo] = mfma32(A, B, gw[L][mo]);
                }
            }
            {  // bias gradient: output 16 wv + c, rows 16 g .. 16 g + 15, then across lane groups
                const half8_t b0 = *reinterpret_cast<const half8_t*>(&S.ds[(16 * wv + c) * kRowT + 16 * g]);
                const half8_t b1 = *reinterpret_cast<const half8_t*>(&S.ds[(16 * wv + c) * kRowT + 16 * g + 8]);
                float sb = 0.f;
#pragma unroll
                for (int j = 0; j < 8; j++) sb += (float)b0[j] + (float)b1[j];
                sb += __shfl_xor(sb, 16);
                sb += __shfl_xor(sb, 32);
                gb[L] += sb;
            }
            __syncthreads();  // the tiles are rewritten for the next layer
            if (L == 0) break;
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
    if (lossAcc) {
        for (int sh = 32; sh > 0; sh >>= 1) lsum += __shfl_xor(lsum, sh);
        if (l == 0) atomicAdd(lossAcc, lsum);
    }
    if ((int)blockIdx.x >= maxBlocks) return;
    for (int L = 0; L < kNrcLayers; L++) {
        float* out = partial + ((int64_t)L * maxBlocks + blockIdx.x) * kPartial;
        const int mo_tiles = L == 4 ? 1 : 4;
        for (int mo = 0; mo < 4; mo++)
            for (int i = 0; i < 4; i++)
                out[(16 * wv + 4 * g + i) * kNrcW + 16 * mo + c] = mo < mo_tiles ? gw[L][mo][i] : 0.f;
        if (g == 0) out[kNrcW * kNrcW + 16 * wv + c] = (L < 4 || wv == 0) ? gb[L] : 0.f;
    }
}

// The gradient (mean over the slice's rows of the per-block partials, added in block order
// by 8 lanes per parameter, then across them) and, with adam, the Adam step on the FP32
// master parameters with the fp16 copies refreshed in both layouts. An empty slice takes
// no step.
constexpr int kReduceLanes = 8;
__global__ void __launch_bounds__(256) k_nrc_reduce_adam(const int32_t* __restrict__ countp, int32_t js, int32_t ns,
                                                         int32_t nfixed, const float* __restrict__ partial,
                                                         int32_t maxBlocks, int32_t used_blocks, float* grad,
                                                         unsigned long long* tot, int32_t adam, float* master,
                                                         float* m1, float* m2, _Float16* w, _Float16* wT, float lr,
                                                         float beta1, float beta2, float eps, float c1, float c2) {
    int off;
    const int n = slice_rows(countp, js, ns, nfixed, &off);
    if (tot && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(tot, (unsigned long long)n);
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gt / kReduceLanes, part = gt % kReduceLanes;
    const bool inRange = i < kNrcParams;
    const int ii = inRange ? i : 0;
    const bool isW = ii < kNrcLayers * kNrcW * kNrcW;
    const int layer = isW ? ii / (kNrcW * kNrcW) : (ii - kNrcLayers * kNrcW * kNrcW) / kNrcW;
    const int e = isW ? ii % (kNrcW * kNrcW) : kNrcW * kNrcW + (ii - kNrcLayers * kNrcW * kNrcW) % kNrcW;
    const int used = min(maxBlocks, used_blocks);
    float sacc = 0.f;
    for (int k = part; k < used; k += kReduceLanes) sacc += partial[((int64_t)layer * maxBlocks + k) * kPartial + e];
    for (int sh = kReduceLanes / 2; sh > 0; sh >>= 1) sacc += __shfl_xor(sacc, sh);
    if (!inRange || part != 0) return;
    const float gr = n > 0 ? sacc / (float)n : 0.f;
    grad[i] = gr;
    if (!adam || n == 0) return;
    const int out = isW ? i % kNrcW : (i - kNrcLayers * kNrcW * kNrcW) % kNrcW;
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
    int32_t maxBlocks = 0;  // training blocks (one partial each)
    float* loss = nullptr;
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
constexpr int kInferBlocks = 768;  // 3 blocks of 40 KB LDS per CU
constexpr int kTrainBlocks = 256;   // one block of 96 KB LDS per CU, each looping over row tiles

// Launches used by the renderer (c2_device.hip) on its stream.
void nrc_launch_infer(c2_nrc* c, const int32_t* list, const int32_t* countp, int32_t maxn, const float4* feats,
                      float4* L, const float4* beta, hipStream_t s) {
    hipLaunchKernelGGL(k_nrc_infer, dim3(nrc_grid(maxn, 4 * 16 * kInferTiles, kInferBlocks)), dim3(kNrcBlock), 0, s,
                       c->dev, list, countp, maxn, feats, L, beta, (float4*)nullptr, c->totals);
}

// One training step on slice js of ns of the *countp rows collected on the device (at most
// maxn rows per slice); a host-sized batch passes countp = nullptr, js = 0, ns = 1.
void nrc_launch_train(c2_nrc* c, const float4* feats, const float4* targets, const int32_t* countp, int32_t js,
                      int32_t ns, int32_t maxn, hipStream_t s, bool adam) {
    if (maxn > c->cap) fail(C2_ERR_INVALID, "radiance cache: training batch exceeds capacity");
    const int blocks = nrc_grid(maxn, kTrainRows, kTrainBlocks);
    hipLaunchKernelGGL(k_nrc_train, dim3(blocks), dim3(kNrcBlock), 0, s, c->dev, feats, targets, countp, js, ns, maxn,
                       c->partial, c->maxBlocks, c->loss);
    float c1 = 1.f, c2v = 1.f;
    if (adam) {
        c->step++;
        c1 = 1.f - (float)std::pow((double)c->opts.beta1, (double)c->step);
        c2v = 1.f - (float)std::pow((double)c->opts.beta2, (double)c->step);
    }
    const c2_nrc_opts& o = c->opts;
    hipLaunchKernelGGL(k_nrc_reduce_adam, dim3((kNrcParams * kReduceLanes + 255) / 256), dim3(256), 0, s, countp, js,
                       ns, maxn, c->partial, c->maxBlocks, blocks, c->grad, c->totals + 1, adam ? 1 : 0, c->master,
                       c->m1, c->m2, c->w, c->wT, o.learning_rate, o.beta1, o.beta2, o.epsilon, c1, c2v);
    NRC_CHECK(hipGetLastError());
}

NrcDev nrc_dev(const c2_nrc* c) { return c->dev; }

// Queries from the host: slots 13 and 14 get the azimuths of the outgoing direction (slots
// 3..5) and the normal (6..8), atan2(z, x) / 2 pi + 0.5, as the renderer writes them.
static std::vector<float> with_azimuths(const float* feats, int64_t n) {
    std::vector<float> q(feats, feats + (size_t)n * kNrcFeat);
    for (int64_t i = 0; i < n; i++) {
        float* r = q.data() + (size_t)i * kNrcFeat;
        r[13] = std::atan2(r[5], r[3]) * 0.15915494f + 0.5f;
        r[14] = std::atan2(r[8], r[6]) * 0.15915494f + 0.5f;
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
        c->maxBlocks = kTrainBlocks;
        c->partial = c->A.alloc<float>((size_t)kNrcLayers * c->maxBlocks * kPartial);
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
        float4* df = T.alloc<float4>((size_t)n * 4);
        float4* dout = T.alloc<float4>(n);
        const std::vector<float> q = with_azimuths(feats, n);
        NRC_CHECK(hipMemcpy(df, q.data(), (size_t)n * kNrcFeat * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_nrc_infer, dim3(nrc_grid(n, 4 * 16 * kInferTiles, kInferBlocks)), dim3(kNrcBlock), 0, 0,
                           c->dev, (const int32_t*)nullptr,
                           (const int32_t*)nullptr, (int32_t)n, df, (float4*)nullptr, (const float4*)nullptr, dout,
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
        float4* df = T.alloc<float4>((size_t)n * 4);
        float4* dt = T.alloc<float4>(n);
        const std::vector<float> q = with_azimuths(feats, n);
        NRC_CHECK(hipMemcpy(df, q.data(), (size_t)n * kNrcFeat * sizeof(float), hipMemcpyHostToDevice));
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
