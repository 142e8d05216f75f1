// Radiance-cache MLP on gfx950: MFMA (v_mfma_f32_16x16x16f16) inference, training
// forward/backward, split-K weight gradients and FP32 Adam. See c2_nrc.h for the
// network, the query format and the fragment layouts.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "c2_host.h"
#include "c2_nrc.h"

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

namespace c2 {

#define NRC_CHECK(x)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) fail(C2_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));  \
    } while (0)

constexpr int kNrcBlock = 256;         // 4 wavefronts
constexpr int kRowsPerWave = 256;      // training rows per wavefront in the weight-gradient pass
constexpr int kPartial = kNrcW * kNrcW + kNrcW;  // dW + db of one layer

__device__ __forceinline__ float4_t mfma16(half4_t a, half4_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

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

// Input encoding: feature f in [0, 64) of raw query r[16].
__device__ __forceinline__ float nrc_feature(const float* r, const NrcDev& N, int f) {
    if (f < 48) {
        int c = f >> 4, w = f & 15, band = w >> 1;
        float u = (r[c] - N.bmin[c]) * N.binv[c];
        float x = u * (float)(1 << band) * 3.14159265f;
        return (w & 1) ? __cosf(x) : __sinf(x);
    }
    int k = f - 48;
    if (k == 0) return atan2f(r[5], r[3]) * 0.15915494f + 0.5f;  // outgoing direction phi
    if (k == 1) return r[4] * 0.5f + 0.5f;                         // outgoing direction cos(theta)
    if (k == 2) return atan2f(r[8], r[6]) * 0.15915494f + 0.5f;  // normal phi
    if (k == 3) return r[7] * 0.5f + 0.5f;
    if (k < 8) {  // one-blob roughness, 4 bins
        float c = ((k - 4) + 0.5f) * 0.25f;
        float d = (r[9] - c) * 4.f;
        return __expf(-d * d);
    }
    if (k < 11) return r[10 + (k - 8)];  // albedo
    if (k == 11) return 1.f;
    return 0.f;
}

__device__ __forceinline__ half4_t to_half4(float4_t v) {
    half4_t h;
    h[0] = (_Float16)v[0];
    h[1] = (_Float16)v[1];
    h[2] = (_Float16)v[2];
    h[3] = (_Float16)v[3];
    return h;
}

// Stage all forward (or backward) weights into LDS: 5 x 64 x 64 halves + 5 x 64 biases.
__device__ __forceinline__ void stage_weights(const _Float16* __restrict__ src, const float* __restrict__ bsrc,
                                              _Float16* sW, float* sB) {
    const uint2* s2 = reinterpret_cast<const uint2*>(src);
    uint2* d2 = reinterpret_cast<uint2*>(sW);
    for (int i = threadIdx.x; i < kNrcLayers * kNrcW * kNrcW / 4; i += blockDim.x) d2[i] = s2[i];
    for (int i = threadIdx.x; i < kNrcLayers * kNrcW; i += blockDim.x) sB[i] = bsrc[i];
    __syncthreads();
}

// Forward pass of one 16-query block held by a wavefront; act[t] holds the input
// encoding on entry. If H != nullptr the encoding and every hidden activation are
// stored (fp16, [row][64]) for training. Returns the output D fragment.
__device__ __forceinline__ float4_t nrc_forward(const _Float16* sW, const float* sB, half4_t act[4], int q, bool valid,
                                                _Float16* H, int64_t rowStride) {
    const int l = threadIdx.x & 63, g = l >> 4, r16 = l & 15;
    for (int layer = 0; layer < 4; layer++) {
        if (H && valid) {
            _Float16* hrow = H + (int64_t)layer * rowStride + (int64_t)q * kNrcW;
            for (int t = 0; t < 4; t++) *reinterpret_cast<half4_t*>(hrow + 16 * t + 4 * g) = act[t];
        }
        const _Float16* Wl = sW + layer * kNrcW * kNrcW;
        half4_t nact[4];
        for (int m = 0; m < 4; m++) {
            float4_t acc = {0.f, 0.f, 0.f, 0.f};
            for (int t = 0; t < 4; t++) {
                half4_t a = *reinterpret_cast<const half4_t*>(Wl + (16 * m + r16) * kNrcW + 16 * t + 4 * g);
                acc = mfma16(a, act[t], acc);
            }
            for (int j = 0; j < 4; j++) acc[j] = fmaxf(acc[j] + sB[layer * kNrcW + 16 * m + 4 * g + j], 0.f);
            nact[m] = to_half4(acc);
        }
        for (int t = 0; t < 4; t++) act[t] = nact[t];
    }
    if (H && valid) {
        _Float16* hrow = H + 4 * rowStride + (int64_t)q * kNrcW;
        for (int t = 0; t < 4; t++) *reinterpret_cast<half4_t*>(hrow + 16 * t + 4 * g) = act[t];
    }
    const _Float16* Wo = sW + 4 * kNrcW * kNrcW;
    float4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 4; t++) {
        half4_t a = *reinterpret_cast<const half4_t*>(Wo + r16 * kNrcW + 16 * t + 4 * g);
        acc = mfma16(a, act[t], acc);
    }
    for (int j = 0; j < 4; j++) acc[j] += sB[4 * kNrcW + 4 * g + j];
    return acc;
}

__device__ __forceinline__ void encode(const NrcDev& N, const float* raw, half4_t act[4]) {
    const int g = (threadIdx.x & 63) >> 4;
    for (int t = 0; t < 4; t++)
        for (int j = 0; j < 4; j++) act[t][j] = (_Float16)nrc_feature(raw, N, 16 * t + 4 * g + j);
}

// Inference. list == nullptr: queries 0..n-1 with results in out[q] (float4).
// Otherwise query i is path list[i]: L[p] += beta[p] * max(0, prediction).
__global__ void __launch_bounds__(kNrcBlock) k_nrc_infer(NrcDev N, const int32_t* __restrict__ list,
                                                         const int32_t* __restrict__ countp, int32_t nfixed,
                                                         const float4* __restrict__ feats, float4* __restrict__ L,
                                                         const float4* __restrict__ beta, float4* __restrict__ out,
                                                         unsigned long long* tot) {
    __shared__ __align__(16) _Float16 sW[kNrcLayers * kNrcW * kNrcW];
    __shared__ float sB[kNrcLayers * kNrcW];
    stage_weights(N.wT, N.b, sW, sB);
    const int n = countp ? min(*countp, nfixed) : nfixed;
    if (tot && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(tot, (unsigned long long)n);
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 4;
    for (int base = (blockIdx.x * 4 + wv) * 16; base < n; base += gridDim.x * 4 * 16) {
        int q = base + (l & 15);
        bool valid = q < n;
        int p = valid ? (list ? list[q] : q) : 0;
        float raw[kNrcFeat];
        for (int k = 0; k < 4; k++) {
            float4 v = valid ? feats[(int64_t)p * 4 + k] : make_float4(0.f, 0.f, 0.f, 0.f);
            raw[4 * k] = v.x; raw[4 * k + 1] = v.y; raw[4 * k + 2] = v.z; raw[4 * k + 3] = v.w;
        }
        half4_t act[4];
        encode(N, raw, act);
        float4_t y = nrc_forward(sW, sB, act, q, valid, nullptr, 0);
        if (valid && g == 0) {
            float r = fmaxf(y[0], 0.f), gg = fmaxf(y[1], 0.f), bb = fmaxf(y[2], 0.f);
            if (list) {
                float4 b4 = beta[p], L4 = L[p];
                L[p] = make_float4(L4.x + b4.x * r, L4.y + b4.y * gg, L4.z + b4.z * bb, L4.w);
            } else {
                out[q] = make_float4(r, gg, bb, 0.f);
            }
        }
    }
}

// Training forward: stores the encoding and hidden activations, and the scaled
// relative-L2 loss gradient dY = 2 (y - t) / (y^2 + 0.01) / 3 (clamped) as fp16 in D5.
__global__ void __launch_bounds__(kNrcBlock) k_nrc_fwd_train(NrcDev N, const float4* __restrict__ feats,
                                                             const float4* __restrict__ targets,
                                                             const int32_t* __restrict__ countp, int32_t js,
                                                             int32_t ns, int32_t nfixed, _Float16* H, _Float16* D5,
                                                             int64_t rowStride, float* lossAcc) {
    __shared__ __align__(16) _Float16 sW[kNrcLayers * kNrcW * kNrcW];
    __shared__ float sB[kNrcLayers * kNrcW];
    stage_weights(N.wT, N.b, sW, sB);
    int off;
    const int n = slice_rows(countp, js, ns, nfixed, &off);
    feats += (int64_t)off * 4;
    targets += off;
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 4;
    float lsum = 0.f;
    for (int base = (blockIdx.x * 4 + wv) * 16; base < n; base += gridDim.x * 4 * 16) {
        int q = base + (l & 15);
        bool valid = q < n;
        float raw[kNrcFeat];
        for (int k = 0; k < 4; k++) {
            float4 v = valid ? feats[(int64_t)q * 4 + k] : make_float4(0.f, 0.f, 0.f, 0.f);
            raw[4 * k] = v.x; raw[4 * k + 1] = v.y; raw[4 * k + 2] = v.z; raw[4 * k + 3] = v.w;
        }
        half4_t act[4];
        encode(N, raw, act);
        float4_t y = nrc_forward(sW, sB, act, q, valid, H, rowStride);
        if (valid) {
            float4 tg = targets[q];
            float tc[3] = {tg.x, tg.y, tg.z};
            half4_t d = {0, 0, 0, 0};
            if (g == 0) {
                for (int c = 0; c < 3; c++) {
                    float e = y[c] - tc[c];
                    float den = y[c] * y[c] + 0.01f;
                    lsum += e * e / den;
                    d[c] = (_Float16)fminf(fmaxf(2.f * e / den / 3.f, -1000.f), 1000.f);
                }
            }
            *reinterpret_cast<half4_t*>(D5 + (int64_t)q * kNrcW + 4 * g) = d;
            for (int t = 1; t < 4; t++)  // columns 16..63 of the padded output gradient
                *reinterpret_cast<half4_t*>(D5 + (int64_t)q * kNrcW + 16 * t + 4 * g) = half4_t{0, 0, 0, 0};
        }
    }
    if (lossAcc) {
        for (int sh = 32; sh > 0; sh >>= 1) lsum += __shfl_xor(lsum, sh);
        if (l == 0) atomicAdd(lossAcc, lsum);
    }
}

// Backward through the layers: D_k = (W_k . D_{k+1}) * (H_k > 0) for k = 4..1, with
// D5 = dY (16 padded columns). D has 5 slabs (D1..D5), H 5 slabs (H0..H4).
__global__ void __launch_bounds__(kNrcBlock) k_nrc_bwd(NrcDev N, const int32_t* __restrict__ countp, int32_t js,
                                                       int32_t ns, int32_t nfixed, const _Float16* __restrict__ H,
                                                       _Float16* D, int64_t rowStride) {
    __shared__ __align__(16) _Float16 sW[kNrcLayers * kNrcW * kNrcW];
    __shared__ float sB[kNrcLayers * kNrcW];
    stage_weights(N.w, N.b, sW, sB);
    int off;
    const int n = slice_rows(countp, js, ns, nfixed, &off);
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 4, r16 = l & 15;
    for (int base = (blockIdx.x * 4 + wv) * 16; base < n; base += gridDim.x * 4 * 16) {
        int q = base + r16;
        bool valid = q < n;
        int qq = valid ? q : 0;
        // B fragments of D5 (one k-tile: 16 output columns)
        half4_t dn[4];
        dn[0] = *reinterpret_cast<const half4_t*>(D + 4 * rowStride + (int64_t)qq * kNrcW + 4 * g);
        int ktiles = 1;
        for (int layer = 4; layer >= 1; layer--) {
            const _Float16* Wl = sW + layer * kNrcW * kNrcW;  // [in][out]
            half4_t dp[4];
            for (int m = 0; m < 4; m++) {
                float4_t acc = {0.f, 0.f, 0.f, 0.f};
                for (int t = 0; t < ktiles; t++) {
                    half4_t a = *reinterpret_cast<const half4_t*>(Wl + (16 * m + r16) * kNrcW + 16 * t + 4 * g);
                    acc = mfma16(a, dn[t], acc);
                }
                half4_t h = *reinterpret_cast<const half4_t*>(H + (int64_t)layer * rowStride + (int64_t)qq * kNrcW +
                                                              16 * m + 4 * g);
                for (int j = 0; j < 4; j++) acc[j] = (float)h[j] > 0.f ? acc[j] : 0.f;
                dp[m] = to_half4(acc);
                if (valid)
                    *reinterpret_cast<half4_t*>(D + (int64_t)(layer - 1) * rowStride + (int64_t)q * kNrcW + 16 * m +
                                                4 * g) = dp[m];
            }
            for (int t = 0; t < 4; t++) dn[t] = dp[t];
            ktiles = 4;
        }
    }
}

// Split-K weight gradients: partial[layer][wave] = sum over the wavefront's rows of
// H_layer^T D_{layer+1} (64 x 64, output layer 64 x 16) and column sums of D_{layer+1}.
__global__ void __launch_bounds__(kNrcBlock) k_nrc_wgrad(const int32_t* __restrict__ countp, int32_t js,
                                                         int32_t ns, int32_t nfixed, const _Float16* __restrict__ H,
                                                         const _Float16* __restrict__ D, int64_t rowStride,
                                                         float* partial, int32_t maxWaves) {
    __shared__ __align__(16) _Float16 sH[4][16 * kNrcW];
    __shared__ __align__(16) _Float16 sD[4][16 * kNrcW];
    const int layer = blockIdx.y;
    int off;
    const int n = slice_rows(countp, js, ns, nfixed, &off);
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 4, r16 = l & 15;
    const int wave = blockIdx.x * 4 + wv;
    const int mo_tiles = layer == 4 ? 1 : 4;
    float4_t acc[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) acc[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    const _Float16* Hl = H + (int64_t)layer * rowStride;
    const _Float16* Dl = D + (int64_t)layer * rowStride;  // D slab `layer` holds D_{layer+1}
    int r0 = wave * kRowsPerWave;
    for (int blk = 0; blk < kRowsPerWave / 16; blk++) {
        int q0 = r0 + blk * 16;
        if (q0 >= n) break;
        // stage 16 rows of H and D (zero rows past n)
        for (int k = l; k < 16 * kNrcW / 4; k += 64) {
            int row = (k * 4) / kNrcW;
            bool ok = q0 + row < n;
            uint2 hv = ok ? *reinterpret_cast<const uint2*>(Hl + (int64_t)q0 * kNrcW + k * 4) : make_uint2(0, 0);
            uint2 dv = ok ? *reinterpret_cast<const uint2*>(Dl + (int64_t)q0 * kNrcW + k * 4) : make_uint2(0, 0);
            *reinterpret_cast<uint2*>(&sH[wv][k * 4]) = hv;
            *reinterpret_cast<uint2*>(&sD[wv][k * 4]) = dv;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        for (int r = 0; r < 16; r++) bsum += (float)sD[wv][r * kNrcW + l];
        for (int mi = 0; mi < 4; mi++) {
            half4_t a;
            for (int j = 0; j < 4; j++) a[j] = sH[wv][(4 * g + j) * kNrcW + 16 * mi + r16];
            for (int mo = 0; mo < mo_tiles; mo++) {
                half4_t b;
                for (int j = 0; j < 4; j++) b[j] = sD[wv][(4 * g + j) * kNrcW + 16 * mo + r16];
                acc[mi][mo] = mfma16(a, b, acc[mi][mo]);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (wave >= maxWaves) return;
    float* out = partial + ((int64_t)layer * maxWaves + wave) * kPartial;
    // acc[mi][mo]: lane holds dW[in = 16 mi + 4 g + j][out = 16 mo + r16]
    for (int mi = 0; mi < 4; mi++)
        for (int mo = 0; mo < 4; mo++)
            for (int j = 0; j < 4; j++)
                out[(16 * mi + 4 * g + j) * kNrcW + 16 * mo + r16] = mo < mo_tiles ? acc[mi][mo][j] : 0.f;
    out[kNrcW * kNrcW + l] = bsum;
}

// Sum the per-wavefront partials (mean over n rows) into grad [layer][in][out] + [layer][64].
__global__ void k_nrc_reduce(const int32_t* __restrict__ countp, int32_t js, int32_t ns, int32_t nfixed,
                             const float* __restrict__ partial, int32_t maxWaves, float* grad,
                             unsigned long long* tot) {
    int off;
    const int n = slice_rows(countp, js, ns, nfixed, &off);
    if (tot && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(tot, (unsigned long long)n);
    const int used = min(maxWaves, (n + kRowsPerWave - 1) / kRowsPerWave);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kNrcLayers * kPartial) return;
    int layer = i / kPartial, e = i % kPartial;
    float s = 0.f;
    for (int w = 0; w < used; w++) s += partial[((int64_t)layer * maxWaves + w) * kPartial + e];
    s = n > 0 ? s / (float)n : 0.f;
    if (e < kNrcW * kNrcW)
        grad[layer * kNrcW * kNrcW + e] = s;
    else
        grad[kNrcLayers * kNrcW * kNrcW + layer * kNrcW + (e - kNrcW * kNrcW)] = s;
}

// Adam on the FP32 master parameters; refresh the fp16 copies in both layouts.
__global__ void k_nrc_adam(const int32_t* __restrict__ countp, int32_t js, int32_t ns, int32_t nfixed,
                           float* master, float* m1, float* m2, const float* __restrict__ grad, _Float16* w,
                           _Float16* wT, float lr, float beta1, float beta2, float eps, float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kNrcParams) return;
    int off;
    if (slice_rows(countp, js, ns, nfixed, &off) == 0) return;  // an empty slice: no step
    bool isW = i < kNrcLayers * kNrcW * kNrcW;
    int layer = isW ? i / (kNrcW * kNrcW) : (i - kNrcLayers * kNrcW * kNrcW) / kNrcW;
    int out = isW ? i % kNrcW : (i - kNrcLayers * kNrcW * kNrcW) % kNrcW;
    bool active = layer < 4 || out < kNrcOut;  // output layer: 3 of 64 padded columns
    float p = master[i];
    if (active) {
        float gr = grad[i];
        float a = beta1 * m1[i] + (1.f - beta1) * gr;
        float b = beta2 * m2[i] + (1.f - beta2) * gr * gr;
        m1[i] = a;
        m2[i] = b;
        p -= lr * (a / c1) / (sqrtf(b / c2) + eps);
        master[i] = p;
    }
    if (isW) {
        int in = (i % (kNrcW * kNrcW)) / kNrcW;
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
    _Float16* H = nullptr;    // 5 slabs [cap][64]
    _Float16* D = nullptr;    // 5 slabs [cap][64]
    float* partial = nullptr;
    int32_t maxWaves = 0;
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

static int nrc_grid(int64_t rows) {
    int64_t blocks = (rows + 4 * 16 - 1) / (4 * 16);
    return (int)std::max<int64_t>(1, std::min<int64_t>(blocks, 2048));
}

// Launches used by the renderer (c2_device.hip) on its stream.
void nrc_launch_infer(c2_nrc* c, const int32_t* list, const int32_t* countp, int32_t maxn, const float4* feats,
                      float4* L, const float4* beta, hipStream_t s) {
    hipLaunchKernelGGL(k_nrc_infer, dim3(nrc_grid(maxn)), dim3(kNrcBlock), 0, s, c->dev, list, countp, maxn, feats,
                       L, beta, (float4*)nullptr, c->totals);
}

// One training step on slice js of ns of the *countp rows collected on the device (at most
// maxn rows per slice); a host-sized batch passes countp = nullptr, js = 0, ns = 1.
void nrc_launch_train(c2_nrc* c, const float4* feats, const float4* targets, const int32_t* countp, int32_t js,
                      int32_t ns, int32_t maxn, hipStream_t s, bool adam) {
    if (maxn > c->cap) fail(C2_ERR_INVALID, "radiance cache: training batch exceeds capacity");
    int64_t rs = (int64_t)c->cap * kNrcW;
    hipLaunchKernelGGL(k_nrc_fwd_train, dim3(nrc_grid(maxn)), dim3(kNrcBlock), 0, s, c->dev, feats, targets, countp,
                       js, ns, maxn, c->H, c->D + 4 * rs, rs, c->loss);
    hipLaunchKernelGGL(k_nrc_bwd, dim3(nrc_grid(maxn)), dim3(kNrcBlock), 0, s, c->dev, countp, js, ns, maxn, c->H,
                       c->D, rs);
    int waves = (maxn + kRowsPerWave - 1) / kRowsPerWave;
    hipLaunchKernelGGL(k_nrc_wgrad, dim3((waves + 3) / 4, kNrcLayers), dim3(kNrcBlock), 0, s, countp, js, ns, maxn,
                       c->H, c->D, rs, c->partial, c->maxWaves);
    int tot = kNrcLayers * kPartial;
    hipLaunchKernelGGL(k_nrc_reduce, dim3((tot + 255) / 256), dim3(256), 0, s, countp, js, ns, maxn, c->partial,
                       c->maxWaves, c->grad, c->totals + 1);
    if (adam) {
        c->step++;
        const c2_nrc_opts& o = c->opts;
        float c1 = 1.f - (float)std::pow((double)o.beta1, (double)c->step);
        float c2v = 1.f - (float)std::pow((double)o.beta2, (double)c->step);
        hipLaunchKernelGGL(k_nrc_adam, dim3((kNrcParams + 255) / 256), dim3(256), 0, s, countp, js, ns, maxn,
                           c->master, c->m1, c->m2, c->grad, c->w, c->wT, o.learning_rate, o.beta1, o.beta2,
                           o.epsilon, c1, c2v);
    }
    NRC_CHECK(hipGetLastError());
}

NrcDev nrc_dev(const c2_nrc* c) { return c->dev; }

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
        c->H = c->A.alloc<_Float16>((size_t)kNrcLayers * c->cap * kNrcW);
        c->D = c->A.alloc<_Float16>((size_t)kNrcLayers * c->cap * kNrcW);
        c->maxWaves = (c->cap + kRowsPerWave - 1) / kRowsPerWave;
        c->partial = c->A.alloc<float>((size_t)kNrcLayers * c->maxWaves * kPartial);
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
        NRC_CHECK(hipMemcpy(df, feats, (size_t)n * kNrcFeat * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_nrc_infer, dim3(nrc_grid(n)), dim3(kNrcBlock), 0, 0, c->dev, (const int32_t*)nullptr,
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
        NRC_CHECK(hipMemcpy(df, feats, (size_t)n * kNrcFeat * sizeof(float), hipMemcpyHostToDevice));
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
