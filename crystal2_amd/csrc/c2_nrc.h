// Radiance cache network shared definitions (device + host).
//
// MLP 64 -> 64 -> 64 -> 64 -> 64 -> 3 (4 hidden ReLU layers, linear output),
// FP16 operands on v_mfma_f32_16x16x32_f16 with FP32 accumulation, FP32 master
// weights and Adam. A query is 16 raw floats: position, outgoing direction,
// facing normal, roughness, albedo; the 64-dim input encoding is built inside
// the kernels (per coordinate 8 frequencies x {sin, cos}, spherical direction and
// normal, 4-bin one-blob roughness, albedo, bias).
//
// Feature-major formulation per wavefront: H_out^T (64 x 16 queries) =
// W^T (64 x 64) * H_in^T (64 x 16). With the MFMA fragment layouts (lane l = 16 g + c)
//   A: lane l holds A[c][8 g + j], j < 8
//   B: lane l holds B[8 g + j][c]
//   D: lane l holds D[4 g + i][c], i < 4
// and the K order of a 32-feature k-tile chosen so that slot j of lane group g is
// feature 32 t + 16 (j / 4) + 4 g + j % 4, the D fragments of output tiles 2t and
// 2t + 1 of a layer, concatenated per lane, are exactly the next layer's B fragment
// of k-tile t: layers chain in registers without transposes (the weights are staged
// with their columns in that order).
#pragma once
#include <cstdint>

namespace c2 {

constexpr int kNrcIn = 64;
constexpr int kNrcW = 64;       // hidden width
constexpr int kNrcLayers = 5;   // 4 hidden + output
constexpr int kNrcOut = 3;
constexpr int kNrcFeat = 16;    // raw floats per query (host API)
// Packed query on the device, kNrcQ4 float4 per query, written by the renderer and by the
// host conversion of raw queries:
//   q0 = (position xyz, roughness)
//   q1 = (azimuth of wo, cos theta of wo * 0.5 + 0.5, azimuth of n, cos theta of n * 0.5 + 0.5)
//   q2 = (albedo rgb, 1)
// (azimuth = atan2(z, x) / 2 pi + 0.5; cos theta = the y component). These are the raw
// encoding inputs every lane group of the MLP kernels needs: q0 for the position bands and
// the roughness one-blob, q1 and q2 for features 48..51 and 56..59, so a lane loads two.
constexpr int kNrcQ4 = 3;
// The renderer's query record adds a fourth float4: (throughput rgb, path index bits); the
// inference kernel adds throughput x prediction to that path's radiance.
constexpr int kNrcRec4 = kNrcQ4 + 1;
constexpr int kNrcParams = kNrcLayers * kNrcW * kNrcW + kNrcLayers * kNrcW;  // padded storage

// Forward FLOP per query: 4 x (64 x 64 x 2) + 64 x 3 x 2
constexpr double kNrcFlopFwd = 4.0 * 64 * 64 * 2 + 64.0 * 3 * 2;

struct NrcDev {
    const _Float16* wT;  // [layer][out][in]  forward A fragments
    const _Float16* w;   // [layer][in][out]  backward A fragments
    const float* b;      // [layer][64]
    float bmin[3], binv[3];  // position normalisation to [0, 1]
};

}  // namespace c2
