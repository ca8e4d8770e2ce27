// The warm-start network for rotamer belief propagation: a multilayer
// perceptron 32 -> 128 -> 128 -> 16 (ReLU hidden layers) with bfloat16
// weights and activations and float32 accumulation on the gfx950 matrix
// cores (v_mfma_f32_16x16x32_bf16).  It maps per-residue features of the
// first BP pass to corrections of that pass's beliefs.  No counterpart in the
// reference, whose solve always starts from the node probabilities
// (src/rotamer.cpp:1110-1133).
//
// Matrix-core layout (16x16x32, wave64): lane l holds A[l%16][8(l/16)+j] and
// B[8(l/16)+j][l%16] (j = 0..7, bf16) and gets D[4(l/16)+i][l%16] (i = 0..3,
// f32).  Every layer is computed transposed, H^T = W X^T (M = output
// features, N = 16 rows of a tile, K = input features), so a lane's D values
// of output blocks 2kb and 2kb+1 are directly its B operand for input block
// kb of the next layer: within a K block the inputs are taken in the order
//   slot (g, j) -> feature (2 kb + j / 4) * 16 + 4 g + j % 4,   g = l / 16,
// and the weights of layers 2 and 3 are stored with the same permutation so
// that each lane's A operand is 16 contiguous bytes.  A sum over K does not
// depend on the order of its terms' positions, only on the pairing, which
// this keeps.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace upside {

constexpr int MLP_IN = 32;      // features per residue
constexpr int MLP_H = 128;      // hidden width (both layers)
constexpr int MLP_OUT = 16;     // outputs (the first 6 are used: one per rotamer state)
constexpr int MLP_USED = 6;
// float32 parameter vector (the trainer's master copy and the C ABI layout):
// W1 [128][32], b1 [128], W2 [128][128], b2 [128], W3 [16][128], b3 [16]
// (weights as [output][input]).
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_OUT * MLP_H;
constexpr int MLP_NPARAM = MLP_OFF_B3 + MLP_OUT;
// Workgroups of the gradient kernel (each writes one partial gradient; the
// partials are summed in workgroup order).
constexpr int MLP_PARTS = 32;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Device copies of the network in the layouts the kernels read.
struct MlpDev {
    const __bf16* w1;    // [128][32]   W1, natural input order
    const __bf16* w2p;   // [128][128]  W2, inputs permuted per K block (see above)
    const __bf16* w3p;   // [16][128]   W3, permuted likewise
    const __bf16* w3tp;  // [128][32]   W3^T for the backward pass: row j, slot (g, jj) = W3[4g+jj][j] (jj < 4), else 0
    const __bf16* w2tp;  // [128][128]  W2^T for the backward pass, permuted: row i, slot (kb, g, jj) = W2[f][i]
    const float* b1;     // [128]
    const float* b2;     // [128]
    const float* b3;     // [16]
};

// Input feature of K block kb taken by slot (g, j).
__host__ __device__ __forceinline__ int mlp_perm(int kb, int g, int j) { return (2 * kb + j / 4) * 16 + 4 * g + (j & 3); }

__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// 8 bf16 from 16 aligned bytes.
__device__ __forceinline__ bf16x8 ld_bf16x8(const __bf16* p) {
    return *reinterpret_cast<const bf16x8*>(p);
}

// B operand of the next layer from two output blocks held in D layout.
__device__ __forceinline__ bf16x8 pack_blocks(f32x4 lo, f32x4 hi) {
    bf16x8 r;
    r[0] = (__bf16)lo[0]; r[1] = (__bf16)lo[1]; r[2] = (__bf16)lo[2]; r[3] = (__bf16)lo[3];
    r[4] = (__bf16)hi[0]; r[5] = (__bf16)hi[1]; r[6] = (__bf16)hi[2]; r[7] = (__bf16)hi[3];
    return r;
}

// Forward pass of one 16-row tile by one wave (all 64 lanes).  X: float
// features [n][MLP_IN] (row-major), rows r0 .. r0+15 (rows >= n read as 0).
// Hidden activations are kept in D layout (h1[ob], h2[ob]: outputs
// ob*16 + 4g + i of row r0 + l%16) for the backward pass; the output y holds
// outputs 4g + i of row r0 + l%16 (bias added).
template <class PX>
__device__ __forceinline__ void mlp_tile_forward(const MlpDev& w, PX X, int n, int r0, f32x4* h1, f32x4* h2,
                                                 f32x4& y) {
    const int l = (int)(threadIdx.x & 63), g = l >> 4, c = l & 15;
    const int row = r0 + c;
    bf16x8 x;
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = (__bf16)(row < n ? X[row * MLP_IN + 8 * g + j] : 0.f);
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ob = 0; ob < MLP_H / 16; ++ob) {
        f32x4 a = mfma_bf16(ld_bf16x8(w.w1 + (ob * 16 + c) * MLP_IN + 8 * g), x, z);
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = fmaxf(a[i] + w.b1[ob * 16 + 4 * g + i], 0.f);
        h1[ob] = a;
    }
    bf16x8 b[MLP_H / 32];
#pragma unroll
    for (int kb = 0; kb < MLP_H / 32; ++kb) b[kb] = pack_blocks(h1[2 * kb], h1[2 * kb + 1]);
#pragma unroll
    for (int ob = 0; ob < MLP_H / 16; ++ob) {
        f32x4 a = z;
#pragma unroll
        for (int kb = 0; kb < MLP_H / 32; ++kb) a = mfma_bf16(ld_bf16x8(w.w2p + (ob * 16 + c) * MLP_H + kb * 32 + 8 * g), b[kb], a);
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = fmaxf(a[i] + w.b2[ob * 16 + 4 * g + i], 0.f);
        h2[ob] = a;
    }
#pragma unroll
    for (int kb = 0; kb < MLP_H / 32; ++kb) b[kb] = pack_blocks(h2[2 * kb], h2[2 * kb + 1]);
    f32x4 a = z;
#pragma unroll
    for (int kb = 0; kb < MLP_H / 32; ++kb) a = mfma_bf16(ld_bf16x8(w.w3p + c * MLP_H + kb * 32 + 8 * g), b[kb], a);
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] += w.b3[4 * g + i];
    y = a;
}

// Forward pass of one 16-row tile without keeping the hidden layers (the
// warm start inside the engine kernel): y as in mlp_tile_forward.
template <class PX>
__device__ __forceinline__ void mlp_tile_infer(const MlpDev& w, PX X, int n, int r0, f32x4& y) {
    const int l = (int)(threadIdx.x & 63), g = l >> 4, c = l & 15;
    const int row = r0 + c;
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
    bf16x8 b[MLP_H / 32];
    {
        bf16x8 x;
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = (__bf16)(row < n ? X[row * MLP_IN + 8 * g + j] : 0.f);
#pragma unroll
        for (int kb = 0; kb < MLP_H / 32; ++kb) {
            f32x4 h[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int ob = 2 * kb + q;
                f32x4 a = mfma_bf16(ld_bf16x8(w.w1 + (ob * 16 + c) * MLP_IN + 8 * g), x, z);
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = fmaxf(a[i] + w.b1[ob * 16 + 4 * g + i], 0.f);
                h[q] = a;
            }
            b[kb] = pack_blocks(h[0], h[1]);
        }
    }
    bf16x8 b2[MLP_H / 32];
#pragma unroll
    for (int kb2 = 0; kb2 < MLP_H / 32; ++kb2) {
        f32x4 h[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int ob = 2 * kb2 + q;
            f32x4 a = z;
#pragma unroll
            for (int kb = 0; kb < MLP_H / 32; ++kb) a = mfma_bf16(ld_bf16x8(w.w2p + (ob * 16 + c) * MLP_H + kb * 32 + 8 * g), b[kb], a);
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = fmaxf(a[i] + w.b2[ob * 16 + 4 * g + i], 0.f);
            h[q] = a;
        }
        b2[kb2] = pack_blocks(h[0], h[1]);
    }
    f32x4 a = z;
#pragma unroll
    for (int kb = 0; kb < MLP_H / 32; ++kb) a = mfma_bf16(ld_bf16x8(w.w3p + c * MLP_H + kb * 32 + 8 * g), b2[kb], a);
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] += w.b3[4 * g + i];
    y = a;
}

}  // namespace upside
